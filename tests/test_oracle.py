"""Pins on the oracle beyond the golden vectors: what general relativity and the closed forms fix.

- the shadow of a Schwarzschild hole, sin(alpha) = sqrt(27) sqrt(1 - 2/r) / r (tests.c:4249-4275);
- the two edges of a Kerr shadow from the prograde and retrograde photon orbits (tests.c:4277-4315);
- the Doppler asymmetry of the disc (tests.c:4342-4362);
- the total deflection of equatorial Schwarzschild rays against the elliptic integral
  phi = int du / sqrt(1/b^2 - u^2 + 2 u^3), within 1e-4 rad (the integrator's own error is ~1e-5;
  a term dropped or a sign wrong in u'' moves it by 0.1 rad or more);
- the FP64 oracle against the FP32 one (pinned bit for bit by the golden vectors);
- the frame worked out in doubles against the reference's float frame;
- the bloom pyramid against an independent numpy pyramid, the tone map against its closed form;
- the analytic disc seen face-on around a = 0: light gain * w(r)^2 * (inner/r)^q * colour(4500 K) at
  the crossing radius the orbit integral gives, with and without the Keplerian shift
  g = sqrt(1 - 3/r) / sqrt(1 - 2/r_eye); the ratio of two emissivity indices; the escape radius
  (the direction a ray leaves in at u = 1/R against the orbit integral); the jitter rule (strata, and
  the oracle tracing each pixel at exactly the rule's places).
These run the FP32 and FP64 oracle: a plausible mistake in either fails one of them. Planted in a copy
of the oracle, an inverted emissivity exponent, an edge twice as wide, an unsquared window, an escape
at 2R and swapped jitter axes each fail the tests below."""
from __future__ import annotations

import math
import warnings

import numpy as np
import pytest
from scipy import integrate, optimize

from artful_amd import abi
from tests.golden import cases

W, H = 320, 200


def scene(native, distance, tilt, fov=.9, **kw):
    case = dict(distance=distance, tilt=tilt, fov=fov, desc=kw)
    desc, cam = cases.make(case)
    return native.Scene(desc, cam)


def render(oracle, s, w, h, precision=32):
    f, n = s.frame(w, h)
    f64 = s.frame64(w, h) if precision == 64 else None
    hdr = oracle.trace(f, s.map(), s.fluid()[0], precision, f64)
    bl = oracle.bloom(f, hdr, n)
    return oracle.tone_map(f, hdr, bl)


def shadow_radius(img):
    least = 1e9
    for i in range(16):
        dx, dy = math.cos(i * .3926991), math.sin(i * .3926991)
        at = 1.0
        while at < H * .5:
            x, y = int(W * .5 + dx * at), int(H * .5 + dy * at)
            if ((int(img[y, x]) >> 16) & 255) >= 24:
                least = min(least, at)
                break
            at += .25
    return least


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("r", [14.0, 30.0, 80.0])
def test_schwarzschild_shadow(native, oracle, r, precision):
    sine = math.sqrt(27) * math.sqrt(1 - 2 / r) / r
    angle = math.asin(sine)
    fov = angle * 2.6
    s = scene(native, r, 1.57079633, fov=fov, discInner=10.0, discOuter=60.0, discGain=200.0, samples=4, stars=0.0)
    img = render(oracle, s, W, H, precision)
    expected = math.tan(angle) / math.tan(fov * .5) * H * .5
    assert abs(shadow_radius(img) - expected) < 1.5


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("a", [.3, .6, .95])
def test_kerr_shadow_edges(native, oracle, a, precision):
    r = 60.0
    s = scene(native, r, 0.0, fov=.36, spin=a, discGain=0.0, discThickness=0.0, stars=1e4)
    img = render(oracle, s, W, H, precision)
    row = img[H // 2]
    left = right = W // 2
    while left > 0 and (int(row[left]) & 255) < 40:
        left -= 1
    while right < W - 1 and (int(row[right]) & 255) < 40:
        right += 1
    big = (r * r + a * a) ** 2 - a * a * (r * r - 2 * r + a * a)
    varpi, alpha, turning = math.sqrt(big) / r, r * math.sqrt(r * r - 2 * r + a * a) / math.sqrt(big), 2 * a * r / big
    edge = []
    for k in range(2):
        orbit = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(a if k else -a)))
        xi = -(orbit ** 3 - 3 * orbit ** 2 + a * a * orbit + a * a) / (a * (orbit - 1.0))
        along = xi * alpha / (varpi * (1.0 - turning * xi))
        edge.append(along / math.sqrt(1.0 - along * along) / math.tan(.18) * H * .5)
    assert abs((right - W // 2) - edge[0]) < 2.0
    assert abs((left + 1 - W // 2) - edge[1]) < 2.0


def brightness(img, x0, x1):
    p = img[:, x0:x1].astype(np.int64)
    return float(((p >> 16) & 255).sum() + ((p >> 8) & 255).sum() + (p & 255).sum()) / (H * (x1 - x0))


def test_doppler(native, oracle):
    img = render(oracle, scene(native, 26.0, .06, doppler=0.0, stars=0.0), W, H)
    left, right = brightness(img, 0, W // 2), brightness(img, W // 2, W)
    assert right > 0 and abs(left - right) / (left + right) < .2
    img = render(oracle, scene(native, 26.0, .06, doppler=1.0, stars=0.0, exposure=.1), W, H)
    assert brightness(img, W // 2, W) > 1.5 * brightness(img, 0, W // 2)


def swept_angle(b, u0):
    """Total angle an incoming equatorial ray from u0 sweeps to infinity, around a = 0."""
    f = lambda u: 1 / b ** 2 - u * u + 2 * u ** 3
    up = optimize.brentq(f, 1e-12, 1 / 3)  # periapsis: the smallest positive root (b > sqrt 27)
    # u = up - t^2 takes out the square-root singularity at the periapsis
    g = lambda t: 2 * t / math.sqrt(max(f(up - t * t), 1e-300))
    inward = integrate.quad(g, 0, math.sqrt(up - u0), limit=200, epsabs=1e-12)[0]
    outward = integrate.quad(g, 0, math.sqrt(up), limit=200, epsabs=1e-12)[0]
    return inward + outward


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("b", [6.0, 8.0, 12.0, 25.0])
def test_schwarzschild_deflection(native, oracle, b, precision):
    r0 = 50.0
    s = scene(native, r0, 0.0, fov=1.2, spin=0.0, discGain=0.0, discThickness=0.0)
    f, _ = s.frame(64, 64)
    f64 = s.frame64(64, 64)
    sin_alpha = b * math.sqrt(1 - 2 / r0) / r0
    tan_x = math.sqrt(sum(v * v for v in f64.right[:3]))
    nx = math.tan(math.asin(sin_alpha)) / tan_x
    m, fl = s.map(), s.fluid()[0]
    lw = oracle.law(max_steps=2000)
    _, steps, path = oracle.ray(f, m, fl, nx, 0.0, 0.0, precision, f64 if precision == 64 else None, lw)
    assert path[4] == 1.0, "the ray must escape"
    away = path[:3] / np.linalg.norm(path[:3])
    start = np.array(f64.er[:3])
    got = math.acos(float(np.clip(away @ start, -1, 1)))
    want = swept_angle(b, 1 / r0)
    # the angle from acos is in [0, pi]: unwrap the exact one onto the same branch
    want_wrapped = abs((want + math.pi) % (2 * math.pi) - math.pi)
    assert abs(got - want_wrapped) < 1e-4, (got, want_wrapped, steps)


def test_fp64_oracle_near_fp32(native, oracle, scene_of):
    s = scene_of("config1")
    f, _ = s.frame(320, 180)
    h32 = oracle.trace(f, s.map(), s.fluid()[0], 32)
    h64 = oracle.trace(f, s.map(), s.fluid()[0], 64, s.frame64(320, 180))
    rel = np.abs(h32.astype(np.float64) - h64) / np.maximum(np.abs(h64), 1e-3)
    assert rel.mean() < 1e-3 and (rel > 1e-3).mean() < .02


def test_frame64_matches_float_frame(scene_of):
    for name in ("config1", "moving_eye", "defaults", "inner_given"):
        s = scene_of(name)
        f, _ = s.frame(101, 67)
        f64 = s.frame64(101, 67)
        for field in ("er", "etheta", "ephi", "right", "up", "forward", "metric", "place", "spin", "boost", "disc"):
            a = np.array(getattr(f, field)[:], np.float64)
            b = np.array(getattr(f64, field)[:], np.float64)
            np.testing.assert_allclose(b, a, rtol=2e-6, atol=2e-6, err_msg=f"{name}.{field}")


def numpy_bloom(frame, hdr):
    """An independent pyramid: thresholded 2x2 means, then 4 bilinear taps down and up."""
    W_, H_ = frame.size[0], frame.size[1]
    levels = [tuple(frame.bloom[k][:3]) for k in range(frame.size[3])]

    def tap(img, sx, sy):
        h_, w_ = img.shape
        bx, by = np.floor(sx), np.floor(sy)
        px, py = (sx - bx).astype(np.float32), (sy - by).astype(np.float32)
        x0 = np.clip(bx.astype(int), 0, w_ - 1)
        x1 = np.clip(bx.astype(int) + 1, 0, w_ - 1)
        y0 = np.clip(by.astype(int), 0, h_ - 1)
        y1 = np.clip(by.astype(int) + 1, 0, h_ - 1)
        top = img[y0, x0] + (img[y0, x1] - img[y0, x0]) * px
        bottom = img[y1, x0] + (img[y1, x1] - img[y1, x0]) * px
        return top + (bottom - top) * py

    out = []
    for c in range(3):
        pyr = []
        for k, (_, lw, lh) in enumerate(levels):
            ys, xs = np.mgrid[0:lh, 0:lw]
            if k == 0:
                br = np.minimum(np.maximum(hdr[c] - 2.0, 0.0), 48.0).astype(np.float32)
                y0, y1 = np.minimum(2 * ys, H_ - 1), np.minimum(2 * ys + 1, H_ - 1)
                x0, x1 = np.minimum(2 * xs, W_ - 1), np.minimum(2 * xs + 1, W_ - 1)
                pyr.append(((br[y0, x0] + br[y0, x1] + br[y1, x0] + br[y1, x1]) * .25).astype(np.float32))
            else:
                cx, cy = (xs * 2 + .5).astype(np.float32), (ys * 2 + .5).astype(np.float32)
                sm = sum(tap(pyr[-1], cx + dx, cy + dy) for dy in (-1, 1) for dx in (-1, 1))
                pyr.append((sm * .25).astype(np.float32))
        for k in range(len(levels) - 2, -1, -1):
            _, lw, lh = levels[k]
            ys, xs = np.mgrid[0:lh, 0:lw]
            sx, sy = ((xs + .5) * .5 - .5).astype(np.float32), ((ys + .5) * .5 - .5).astype(np.float32)
            sm = sum(tap(pyr[k + 1], sx + dx, sy + dy) for dy in (-.5, .5) for dx in (-.5, .5))
            pyr[k] = (pyr[k] + sm * np.float32(.25 * .85)).astype(np.float32)
        out.append(np.concatenate([p.ravel() for p in pyr]))
    return np.stack(out)


@pytest.mark.parametrize("size", [(37, 23), (64, 36), (5, 4)])
def test_bloom_against_numpy(native, oracle, size):
    s = scene(native, 26.0, .06)
    f, n = s.frame(*size)
    rng = np.random.default_rng(7)
    hdr = (rng.random((3, size[1], size[0]), np.float32) * 60).astype(np.float32)
    got = oracle.bloom(f, hdr, n)
    np.testing.assert_allclose(got, numpy_bloom(f, hdr), rtol=2e-5, atol=1e-5)


def hash3(x, y, z):
    x, y, z = (np.asarray(v, np.uint64) for v in (x, y, z))
    m = np.uint64(0xFFFFFFFF)
    h = ((x * np.uint64(73856093)) ^ (y * np.uint64(19349663)) ^ (z * np.uint64(83492791))) & m
    h = ((h ^ (h >> np.uint64(15))) * np.uint64(0x2C1B3C6D)) & m
    h = ((h ^ (h >> np.uint64(12))) * np.uint64(0x297A2D39)) & m
    return h ^ (h >> np.uint64(15))


def test_tone_map_closed_form(native, oracle):
    """Narkowicz's ACES fit, a square root, x 255, plus a hashed dither, truncated; bloom off."""
    s = scene(native, 26.0, .06, exposure=1.7)
    w, h = 50, 30
    f, n = s.frame(w, h)
    rng = np.random.default_rng(3)
    hdr = (rng.random((3, h, w), np.float32) * 3).astype(np.float32)
    got = oracle.tone_map(f, hdr, np.zeros((3, n), np.float32))
    x = hdr.astype(np.float64) * 1.7
    v = np.sqrt(np.minimum(x * (2.51 * x + .03) / (x * (2.43 * x + .59) + .14), 1.0))
    ys, xs = np.mgrid[0:h, 0:w]
    dither = (hash3(xs + ys * 8192, 7, 13) & np.uint64(255)).astype(np.float64) / 255
    byte = np.minimum(v * 255 + dither, 255).astype(np.int64)
    want = (0xFF << 24) | (byte[0] << 16) | (byte[1] << 8) | byte[2]
    diff = np.abs(((got[None].astype(np.int64) >> np.array([16, 8, 0])[:, None, None]) & 255) - byte)
    assert diff.max() <= 1 and (got == want).mean() > .99


# ---- the analytic disc, the escape radius and the jitter, pinned by closed forms ----
#
# The analytic disc is this project's extension (the reference's disc is painted): a thin disc whose
# light at radius r is gain * (inner / r)^q, inside soft edges a mass wide, w(r) = s(r - inner + 1/2) *
# (1 - s(r - outer + 1/2)) with s the smoothstep on [0, 1], of opacity min(thickness, 1) * w. A ray's
# first crossing therefore brings gain * w^2 * (inner / r)^q * colour(4500 K) (w for the light, w for
# its opacity). Seen from the axis of a Schwarzschild hole each ray stays in a plane through the axis
# and reaches the disc where its orbit has swept pi / 2, so the crossing radius follows from the orbit
# integral phi = int du / sqrt(1/b^2 - u^2 + 2 u^3); and the light's frequency shift there is the
# textbook one for a face-on Keplerian disc, g = sqrt(1 - 3 / r) / sqrt(1 - 2 / r_eye) (the ray has no
# angular momentum about the axis). A flipped exponent, an edge a mass off or twice as wide, an
# unsquared window or a g^3 fails these by orders of magnitude more than the integrator's own error
# (~3e-6 relative on the flat part of the disc, ~2e-4 on the steep edges, measured).

FACE_R0, FACE_INNER, FACE_OUTER, FACE_GAIN = 50.0, 8.0, 16.0, 2.0


def orbit_integrand(b):
    f = lambda u: 1 / b ** 2 - u * u + 2 * u ** 3
    up = optimize.brentq(f, 1e-12, 1 / 3)  # periapsis (b > sqrt 27)
    return up, (lambda t: 2 * t / math.sqrt(max(f(up - t * t), 1e-300)))  # u = up - t^2


def swept_to(b, u0, angle):
    """1/u where an incoming ray from u0 has swept `angle`: on its way in when the bending brings that
    before the periapsis, else on its way out."""
    up, g = orbit_integrand(b)

    def swept(u):  # from u to the periapsis (quad warns when the root finder's bracket probes u next to it)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", integrate.IntegrationWarning)
            return integrate.quad(g, 0, math.sqrt(up - u), limit=200, epsabs=1e-13)[0]
    inward = swept(u0)
    if inward > angle:
        return 1 / optimize.brentq(lambda u: inward - swept(u) - angle, u0, up)
    return 1 / optimize.brentq(lambda u: swept(u) - (angle - inward), 1e-12, up)


def smooth(t):
    t = min(max(t, 0.0), 1.0)
    return t * t * (3 - 2 * t)


def face_window(r):
    return smooth(r - FACE_INNER + .5) * (1 - smooth(r - FACE_OUTER + .5))


def colour_at(colors, heat):
    """The black-body table (pinned bit for bit by the golden map digests) read at heat * 128, linearly."""
    at = min(max(heat * 128, 0.0), 255 - 1e-3)
    c0 = int(math.floor(at))
    return colors[c0] + (colors[c0 + 1] - colors[c0]) * (at - c0)


@pytest.fixture(scope="module")
def face_on(native):
    out = {}
    for doppler in (0.0, 1.0):
        case = dict(distance=FACE_R0, tilt=1.57079633, fov=.9,
                    desc=dict(spin=0.0, discInner=FACE_INNER, discOuter=FACE_OUTER, discGain=FACE_GAIN, doppler=doppler,
                              stars=0.0, discThickness=1.0))
        s = native.Scene(*cases.make(case))
        f, _ = s.frame(64, 64)
        out[doppler] = (s, f, s.frame64(64, 64))
    return out


def nx_for(f64, b, r0=FACE_R0):
    """The image coordinate of impact parameter b: a static eye at r0 sees it at sin(alpha) = b sqrt(1 - 2/r0) / r0."""
    tan_x = math.sqrt(sum(v * v for v in f64.right[:3]))
    return math.tan(math.asin(b * math.sqrt(1 - 2 / r0) / r0)) / tan_x


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("doppler", [0.0, 1.0])
def test_analytic_disc_face_on(face_on, oracle, precision, doppler):
    s, f, f64 = face_on[doppler]
    m, fl = s.map(), s.fluid()[0]
    assert list(f.look)[:2] == [FACE_GAIN, doppler] and list(f64.disc)[:2] == [FACE_INNER, FACE_OUTER]
    colors = m[abi.COLORS_AT:abi.COLORS_AT + 256 * 3].reshape(256, 3).astype(np.float64)
    seen = {"dark": 0, "edge": 0, "flat": 0}
    for q in (3.0, 2.5):
        for b in np.linspace(6.5, 17.75, 46):
            rc = swept_to(b, 1 / FACE_R0, math.pi / 2)
            g = math.sqrt(1 - 3 / rc) / math.sqrt(1 - 2 / FACE_R0) if doppler else 1.0
            w = face_window(rc)
            want = FACE_GAIN * w * w * (FACE_INNER / rc) ** q * g ** 4 * colour_at(colors, 4500 / 6500 * g)
            rgb, _, _ = oracle.ray(f, m, fl, nx_for(f64, b), 0.0, 0.0, precision, f64 if precision == 64 else None,
                                   oracle.law(disc="analytic", emissivity=q))
            if w == 0.0 and abs(rc - FACE_INNER) > .52 and abs(rc - FACE_OUTER) > .52:
                assert not rgb.any(), (b, rc, rgb)
                seen["dark"] += 1
            elif w == 1.0 and min(abs(rc - FACE_INNER), abs(rc - FACE_OUTER)) > .52:
                np.testing.assert_allclose(rgb, want, rtol=2e-5, err_msg=f"b {b} r {rc}")
                seen["flat"] += 1
            elif w > .02:
                np.testing.assert_allclose(rgb, want, rtol=1e-3, err_msg=f"b {b} r {rc} w {w}")
                seen["edge"] += 1
    assert min(seen.values()) >= 4, seen  # inside the inner edge, both edges and the flat part all met


@pytest.mark.parametrize("precision", [32, 64])
def test_analytic_disc_emissivity_ratio(face_on, oracle, precision):
    """q changes the light, not the path: HDR(q1) / HDR(q2) = (inner / r)^(q1 - q2), and the crossing
    radius that ratio implies is the orbit integral's."""
    s, f, f64 = face_on[0.0]
    m, fl = s.map(), s.fluid()[0]
    for b in (10.0, 12.5, 15.0):
        rc = swept_to(b, 1 / FACE_R0, math.pi / 2)
        got = [oracle.ray(f, m, fl, nx_for(f64, b), 0.0, 0.0, precision, f64 if precision == 64 else None,
                          oracle.law(disc="analytic", emissivity=q))[0].astype(np.float64) for q in (3.5, 1.0)]
        ratio = got[0] / got[1]
        np.testing.assert_allclose(ratio, (FACE_INNER / rc) ** 2.5, rtol=2e-5)
        implied = FACE_INNER * ratio.mean() ** (-1 / 2.5)
        assert abs(implied - rc) < 2e-5 * rc, (implied, rc)


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("radius", [60.0, 100.0, 1000.0])
def test_escape_radius_law(native, oracle, radius, precision):
    """With an escape radius R a ray leaves where u = 1/R on its way out: the direction it leaves in has
    swept the orbit integral from u0 through the periapsis to 1/R (measured error <= 5.4e-4 rad; leaving
    at 2R instead moves it by 0.1 rad or more, and not interpolating within the step by up to 0.13)."""
    r0 = 50.0
    s = scene(native, r0, 0.0, fov=1.2, spin=0.0, discGain=0.0, discThickness=0.0)
    f, _ = s.frame(64, 64)
    f64 = s.frame64(64, 64)
    m, fl = s.map(), s.fluid()[0]
    steps_at = {}
    for b in (6.0, 8.0, 12.0, 25.0):
        up, g = orbit_integrand(b)
        want = integrate.quad(g, 0, math.sqrt(up - 1 / r0), limit=200, epsabs=1e-12)[0] + \
            integrate.quad(g, 0, math.sqrt(up - 1 / radius), limit=200, epsabs=1e-12)[0]
        want = abs((want + math.pi) % (2 * math.pi) - math.pi)
        _, steps, path = oracle.ray(f, m, fl, nx_for(f64, b, r0), 0.0, 0.0, precision, f64 if precision == 64 else None,
                                    oracle.law(max_steps=2000, escape_radius=radius))
        assert path[4] == 1.0, "the ray must leave"
        away = path[:3] / np.linalg.norm(path[:3])
        got = math.acos(float(np.clip(away @ np.array(f64.er[:3]), -1, 1)))
        assert abs(got - want) < 1e-3, (b, got, want, steps)
        steps_at[b] = steps
        _, steps_far, _ = oracle.ray(f, m, fl, nx_for(f64, b, r0), 0.0, 0.0, precision, f64 if precision == 64 else None,
                                     oracle.law(max_steps=2000))
        assert steps <= steps_far


def jitter_places(seed, width, height, side):
    """The project's jitter rule: ray (sx, sy) of pixel (px, py) at ((sx + jx) / side, (sy + jy) / side) with
    h = hash(px ^ (seed * 0x9E3779B1), py, sy * side + sx + seed), jx = (h & 0xFFFF + 1/2) / 2^16,
    jy = (h >> 16 + 1/2) / 2^16 (the reference's own hash, lens.inc:56-61). Arrays (side^2, H, W)."""
    ys, xs = np.mgrid[0:height, 0:width]
    dx = np.zeros((side * side, height, width), np.float32)
    dy = np.zeros_like(dx)
    for sy in range(side):
        for sx in range(side):
            h = hash3(xs.astype(np.uint64) ^ np.uint64((seed * 0x9E3779B1) & 0xFFFFFFFF), ys, sy * side + sx + seed)
            jx = ((h & np.uint64(0xFFFF)).astype(np.float32) + np.float32(.5)) * np.float32(1 / 65536)
            jy = ((h >> np.uint64(16)).astype(np.float32) + np.float32(.5)) * np.float32(1 / 65536)
            dx[sy * side + sx] = (np.float32(sx) + jx) / np.float32(side)
            dy[sy * side + sx] = (np.float32(sy) + jy) / np.float32(side)
    return dx, dy


@pytest.mark.parametrize("seed", [2, 9])
def test_jitter_places(native, oracle, seed):
    """Every jittered ray lies inside its stratum, the places spread over it, and the oracle traces each
    pixel at exactly the places the rule gives: FP32 with one ray a pixel bit for bit (the pixel is its
    ray), FP64 with 2x2 rays as the mean of the four rays traced one by one."""
    w, h = 24, 16
    s = scene(native, 26.0, .3, fov=.9, spin=.7, doppler=1.0, stars=2.0)
    m, fl = s.map(), s.fluid()[0]
    for side in (1, 2, 3):
        dx, dy = jitter_places(seed, w, h, side)
        for k in range(side * side):
            sx, sy = k % side, k // side
            assert (dx[k] > sx / side).all() and (dx[k] < (sx + 1) / side).all()
            assert (dy[k] > sy / side).all() and (dy[k] < (sy + 1) / side).all()
        u = (dx * side) % 1.0
        assert abs(float(u.mean()) - .5) < .08 and u.min() < .1 and u.max() > .9  # not a fixed offset
    lw = oracle.law(jitter=seed)
    # FP32, one ray a pixel: the reference's way of forming x (lensTile's group of 8, lens.inc:367)
    s1 = scene(native, 26.0, .3, fov=.9, spin=.7, doppler=1.0, stars=2.0, samples=1)
    f, _ = s1.frame(w, h)
    hdr = oracle.trace(f, m, fl, 32, None, lw)
    dx, dy = jitter_places(seed, w, h, 1)
    for py in range(h):
        for px in range(w):
            nx = (np.float32(px & 7) + (np.float32(px & ~7) + dx[0, py, px])) * np.float32(f.image[1]) - np.float32(1)
            ny = np.float32(1) - (np.float32(py) + dy[0, py, px]) * np.float32(f.image[2])
            rgb, _, _ = oracle.ray(f, m, fl, float(nx), float(ny), float(f.disc[2]), 32, None, lw)
            assert np.array_equal(rgb.view(np.uint32), hdr[:, py, px].view(np.uint32)), (px, py)
    # FP64, 2x2 rays a pixel
    s2 = scene(native, 26.0, .3, fov=.9, spin=.7, doppler=1.0, stars=2.0, samples=2)
    f, _ = s2.frame(w, h)
    f64 = s2.frame64(w, h)
    hdr = oracle.trace(f, m, fl, 64, f64, lw)
    dx, dy = jitter_places(seed, w, h, 2)
    for py in range(0, h, 3):
        for px in range(0, w, 3):
            acc = np.zeros(3)
            for k in range(4):
                time = f.disc[2] + f.spin[3] * ((float((k * 5) % 4) + .5) * f.image[3] - .5)
                nx = (px + float(dx[k, py, px])) * (2.0 / w) - 1.0
                ny = 1.0 - (py + float(dy[k, py, px])) * (2.0 / h)
                acc += oracle.ray(f, m, fl, nx, ny, time, 64, f64, lw)[0]
            np.testing.assert_allclose(hdr[:, py, px], acc / 4, rtol=1e-6, atol=1e-9, err_msg=f"{px},{py}")
