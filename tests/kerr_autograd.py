"""A test harness: the differentiable analytic-disc renderer (include/artful/adjoint.h, through ctypes)
as a torch.autograd.Function of (spin, inclination in degrees, disc inner, disc outer, q).

The scene is the adjoint workload's: 512x512 by default, one ray a pixel, eye at r = 50 looking at the
hole, thin analytic disc (emissivity (inner/r)^q), Doppler 1, no stars, no bloom; 2000 steps, escape
at r > 100. torch is used here as the optimiser's plumbing only: the gradients come from the HIP
adjoint kernel."""
from __future__ import annotations

import math

import numpy as np
import torch

from artful_amd import native

RADIUS, FOV = 50.0, .9
PARAMS = ("spin", "inclination", "inner", "outer", "q")
HIDDEN = (.7, 35.0, 3.4, 18.0, 2.5)
START = (.3, 60.0, 5.0, 25.0, 3.0)


def scene(params, samples=1):
    spin, incl, inner, outer, q = (float(v) for v in params)
    desc = native.hole_desc(spin=spin, axis=(0, 1, 0), discInner=inner, discOuter=outer, doppler=1.0, stars=0.0, bloom=0.0,
                            samples=samples)
    i = math.radians(incl)
    up = (0, 1, 0) if math.pi / 2 - i <= 1 else (0, 0, -1)
    cam = native.camera((0.0, RADIUS * math.cos(i), RADIUS * math.sin(i)), (0, 0, 0), up, FOV, .1)
    return desc, cam, q


def tracing(q, precision="fp32"):
    return native.tracing(precision=precision, disc="analytic", emissivity=q, max_steps=2000, escape_radius=100.0)


def render(ctx: native.HoleAdjoint, params, precision="fp32") -> np.ndarray:
    desc, cam, q = scene(params)
    return ctx.render(desc, cam, tracing(q, precision))


def gradient(ctx: native.HoleAdjoint, params, seed: np.ndarray, precision="fp32"):
    """dL/d(spin, inclination [per degree], inner, outer, q) for dL/dHDR = seed; and the image."""
    desc, cam, q = scene(params)
    g, hdr = ctx.gradient(desc, cam, tracing(q, precision), seed)
    i = math.radians(float(params[1]))
    # eye = r (0, cos i, sin i): d eye / d i = r (0, -sin i, cos i), per degree
    d_incl = (g.position[1] * -RADIUS * math.sin(i) + g.position[2] * RADIUS * math.cos(i)) * math.pi / 180
    return np.array([g.spin, d_incl, g.discInner, g.discOuter, g.emissivity]), hdr


class KerrImage(torch.autograd.Function):
    """image = KerrImage.apply(params, ctx): params a float64 tensor of 5; image (3, H, W) float32."""

    @staticmethod
    def forward(ctx_, params, renderer):
        p = params.detach().cpu().numpy()
        ctx_.renderer, ctx_.p = renderer, p
        return torch.from_numpy(render(renderer, p))

    @staticmethod
    def backward(ctx_, grad_output):
        seed = grad_output.detach().cpu().numpy().astype(np.float32)
        g, _ = gradient(ctx_.renderer, ctx_.p, seed)
        return torch.from_numpy(g), None


def mse(image, target):
    return ((image - target) ** 2).mean()


# central-difference steps of the forward renders the preconditioner is built from (spin, degrees, masses, masses, q)
FD_STEPS = (2e-3, .1, 5e-3, .02, 2e-3)


def gauss_newton(renderer, params, steps=FD_STEPS) -> np.ndarray:
    """The loss's Gauss-Newton matrix (2/N) J^T J at params, J = d(image)/d(params) by central differences
    of forward renders (ten of them): the metric the fit's preconditioner whitens."""
    cols = []
    for i, h in enumerate(steps):
        e = np.zeros(5)
        e[i] = h
        cols.append(((render(renderer, params + e).astype(np.float64) - render(renderer, params - e)) / (2 * h)).ravel())
    J = np.stack(cols, 1)
    return 2.0 / J.shape[0] * (J.T @ J)


# where the fit's parameters are held (spin, inclination in degrees, inner, outer, q); outer stays at
# least OUTER_GAP beyond inner. Past a spin of .998 the scene clamps and the image stops changing, so an
# iterate that left these would find no gradient back.
BOUNDS = np.array([[0.0, .99], [1.0, 89.0], [2.0, 10.0], [8.0, 60.0], [.5, 6.0]])
OUTER_GAP = 2.0


def within_bounds(params: np.ndarray) -> np.ndarray:
    p = np.clip(params, BOUNDS[:, 0], BOUNDS[:, 1])
    p[3] = max(p[3], p[2] + OUTER_GAP)
    return p


def fit(renderer, target: np.ndarray, iterations=200, start=START, lr=.01, refresh=25, floor=1e-3):
    """Adam on the five parameters, from start toward the image target, on the adjoint's gradients; the
    loss and parameters at each iteration.

    Spin and inclination pull the image alike (the loss along spin alone is shallow, and a fit with a step
    size per parameter stalls in the valley between them), so Adam runs in whitened coordinates: every
    `refresh` iterations the Gauss-Newton matrix H = V diag(lam) V^T is taken at the current point
    (eigenvalues held above `floor` x the largest), and p = p_ref + V diag(lam^-1/2) z, z optimised from 0
    by a fresh Adam whose learning rate follows a cosine from lr down to lr / 20 over the whole fit. An
    iterate outside BOUNDS is moved back onto them (z set to the bounded point's coordinates).

    Along the valley the loss is rippled (the integrator's discrete steps move against the disc as the
    parameters change): its local slope, the adjoint's and a finite difference's alike, is ten times the
    valley's overall slope, so where along the valley the fit passes closest to the target and where it
    stops depend on the start and on the gradients' last bits. Its result is the iterate of least loss
    (`best`), which lies nearer the target than the last one: each refresh's fresh Adam first moves
    every coordinate by the learning rate, and near the target that step costs about what is left."""
    t = torch.from_numpy(np.asarray(target, np.float64))
    p = np.array(start, np.float64)
    history = []
    k = 0
    while k < iterations:
        lam, V = np.linalg.eigh(gauss_newton(renderer, p))
        lam = np.maximum(lam, lam.max() * floor)
        to_params = torch.from_numpy(V @ np.diag(lam ** -.5))
        from_params = torch.from_numpy(np.diag(lam ** .5) @ V.T)
        p_ref = torch.from_numpy(p.copy())
        z = torch.zeros(5, dtype=torch.float64, requires_grad=True)
        opt = torch.optim.Adam([z], lr=lr)
        for _ in range(min(refresh, iterations - k)):
            for group in opt.param_groups:
                group["lr"] = lr * (.05 + .95 * .5 * (1 + math.cos(math.pi * k / iterations)))
            opt.zero_grad()
            params = p_ref + to_params @ z
            loss = mse(KerrImage.apply(params, renderer).double(), t)
            loss.backward()
            history.append((float(loss.detach()), params.detach().numpy().copy()))
            opt.step()
            with torch.no_grad():
                moved = (p_ref + to_params @ z).numpy()
                held = within_bounds(moved)
                if np.any(held != moved):
                    z.copy_(from_params @ torch.from_numpy(held - p_ref.numpy()))
            k += 1
        p = (p_ref + to_params @ z).detach().numpy()
    return history


def best(history):
    """The fit's result: the (loss, parameters) of its iterate of least loss."""
    return min(history, key=lambda h: h[0])
