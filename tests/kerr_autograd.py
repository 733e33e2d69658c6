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


# a step size per parameter (spin, degrees, masses, masses, q): Adam moves each by about this a step
STEP_SIZES = (.01, .5, .03, .2, .01)


def fit(renderer, target: np.ndarray, iterations=200, start=START, steps=STEP_SIZES):
    """Adam on the five parameters, from start toward the image target; the loss and parameters at each
    iteration. The parameters are p = start + steps * z, z optimised with unit learning rate."""
    t = torch.from_numpy(np.asarray(target, np.float64))
    s = torch.tensor(steps, dtype=torch.float64)
    p0 = torch.tensor(start, dtype=torch.float64)
    z = torch.zeros(5, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([z], lr=1.0)
    history = []
    for _ in range(iterations):
        opt.zero_grad()
        params = p0 + s * z
        loss = mse(KerrImage.apply(params, renderer).double(), t)
        loss.backward()
        history.append((float(loss), params.detach().numpy().copy()))
        opt.step()
    return history
