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


def fit(renderer, target: np.ndarray, iterations=200, start=START, lr=.01, refresh=25, floor=1e-3):
    """Adam on the five parameters, from start toward the image target, on the adjoint's gradients; the
    loss and parameters at each iteration.

    Spin and inclination pull the image alike (the loss along spin alone is shallow, and a fit with a step
    size per parameter stalls in the valley between them), so Adam runs in whitened coordinates: every
    `refresh` iterations the Gauss-Newton matrix H = V diag(lam) V^T is taken at the current point
    (eigenvalues held above `floor` x the largest), and p = p_ref + V diag(lam^-1/2) z, z optimised from 0
    by a fresh Adam whose learning rate follows a cosine from lr down to lr / 20 over the whole fit."""
    t = torch.from_numpy(np.asarray(target, np.float64))
    p = np.array(start, np.float64)
    history = []
    k = 0
    while k < iterations:
        lam, V = np.linalg.eigh(gauss_newton(renderer, p))
        lam = np.maximum(lam, lam.max() * floor)
        to_params = torch.from_numpy(V @ np.diag(lam ** -.5))
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
            k += 1
        p = (p_ref + to_params @ z).detach().numpy()
    return history
