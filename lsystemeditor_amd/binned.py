"""Material colours and light emission as 16-bin spectra for the differentiable renderer.

A bin value is a reflectance (or, for a light, a relative emission) that is constant over one of
16 equal wavelength intervals of 360..830 nm. The renderer takes linear-sRGB colours, so a binned
spectrum enters it through its sRGB projection

    M[c, b] = (sum over the 1-nm samples of bin b of response_c) / (sum over 360..830 nm of response_c)

with response = CIE 1931 colour matching x D65 -> linear sRGB, the response the hero-wavelength
mode integrates (rgb2spec.rgb_response); a constant spectrum 1 projects to (1, 1, 1). Gradients
with respect to the bins are M^T times the renderer's gradients with respect to the colour
(torch autograd through the projection), so the adjoint kernel serves them unchanged. Light
transport itself stays RGB: a product of two binned spectra is rendered as the product of their
projections.

The spectra start from the materials' RGB through the rgb2spec model (a sigmoid of a quadratic in
wavelength fitted to the colour, the model of the reference's spectral mode), averaged per bin,
so they project back to about the original colours.
"""
from __future__ import annotations

import numpy as np
import torch

from . import diff, rgb2spec

BINS = 16


def projection(bins: int = BINS) -> np.ndarray:
    """(3, bins) float64: linear sRGB of a spectrum that is 1 on bin b and 0 elsewhere."""
    lam = np.arange(rgb2spec.LAMBDA_MIN, rgb2spec.LAMBDA_MAX + 1.0, 1.0)
    resp = rgb2spec.rgb_response(lam)
    norm = resp.sum(0)
    b = np.minimum(((lam - rgb2spec.LAMBDA_MIN) / rgb2spec.LAMBDA_RANGE * bins).astype(np.int64), bins - 1)
    m = np.zeros((3, bins))
    for k in range(bins):
        m[:, k] = resp[b == k].sum(0) / norm
    return m


def projection_over(lo: float, hi: float, bins: int = BINS) -> np.ndarray:
    """(3, bins) float64: linear sRGB of a spectrum that is 1 on bin b of [lo, hi] nm and 0
    elsewhere, normalised so that a constant 1 over [lo, hi] projects to (1, 1, 1)."""
    lam = np.arange(lo, hi + 1.0, 1.0)
    resp = rgb2spec.rgb_response(lam)
    b = np.minimum(((lam - lo) / (hi - lo) * bins).astype(np.int64), bins - 1)
    m = np.zeros((3, bins))
    for k in range(bins):
        m[:, k] = resp[b == k].sum(0) / resp.sum(0)
    return m


def uplift(rgb, bins: int = BINS, lo: float = rgb2spec.LAMBDA_MIN, hi: float = rgb2spec.LAMBDA_MAX) -> np.ndarray:
    """(bins,) float64: the rgb2spec spectrum of a linear-sRGB colour in [0, 1]^3, averaged over
    each of `bins` equal intervals of [lo, hi] nm (1-nm samples)."""
    fit = rgb2spec._Fit()
    target = np.clip(np.asarray(rgb, np.float64), 0.0, 1.0)[None, :]
    c, _ = fit.solve(np.zeros((1, 3)), target, 200)
    lam = np.arange(lo, hi + 1.0, 1.0)
    t = (lam - rgb2spec.LAMBDA_MIN) / rgb2spec.LAMBDA_RANGE
    x = c[0, 0] * t * t + c[0, 1] * t + c[0, 2]
    s = 0.5 + 0.5 * x / np.sqrt(1.0 + x * x)
    b = np.minimum(((lam - lo) / (hi - lo) * bins).astype(np.int64), bins - 1)
    return np.array([s[b == k].mean() for k in range(bins)])


def hero_spectra(materials: np.ndarray, lights: list[int], lo: float = 380.0, hi: float = 720.0,
                 bins: int = BINS) -> np.ndarray:
    """(materials, 2 bins) float32 table for the hero-wavelength renderer (ptSceneSetSpectra,
    bins over the hero range [lo, hi]): every material's reflectance is the uplift of its base
    colour; the lights' emission shapes are the uplift of their colours (their emission scales
    them, as it does r, g, b); the other materials keep the rgb2spec emission (-1)."""
    m = np.asarray(materials, np.float32).reshape(-1, 24)
    out = np.full((m.shape[0], 2 * bins), -1.0, np.float32)
    for i in range(m.shape[0]):
        out[i, :bins] = uplift(m[i, 0:3], bins, lo, hi)
        if i in lights:
            out[i, bins:] = out[i, :bins]
    return out


class BinnedParams:
    """Per material group: 16 reflectance bins + roughness, metallic, specular (surfaces), or 16
    emission bins (lights, whose emission scale stays fixed); theta() gives the renderer's
    (materials, 7) parameters (diff.PARAM_NAMES) as a differentiable function of them."""

    def __init__(self, theta0: torch.Tensor, lights: list[int], bins: int = BINS):
        self.theta0 = theta0.detach().to(torch.float32).clone()
        self.bins = bins
        self.lights = list(lights)
        self.surfaces = [i for i in range(theta0.shape[0]) if i not in self.lights]
        self.m = torch.from_numpy(projection(bins)).to(torch.float32)
        t0 = self.theta0.numpy()
        self.refl = torch.tensor(np.array([uplift(t0[i, 0:3], bins) for i in self.surfaces]), dtype=torch.float32,
                                 requires_grad=True)
        self.scal = self.theta0[self.surfaces][:, 3:6].clone().requires_grad_(True)
        self.emis = torch.tensor(np.array([uplift(t0[i, 0:3], bins) for i in self.lights]), dtype=torch.float32,
                                 requires_grad=True)

    def parameters(self) -> list[torch.Tensor]:
        return [self.refl, self.scal, self.emis]

    @property
    def count(self) -> int:
        return sum(p.numel() for p in self.parameters())

    def theta(self, refl=None, scal=None, emis=None) -> torch.Tensor:
        refl = self.refl if refl is None else refl
        scal = self.scal if scal is None else scal
        emis = self.emis if emis is None else emis
        rows = [None] * self.theta0.shape[0]
        srgb = refl @ self.m.T    # (surfaces, 3)
        lrgb = emis @ self.m.T    # (lights, 3)
        for k, i in enumerate(self.surfaces):
            rows[i] = torch.cat([srgb[k], scal[k], self.theta0[i, 6:7]])
        for k, i in enumerate(self.lights):
            rows[i] = torch.cat([lrgb[k], self.theta0[i, 3:7]])
        return torch.stack(rows)

    def perturbed(self, seed: int = 11, spread: float = 0.2) -> list[torch.Tensor]:
        """Every parameter scaled by an independent seeded factor in [1 - spread, 1 + spread]."""
        g = np.random.default_rng(seed)
        return [(p.detach() * torch.from_numpy(g.uniform(1.0 - spread, 1.0 + spread, tuple(p.shape))).float())
                for p in self.parameters()]


def gradient_check(r: diff.DiffRenderer, bp: BinnedParams, target: torch.Tensor, h: float = 1e-2) -> dict:
    """Gradient of L = sum (image - target)^2 with respect to every binned parameter: the adjoint
    chained through the projection, against central differences (step h) of the detached
    estimator with the parameters of the forward's samples."""
    theta = bp.theta()
    img = r.forward(theta.detach())
    g_theta = r.adjoint((2.0 * (img - target)).contiguous()).to(torch.float32).cpu()
    grads = torch.autograd.grad(theta, bp.parameters(), grad_outputs=g_theta)
    adj = torch.cat([g.reshape(-1) for g in grads]).double().numpy()
    base = [p.detach().clone() for p in bp.parameters()]
    fd = np.zeros_like(adj)
    k = 0
    for which, p in enumerate(base):
        flat = p.reshape(-1)
        for j in range(flat.numel()):
            losses = []
            for sgn in (1.0, -1.0):
                q = [b.clone() for b in base]
                q[which].view(-1)[j] += sgn * h
                im = r.detached(bp.theta(*q).detach())
                losses.append(float(((im.double() - target.double()) ** 2).sum().item()))
            fd[k] = (losses[0] - losses[1]) / (2.0 * h)
            k += 1
    floor = 1e-3 * float(np.abs(fd).max())
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), floor)
    return {"adjoint": adj, "fd": fd, "rel_err": rel, "max_rel_err": float(rel.max()),
            "median_rel_err": float(np.median(rel)), "params": int(adj.size)}
