"""Batched inverse rendering (workload C4): recover material parameters of many L-system scenes from
target images with Adam through the adjoint kernel.

Each scene is a stochastic bush of the C1 generator (seed s, iterations drawn from 3..6, 256x256,
16 spp, max depth 8) with the C2 materials. The optimised parameters of its bark, leaf and ground
groups are, by default, a 16-bin base-colour spectrum, roughness and metallic (binned.py: the
bins enter the renderer through their linear-sRGB projection); with binned_colour=False, the RGB
base colour, roughness and metallic. Specular and the lights stay fixed. The target parameters are
a seeded perturbation (every optimised value scaled by an independent factor in [0.8, 1.2]); the
target image is their render converged over 16 fixed-spp renders (256 spp, seeds disjoint from
every step's). The optimisation starts from the unperturbed parameters and minimises the summed
L2 image loss over all scenes with Adam (lr 1e-2).

Every step renders the loss image with a fresh sample set and replays another fresh, independent
set in the adjoint, so each step's gradient is an unbiased estimate of the gradient of the
expected-image loss |E[image](theta) - target|^2. With the same samples for the residual and the
derivatives the gradient would carry the covariance of image noise and its derivative, and with
fixed sample sets the sampling error would be a fixed function of the parameters: at 16 spp
either drives poorly determined parameters away from the target. A 16-spp target would be such a
fixed error too (its noise moves the loss minimum), hence the converged target. All scenes render
in one batch wavefront and differentiate in one batch adjoint.

Convergence is measured in image space as well as in parameter space: the loss of 256-spp renders
(one seed set, disjoint from the target's and the steps') at the start, at the end and at the
target parameters, whose loss is the noise floor of two 256-spp images. Roughness and metallic
are nearly degenerate with the base colours in these images (the recovered colours with the true
roughness and metallic render further from the target than the recovered parameters do; the
loss of that combination is reported too), so their parameter errors stay well above the image
loss's. RGB light transport sees a binned spectrum only through its projection (3 of its 16
degrees of freedom), so the bins' error is reported beside the error of their projected colour.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import abi, binned, diff, scenes

OPT_PARAMS = (0, 1, 2, 3, 4)          # base r, g, b, roughness, metallic (diff.PARAM_NAMES order)
OPT_GROUPS = (0, 1, 2)                # bark, leaf, ground
LIGHT_GROUPS = (3, 4)                 # the two lights stay fixed
TARGET_SEED_BASE = 1 << 24            # target renders; steps use seed bases below (2 steps + 2) spp
EVAL_SEED_BASE = 1 << 25              # image-loss evaluation renders
CONVERGED_PASSES = 16                 # fixed-spp renders averaged into a target or evaluation image


def c4_specs(count: int = 32, res: int = 256, spp: int = 16, depth: int = 8, iterations=(3, 6),
             first_seed: int = 0) -> list[scenes.SceneSpec]:
    """Scene s: iterations drawn with SplitMix64 from the seed; the C1 bush with that seed and the C2
    materials (every optimised parameter strictly inside its clamp range)."""
    out = []
    lo, hi = iterations
    for s in range(first_seed, first_seed + count):
        g = scenes.SplitMix64(0xC4000000 + s)
        it = lo + g.next() % (hi - lo + 1)
        c1 = scenes.config1(width=res, height=res, iterations=int(it), seed=s, min_spp=spp, max_spp=spp)
        spec = scenes.config2(width=res, height=res, spp=spp, c1=c1)
        spec.params[4:5] = np.array([depth], np.int32).view(np.float32)
        spec.name = "C4"
        spec.info = dict(spec.info, iterations=int(it), seed=s)
        out.append(spec)
    return out


def target_theta(theta0: np.ndarray, seed: int, spread: float = 0.2) -> np.ndarray:
    """The unperturbed parameters with the optimised RGB ones replaced by their seeded perturbation."""
    pert = scenes.perturbed_target(theta0, seed=seed, spread=spread)
    out = theta0.copy()
    for gidx in OPT_GROUPS:
        out[gidx, list(OPT_PARAMS)] = pert[gidx, list(OPT_PARAMS)]
    return out


class _RgbParams:
    """One scene's (materials, 7) parameter table, optimised on OPT_GROUPS x OPT_PARAMS."""

    def __init__(self, theta0: np.ndarray, seed: int):
        self.leaf = torch.from_numpy(theta0.copy()).requires_grad_(True)
        self.true = torch.from_numpy(target_theta(theta0, seed))
        mask = torch.zeros(theta0.shape)
        for gidx in OPT_GROUPS:
            mask[gidx, list(OPT_PARAMS)] = 1.0
        self.mask = mask

    def leaves(self) -> list[torch.Tensor]:
        return [self.leaf]

    def theta(self) -> torch.Tensor:
        return self.leaf

    def theta_true(self) -> torch.Tensor:
        return self.true

    def theta_true_scalars(self) -> torch.Tensor:
        """The current base colours with the target roughness and metallic."""
        out = self.leaf.detach().clone()
        out[:, 3:5] = self.true[:, 3:5]
        return out

    def mask_grads_(self) -> None:
        self.leaf.grad.mul_(self.mask)

    def clamp_(self) -> None:
        with torch.no_grad():
            self.leaf[:, 0:3].clamp_(0.0, 1.0)
            self.leaf[:, 3:5].clamp_(0.01, 0.99)

    def errors(self) -> dict[str, torch.Tensor]:
        d = (self.leaf.detach() - self.true).abs()[list(OPT_GROUPS)]
        return {"base_color": d[:, 0:3].reshape(-1), "roughness": d[:, 3], "metallic": d[:, 4]}


class _BinnedParams:
    """One scene's 16-bin base-colour spectra, roughness and metallic (binned.BinnedParams; specular
    and the lights' emission bins stay fixed)."""

    def __init__(self, theta0: np.ndarray, seed: int):
        self.bp = binned.BinnedParams(torch.from_numpy(theta0), lights=list(LIGHT_GROUPS))
        assert self.bp.surfaces == list(OPT_GROUPS)
        q = self.bp.perturbed(seed=0xC4B00000 + seed)
        q[1][:, 2] = self.bp.scal.detach()[:, 2]    # specular fixed
        q[2] = self.bp.emis.detach().clone()        # emission fixed
        self.true = q
        self.m = torch.from_numpy(binned.projection()).float()

    def leaves(self) -> list[torch.Tensor]:
        return [self.bp.refl, self.bp.scal]

    def theta(self) -> torch.Tensor:
        return self.bp.theta()

    def theta_true(self) -> torch.Tensor:
        return self.bp.theta(*self.true).detach()

    def theta_true_scalars(self) -> torch.Tensor:
        """The current base-colour spectra with the target roughness and metallic."""
        scal = self.bp.scal.detach().clone()
        scal[:, 0:2] = self.true[1][:, 0:2]
        return self.bp.theta(self.bp.refl.detach(), scal, self.bp.emis.detach()).detach()

    def mask_grads_(self) -> None:
        self.bp.scal.grad[:, 2] = 0.0

    def clamp_(self) -> None:
        with torch.no_grad():
            self.bp.refl.clamp_(0.0, 1.0)
            self.bp.scal[:, 0:2].clamp_(0.01, 0.99)

    def errors(self) -> dict[str, torch.Tensor]:
        dr = self.bp.refl.detach() - self.true[0]
        ds = (self.bp.scal.detach() - self.true[1]).abs()
        return {"base_color": (dr @ self.m.T).abs().reshape(-1), "bins": dr.abs().reshape(-1),
                "roughness": ds[:, 0], "metallic": ds[:, 1]}


HERO_RANGE = (380.0, 720.0)           # hero-wavelength range of the spectral variant (bins over it)


class _HeroParams:
    """One scene's (materials, 36) spectral parameter rows (diff.SpectralDiffRenderer: 16
    reflectance bins, 16 emission bins, roughness, metallic, specular, emission), optimised on the
    reflectance bins, roughness and metallic of OPT_GROUPS; the target scales each of those by an
    independent seeded factor in [0.8, 1.2]."""

    ROUGH, METAL = abi.SPECTRUM_WORDS, abi.SPECTRUM_WORDS + 1

    def __init__(self, theta0: np.ndarray, seed: int):
        self.leaf = torch.from_numpy(theta0.copy()).requires_grad_(True)
        mask = torch.zeros(theta0.shape)
        for gidx in OPT_GROUPS:
            mask[gidx, :abi.SPECTRUM_BINS] = 1.0
            mask[gidx, [self.ROUGH, self.METAL]] = 1.0
        self.mask = mask
        g = np.random.default_rng(0xC4B00000 + seed)
        factor = torch.from_numpy(g.uniform(0.8, 1.2, theta0.shape)).float()
        self.true = torch.where(mask > 0, self.leaf.detach() * factor, self.leaf.detach())
        self.m = torch.from_numpy(binned.projection_over(*HERO_RANGE)).float()

    def leaves(self) -> list[torch.Tensor]:
        return [self.leaf]

    def theta(self) -> torch.Tensor:
        return self.leaf

    def theta_true(self) -> torch.Tensor:
        return self.true

    def theta_true_scalars(self) -> torch.Tensor:
        """The current spectra with the target roughness and metallic."""
        out = self.leaf.detach().clone()
        out[:, [self.ROUGH, self.METAL]] = self.true[:, [self.ROUGH, self.METAL]]
        return out

    def mask_grads_(self) -> None:
        self.leaf.grad.mul_(self.mask)

    def clamp_(self) -> None:
        with torch.no_grad():
            self.leaf[list(OPT_GROUPS), :abi.SPECTRUM_BINS] = self.leaf[list(OPT_GROUPS), :abi.SPECTRUM_BINS].clamp(0, 1)
            self.leaf[:, [self.ROUGH, self.METAL]] = self.leaf[:, [self.ROUGH, self.METAL]].clamp(0.01, 0.99)

    def errors(self) -> dict[str, torch.Tensor]:
        d = (self.leaf.detach() - self.true)[list(OPT_GROUPS)]
        dr = d[:, :abi.SPECTRUM_BINS]
        return {"base_color": (dr @ self.m.T).abs().reshape(-1), "bins": dr.abs().reshape(-1),
                "roughness": d[:, self.ROUGH].abs(), "metallic": d[:, self.METAL].abs()}


class InverseProblem:
    """Scenes, renderers, target images and the parameters being optimised. transport="hero"
    renders and differentiates at four hero wavelengths (HERO_RANGE) with the base colours as
    16-bin reflectance spectra over that range (the lights' emission as the uplift of their
    colours, fixed); the hero-wavelength state (abi.Spectral) must be enabled with that range
    while the problem is used."""

    def __init__(self, specs: list[scenes.SceneSpec], bvh: str = "sah", binned_colour: bool = True,
                 transport: str = "rgb", fd_checker=None):
        if transport not in ("rgb", "hero"):
            raise ValueError("transport is 'rgb' or 'hero'")
        self.specs = specs
        self.fd_checker = fd_checker
        self.hero = transport == "hero"
        self.binned = binned_colour or self.hero
        self.renderers = []
        self.params = []
        for i, spec in enumerate(specs):
            sc = abi.Scene(spec.triangles, spec.materials, bvh=bvh)
            seed = int(spec.info.get("seed", i))
            if self.hero:
                sp = binned.hero_spectra(spec.materials, list(LIGHT_GROUPS), *HERO_RANGE)
                r = diff.SpectralDiffRenderer(sc, spec.materials, sp, spec.camera, spec.params, spec.width,
                                              spec.height, spec.spp)
                self.params.append(_HeroParams(r.theta0().numpy(), seed))
            else:
                r = diff.DiffRenderer(sc, spec.materials, spec.camera, spec.params, spec.width, spec.height, spec.spp)
                th0 = r.theta0().numpy()
                self.params.append(_BinnedParams(th0, seed) if binned_colour else _RgbParams(th0, seed))
            self.renderers.append(r)
        self.samples_per_render = [s.width * s.height * s.spp for s in specs]
        # every scene's loss render and adjoint in one batch wavefront
        self.batch = diff.BatchDiffRenderer(self.renderers)
        self.target_spp = CONVERGED_PASSES * self.batch.spp
        self.target = self.converged_image([p.theta_true() for p in self.params], TARGET_SEED_BASE)
        self.targets = [self.target[self.batch.pixels(i)] for i in range(len(specs))]

    def converged_image(self, thetas, seed_base: int, passes: int = CONVERGED_PASSES) -> torch.Tensor:
        """(total pixels, 3) mean of `passes` batch renders at thetas, render j with seed base
        seed_base + j spp (passes x spp samples per pixel)."""
        acc = None
        for j in range(passes):
            im = self.batch.forward([t.detach() for t in thetas], seed_base=seed_base + j * self.batch.spp)
            acc = im if acc is None else acc.add_(im)
        return acc / passes

    def image_loss(self, thetas) -> float:
        """Summed L2 loss of the thetas' converged image against the target (one evaluation seed set
        for every call, so differences between calls carry little sampling noise)."""
        return float(((self.converged_image(thetas, EVAL_SEED_BASE) - self.target) ** 2).sum())

    def image_losses(self) -> dict:
        """image_loss at the current parameters, at the target parameters (the noise floor), and
        at the current base colours with the target roughness and metallic."""
        return {"current": self.image_loss([p.theta() for p in self.params]),
                "target_params": self.image_loss([p.theta_true() for p in self.params]),
                "target_roughness_metallic": self.image_loss([p.theta_true_scalars() for p in self.params])}

    def leaves(self) -> list[torch.Tensor]:
        return [t for p in self.params for t in p.leaves()]

    def parameter_count(self) -> int:
        """Optimised values over all scenes."""
        per = self.params[0].errors()
        return len(self.params) * sum(per[k].numel() for k in per if k != ("base_color" if self.binned else ""))

    def loss(self, step: int = 0) -> torch.Tensor:
        """Summed L2 image loss of step `step`: the loss render uses seed base (2 step + 1) spp,
        the adjoint replays seed base (2 step + 2) spp (the target is seed base 0)."""
        spp = self.batch.spp
        thetas = [p.theta() for p in self.params]
        img = diff.render_batch_decorrelated(thetas, self.batch, (2 * step + 1) * spp, (2 * step + 2) * spp)
        return ((img - self.target) ** 2).sum().double()

    def mask_grads_(self) -> None:
        for p in self.params:
            p.mask_grads_()

    def clamp_(self) -> None:
        for p in self.params:
            p.clamp_()

    def param_errors(self) -> dict:
        """Mean absolute error against the targets: every optimised value ("all"), and per kind
        (base_color: RGB, or for binned spectra their projected colour; bins; roughness; metallic)."""
        per = [p.errors() for p in self.params]
        out = {k: float(torch.cat([e[k] for e in per]).mean()) for k in per[0]}
        raw = [torch.cat([e["bins" if self.binned else "base_color"], e["roughness"], e["metallic"]]) for e in per]
        out["all"] = float(torch.cat(raw).mean())
        return out

    def param_error(self) -> float:
        return self.param_errors()["all"]

    def fd_check(self, fd_step: float) -> dict:
        """Central-difference spot check of scene 0's gradient (detached estimator). At hero
        wavelengths the caller's fd_checker(problem) makes it (no GPU spectral detached estimator
        exists; the bench passes one built on an FP64 reference)."""
        r0, p0 = self.renderers[0], self.params[0]
        if self.hero:
            if self.fd_checker is None:
                return {"params": 0, "max_rel_err": 0.0, "median_rel_err": 0.0, "reference": "none given"}
            return self.fd_checker(self)
        if self.binned:
            res = binned.gradient_check(r0, p0.bp, self.targets[0], h=fd_step)
            rel = res["rel_err"]
        else:
            res = diff.gradient_check(r0, p0.theta().detach().clone(), self.targets[0], h=fd_step)
            live = res["live"] & p0.mask.numpy().astype(bool)
            rel = res["rel_err"][live]
        return {"params": int(rel.size), "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel))}


def optimise(problem: InverseProblem, steps: int = 200, lr: float = 1e-2, check_every: int = 50,
             fd_step: float = 1e-2, log=None) -> dict:
    """Adam over every scene's parameters; central-difference spot checks of scene 0's gradient
    (detached estimator, step fd_step) every check_every steps."""
    opt = torch.optim.Adam(problem.leaves(), lr=lr)
    checks = []
    err0 = problem.param_errors()
    img0 = problem.image_losses()
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    render_s = 0.0
    for step in range(steps):
        opt.zero_grad()
        tr = time.perf_counter()
        loss = problem.loss(step)
        loss.backward()
        torch.cuda.synchronize()
        render_s += time.perf_counter() - tr
        losses.append(float(loss.detach()))
        problem.mask_grads_()
        opt.step()
        problem.clamp_()
        if check_every and (step + 1) % check_every == 0:
            checks.append(dict(step=step + 1, **problem.fd_check(fd_step)))
            if log:
                log(f"step {step + 1}: loss {losses[-1]:.6g}, parameter error {problem.param_error():.4g}, "
                    f"FD check max rel err {checks[-1]['max_rel_err']:.3g}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    img1 = problem.image_losses()
    floor = img1["target_params"]
    samples = steps * sum(problem.samples_per_render)
    window = max(1, min(10, steps // 4))
    return {"steps": steps, "lr": lr, "wall_s": wall, "render_and_adjoint_s": render_s,
            "parameters": ("16-bin base-colour spectra rendered at four hero wavelengths in 380..720 nm, roughness, "
                           "metallic") if problem.hero else
                          ("16-bin base-colour spectra (linear-sRGB projection), roughness, metallic"
                           if problem.binned else "RGB base colour, roughness, metallic"),
            "renders_per_step": "loss render of a fresh sample set + the adjoint's render and replay of another",
            "grad_samples_per_s": samples / render_s,
            "loss_first_steps": float(np.mean(losses[:window])), "loss_last_steps": float(np.mean(losses[-window:])),
            "loss_window": window, "param_mae_initial": err0, "param_mae_final": problem.param_errors(),
            "target_spp": problem.target_spp,
            "image_loss": {"evaluation_spp": problem.target_spp, "initial": img0["current"],
                           "final": img1["current"], "target_params": floor,
                           "final_colours_target_roughness_metallic": img1["target_roughness_metallic"],
                           "final_excess_fraction": (img1["current"] - floor) / max(img0["current"] - floor, 1e-30)},
            "fd_checks": checks}
