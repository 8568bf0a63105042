"""Batched inverse rendering (workload C4): recover material parameters of many L-system scenes from
target images with Adam through the adjoint kernel.

Each scene is a stochastic bush of the C1 generator (seed s, iterations drawn from 3..6, 256x256,
16 spp, max depth 8) whose bark, leaf and ground groups carry a seeded target material set (every
optimised parameter scaled by an independent factor in [0.8, 1.2]). The target image is the
fixed-spp render at the target parameters; the optimisation starts from the unperturbed
parameters and minimises the summed L2 image loss over all scenes with Adam (lr 1e-2) on base
colour, roughness and metallic. The target is rendered once (seed base 0); every step renders
the loss image with a fresh sample set and replays another fresh, independent set in the
adjoint, so each step's gradient is an unbiased estimate of the gradient of the expected-image
loss |E[image](theta) - target|^2. With the same samples for the residual and the derivatives
the gradient would carry the covariance of image noise and its derivative, and with fixed
sample sets the sampling error would be a fixed function of the parameters: at 16 spp either
drives poorly determined parameters away from the target.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import abi, diff, scenes

OPT_PARAMS = (0, 1, 2, 3, 4)          # base r, g, b, roughness, metallic (diff.PARAM_NAMES order)
OPT_GROUPS = (0, 1, 2)                # bark, leaf, ground; the two lights stay fixed


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
    """The unperturbed parameters with the optimised ones replaced by their seeded perturbation."""
    pert = scenes.perturbed_target(theta0, seed=seed, spread=spread)
    out = theta0.copy()
    for gidx in OPT_GROUPS:
        out[gidx, list(OPT_PARAMS)] = pert[gidx, list(OPT_PARAMS)]
    return out


class InverseProblem:
    """Scenes, renderers, target images and the parameters being optimised."""

    def __init__(self, specs: list[scenes.SceneSpec], bvh: str = "sah"):
        self.specs = specs
        self.renderers = []
        self.targets = []
        self.theta_true = []
        self.theta = []
        mask = torch.zeros((specs[0].materials.shape[0], len(diff.PARAM_NAMES)))
        for gidx in OPT_GROUPS:
            mask[gidx, list(OPT_PARAMS)] = 1.0
        self.mask = mask
        for i, spec in enumerate(specs):
            sc = abi.Scene(spec.triangles, spec.materials, bvh=bvh)
            r = diff.DiffRenderer(sc, spec.materials, spec.camera, spec.params, spec.width, spec.height, spec.spp)
            th0 = r.theta0().numpy()
            tt = torch.from_numpy(target_theta(th0, seed=int(spec.info.get("seed", i))))
            self.targets.append(r.forward(tt, seed_base=0).detach().clone())
            self.theta_true.append(tt)
            self.theta.append(torch.from_numpy(th0.copy()).requires_grad_(True))
            self.renderers.append(r)
        self.samples_per_render = [s.width * s.height * s.spp for s in specs]
        # every scene's loss render and adjoint in one batch wavefront
        self.batch = diff.BatchDiffRenderer(self.renderers)
        self.target = torch.cat(self.targets)

    def loss(self, step: int = 0) -> torch.Tensor:
        """Summed L2 image loss of step `step`: the loss render uses seed base (2 step + 1) spp,
        the adjoint replays seed base (2 step + 2) spp (the target is seed base 0); all scenes in
        one batch render and one batch adjoint."""
        spp = self.batch.spp
        img = diff.render_batch_decorrelated(self.theta, self.batch, (2 * step + 1) * spp, (2 * step + 2) * spp)
        return ((img - self.target) ** 2).sum().double()

    def param_error(self, params=OPT_PARAMS) -> float:
        """Mean absolute error against the targets over the optimised groups and the given
        parameter columns (default: every optimised parameter)."""
        errs = []
        cols = list(params)
        for th, tt in zip(self.theta, self.theta_true):
            d = (th.detach() - tt).abs()[list(OPT_GROUPS)][:, cols]
            errs.append(d.reshape(-1))
        return float(torch.cat(errs).mean())

    def param_errors(self) -> dict:
        return {"all": self.param_error(), "base_color": self.param_error((0, 1, 2)),
                "roughness": self.param_error((3,)), "metallic": self.param_error((4,))}

    def clamp_(self) -> None:
        with torch.no_grad():
            for th in self.theta:
                th[:, 0:3].clamp_(0.0, 1.0)
                th[:, 3:5].clamp_(0.01, 0.99)


def optimise(problem: InverseProblem, steps: int = 200, lr: float = 1e-2, check_every: int = 50,
             fd_step: float = 1e-2, log=None) -> dict:
    """Adam over every scene's parameters; central-difference spot checks of scene 0's gradient
    (detached estimator, step fd_step) every check_every steps."""
    opt = torch.optim.Adam(problem.theta, lr=lr)
    mask = problem.mask
    checks = []
    err0 = problem.param_errors()
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
        for th in problem.theta:
            th.grad.mul_(mask)
        opt.step()
        problem.clamp_()
        if check_every and (step + 1) % check_every == 0:
            r0 = problem.renderers[0]
            res = diff.gradient_check(r0, problem.theta[0].detach().clone(), problem.targets[0], h=fd_step)
            live = res["live"] & mask.numpy().astype(bool)
            rel = res["rel_err"][live]
            checks.append({"step": step + 1, "params": int(live.sum()), "max_rel_err": float(rel.max()),
                           "median_rel_err": float(np.median(rel))})
            if log:
                log(f"step {step + 1}: loss {losses[-1]:.6g}, parameter error {problem.param_error():.4g}, "
                    f"FD check max rel err {checks[-1]['max_rel_err']:.3g}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    samples = steps * sum(problem.samples_per_render)
    window = max(1, min(10, steps // 4))
    return {"steps": steps, "lr": lr, "wall_s": wall, "render_and_adjoint_s": render_s,
            "renders_per_step": "loss render of a fresh sample set + the adjoint's render and replay of another",
            "grad_samples_per_s": samples / render_s,
            "loss_first_steps": float(np.mean(losses[:window])), "loss_last_steps": float(np.mean(losses[-window:])),
            "loss_window": window, "param_mae_initial": err0, "param_mae_final": problem.param_errors(),
            "fd_checks": checks}
