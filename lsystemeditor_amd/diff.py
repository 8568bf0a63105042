"""Differentiable rendering: material parameters in, image out, gradients by the adjoint kernel.

The forward is the renderer's fixed-spp render (sample k of every pixel: sample index k, seed
k + 1); the backward is pathTracerAdjointRender (path replay with detached sampling, FP32 per
vertex, FP64 reduction). Parameters per material, in this order: base r, g, b, roughness,
metallic, specular, emission -- the MaterialGpu fields they overwrite are listed in PARAM_WORDS.
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi, render

PARAM_NAMES = ("r", "g", "b", "roughness", "metallic", "specular", "emission")
PARAM_WORDS = (0, 1, 2, 3, 4, 14, 9)  # word offsets in the 24-word MaterialGpu record


class DiffRenderer:
    """Fixed-spp renderer over one scene whose material table is the parameter vector."""

    def __init__(self, scene: abi.Scene, materials: np.ndarray, camera: np.ndarray, params: np.ndarray, width: int,
                 height: int, spp: int, env: abi.Environment | None = None, stream=None):
        self.scene = scene
        self.base = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, abi.MATERIAL_WORDS).copy()
        if self.base.shape[0] > 64:
            raise ValueError("the adjoint supports at most 64 materials")
        p = np.ascontiguousarray(params, dtype=np.float32).copy()
        p[6:8] = np.array([spp, spp], np.int32).view(np.float32)  # minSamples = maxSamplesPerPixel = spp
        self.params_host = p
        self.inputs = render.DeviceInputs(camera, p, env=env)
        self.width, self.height, self.spp = width, height, spp
        self.acc = render.Accumulator.allocate(width, height)
        self.stream = stream
        # per-sample primal radiance of the last forward, so the adjoint replays without re-tracing
        self.radiance = torch.empty((width * height * spp, 3), dtype=torch.float32, device="cuda")
        self.radiance_valid = False
        self.uploaded: torch.Tensor | None = None  # the parameters the scene's materials hold

    @property
    def material_count(self) -> int:
        return self.base.shape[0]

    def theta0(self) -> torch.Tensor:
        """Current parameters as a (materials, 7) float32 tensor."""
        return torch.from_numpy(self.base[:, list(PARAM_WORDS)].copy())

    def set_theta(self, theta: torch.Tensor) -> None:
        self.radiance_valid = False
        t = theta.detach().to("cpu", torch.float32)
        mats = self.base.copy()
        mats[:, list(PARAM_WORDS)] = t.numpy()
        abi.check(abi.lib().ptSceneUpdateMaterials(self.scene.handle, mats.ctypes.data, mats.shape[0],
                                                   render._stream_handle(self.stream)), "ptSceneUpdateMaterials")
        self.uploaded = t.clone()

    def forward(self, theta: torch.Tensor, seed_base: int = 0) -> torch.Tensor:
        """(height * width, 3) float32 image of the fixed-spp render at parameters theta. Sample k
        of a pixel uses seed seed_base + k + 1; the adjoint's samples are those of seed_base 0,
        whose per-sample radiance this render then stores for it."""
        self.set_theta(theta)
        capture = seed_base == 0
        if capture:
            abi.check(abi.lib().pathTracerWavefrontCaptureRadiance(self.radiance.data_ptr(), self.spp),
                      "pathTracerWavefrontCaptureRadiance")
        try:
            render.render_wavefront(self.scene, self.acc, self.inputs, first_iter=seed_base, passes=self.spp,
                                    stream=self.stream)
        finally:
            if capture:
                abi.lib().pathTracerWavefrontCaptureRadiance(None, 0)
        self.radiance_valid = capture
        self.last_theta = theta.detach().to("cpu", torch.float32).clone()
        return self.acc.acc[:, :3].clone()

    def adjoint(self, grad_image: torch.Tensor, reuse_forward: bool = True, seed_base: int = 0) -> torch.Tensor:
        """dLoss/dtheta (materials, 7) float64 for dLoss/dImage = grad_image (height * width, 3),
        at the current parameters, replaying the samples of seed base `seed_base` (sample k uses
        seed seed_base + k + 1). For seed base 0 the last forward's per-sample radiance is reused
        unless reuse_forward is False; other seed bases render their samples first."""
        g = grad_image.detach().to("cuda", torch.float32).contiguous()
        out = torch.zeros((self.material_count, 7), dtype=torch.float64, device="cuda")
        rad = self.radiance.data_ptr() if (reuse_forward and self.radiance_valid and seed_base == 0) else None
        abi.check(abi.lib().pathTracerAdjointRenderSeeded(
            g.data_ptr(), self.width, self.height, self.spp, self.inputs.camera.data_ptr(), self.scene.device,
            self.inputs.env_ptr, self.inputs.params.data_ptr(), self.material_count, rad, out.data_ptr(),
            seed_base + 1, render._stream_handle(self.stream)), "pathTracerAdjointRenderSeeded")
        return out

    def detached(self, eval_theta: torch.Tensor) -> torch.Tensor:
        """Image of the detached estimator: paths sampled with the scene's current materials,
        values evaluated with eval_theta (pathTracerDetachedRender)."""
        t = eval_theta.detach().to("cpu", torch.float32).numpy()
        mats = self.base.copy()
        mats[:, list(PARAM_WORDS)] = t
        d_mats = torch.from_numpy(mats).cuda()
        img = torch.empty((self.width * self.height, 3), dtype=torch.float32, device="cuda")
        abi.check(abi.lib().pathTracerDetachedRender(
            img.data_ptr(), self.width, self.height, self.spp, self.inputs.camera.data_ptr(), self.scene.device,
            d_mats.data_ptr(), self.inputs.env_ptr, self.inputs.params.data_ptr(), render._stream_handle(self.stream)),
            "pathTracerDetachedRender")
        torch.cuda.synchronize()
        return img

    def live_mask(self) -> np.ndarray:
        """(materials, 7) bool: parameters the forward is differentiable in."""
        t = self.base[:, list(PARAM_WORDS)]
        live = np.ones(t.shape, bool)
        live[:, 3:6] = (t[:, 3:6] > 0.0) & (t[:, 3:6] < 1.0)
        live[:, 6] = t[:, 6] > 0.0
        return live


SPECTRAL_PARAMS = abi.SPECTRUM_WORDS + 4  # kPtSpectrumGradParams
SPECTRAL_SCALAR_WORDS = (3, 4, 14, 9)       # roughness, metallic, specular, emission


class SpectralDiffRenderer(DiffRenderer):
    """Fixed-spp hero-wavelength render of a scene with binned spectra: the parameters of a
    material are a row of 36 -- its 16 reflectance bins, 16 emission bins (ptSceneSetSpectra),
    roughness, metallic, specular and emission -- and the backward is
    pathTracerAdjointRenderSpectral. Hero-wavelength sampling must be on
    (abi.Spectral.enable) while it renders. Materials whose spectra row starts with a negative
    value keep their rgb2spec colours; their bins are not parameters."""

    def __init__(self, scene: abi.Scene, materials: np.ndarray, spectra: np.ndarray, camera: np.ndarray,
                 params: np.ndarray, width: int, height: int, spp: int, env: abi.Environment | None = None,
                 stream=None):
        super().__init__(scene, materials, camera, params, width, height, spp, env=env, stream=stream)
        self.spectra = np.ascontiguousarray(spectra, dtype=np.float32).reshape(self.material_count,
                                                                                abi.SPECTRUM_WORDS).copy()
        self.radiance = torch.empty((width * height * spp, 4), dtype=torch.float32, device="cuda")
        scene.set_spectra(self.spectra)

    def theta0(self) -> torch.Tensor:
        """(materials, 36) float32: spectra, then roughness, metallic, specular, emission."""
        t = np.concatenate([self.spectra, self.base[:, list(SPECTRAL_SCALAR_WORDS)]], axis=1)
        return torch.from_numpy(np.ascontiguousarray(t))

    def _tables(self, theta: torch.Tensor) -> tuple[np.ndarray, np.ndarray]:
        t = theta.detach().to("cpu", torch.float32).numpy()
        mats = self.base.copy()
        mats[:, list(SPECTRAL_SCALAR_WORDS)] = t[:, abi.SPECTRUM_WORDS:]
        sp = self.spectra.copy()
        refl = sp[:, 0] >= 0.0
        emis = sp[:, abi.SPECTRUM_BINS] >= 0.0
        sp[refl, :abi.SPECTRUM_BINS] = t[refl, :abi.SPECTRUM_BINS]
        sp[emis, abi.SPECTRUM_BINS:] = t[emis, abi.SPECTRUM_BINS:abi.SPECTRUM_WORDS]
        return np.ascontiguousarray(mats), np.ascontiguousarray(sp)

    def set_theta(self, theta: torch.Tensor) -> None:
        self.radiance_valid = False
        mats, sp = self._tables(theta)
        abi.check(abi.lib().ptSceneUpdateMaterials(self.scene.handle, mats.ctypes.data, mats.shape[0],
                                                   render._stream_handle(self.stream)), "ptSceneUpdateMaterials")
        self.scene.set_spectra(sp, stream=render._stream_handle(self.stream))
        self.uploaded = theta.detach().to("cpu", torch.float32).clone()

    def adjoint(self, grad_image: torch.Tensor, reuse_forward: bool = True, seed_base: int = 0) -> torch.Tensor:
        """dLoss/dtheta (materials, 36) float64 for dLoss/dImage = grad_image, as DiffRenderer.adjoint
        (pathTracerAdjointRenderSpectral)."""
        g = grad_image.detach().to("cuda", torch.float32).contiguous()
        out = torch.zeros((self.material_count, SPECTRAL_PARAMS), dtype=torch.float64, device="cuda")
        rad = self.radiance.data_ptr() if (reuse_forward and self.radiance_valid and seed_base == 0) else None
        abi.check(abi.lib().pathTracerAdjointRenderSpectral(
            g.data_ptr(), self.width, self.height, self.spp, self.inputs.camera.data_ptr(), self.scene.device,
            self.inputs.env_ptr, self.inputs.params.data_ptr(), self.material_count, rad, out.data_ptr(),
            seed_base + 1, render._stream_handle(self.stream)), "pathTracerAdjointRenderSpectral")
        return out

    def detached(self, eval_theta: torch.Tensor) -> torch.Tensor:
        raise NotImplementedError("the spectral detached estimator is the oracle's (orc_render_detached_spectral)")

    def live_mask(self) -> np.ndarray:
        """(materials, 36) bool: the bins of binned spectra (reflectance of opaque materials,
        emission of emitting ones), roughness / metallic / specular inside (0, 1), and the emission
        scale of a material with binned emission."""
        live = np.zeros((self.material_count, SPECTRAL_PARAMS), bool)
        b = self.base
        opaque = (np.clip(b[:, 5], 0, 1) == 0.0) & (np.clip(b[:, 8], 0, 1) == 0.0)
        emits = b[:, 9] > 0.0
        live[:, :abi.SPECTRUM_BINS] = ((self.spectra[:, 0] >= 0.0) & opaque)[:, None]
        emis = (self.spectra[:, abi.SPECTRUM_BINS] >= 0.0) & emits
        live[:, abi.SPECTRUM_BINS:abi.SPECTRUM_WORDS] = emis[:, None]
        for q, wd in enumerate(SPECTRAL_SCALAR_WORDS[:3]):
            live[:, abi.SPECTRUM_WORDS + q] = (b[:, wd] > 0.0) & (b[:, wd] < 1.0)
        live[:, abi.SPECTRUM_WORDS + 3] = emis
        return live


def gradient_check(r: DiffRenderer, theta0: torch.Tensor, target: torch.Tensor, h: float = 1e-3) -> dict:
    """Adjoint gradient of L = sum (image - target)^2 at theta0 against central differences
    (step h) of the detached estimator, for every live parameter."""
    img = r.forward(theta0)
    grad_img = (2.0 * (img - target)).contiguous()
    adj = r.adjoint(grad_img).cpu().numpy()
    live = r.live_mask()
    fd = np.zeros_like(adj)
    for i in range(adj.shape[0]):
        for k in range(adj.shape[1]):
            if not live[i, k]:
                continue
            losses = []
            for sgn in (1.0, -1.0):
                th = theta0.detach().cpu().clone()
                th[i, k] += sgn * h
                im = r.detached(th)
                losses.append(float(((im.double() - target.double()) ** 2).sum().item()))
            fd[i, k] = (losses[0] - losses[1]) / (2.0 * h)
    floor = 1e-3 * float(np.abs(fd[live]).max())
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), floor)
    rel[~live] = 0.0
    return {"adjoint": adj, "fd": fd, "live": live, "rel_err": rel, "max_rel_err": float(rel[live].max()),
            "median_rel_err": float(np.median(rel[live]))}


class RenderFunction(torch.autograd.Function):
    """image = render(theta); backward through the adjoint kernel."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, renderer: DiffRenderer) -> torch.Tensor:
        ctx.renderer = renderer
        ctx.theta = theta.detach().clone()
        return renderer.forward(theta)

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        # the adjoint replays the paths of the primal parameters; the stored per-sample radiance
        # is reused when no other forward ran in between
        r = ctx.renderer
        if not r.radiance_valid or not torch.equal(r.last_theta, ctx.theta.cpu()):
            r.set_theta(ctx.theta)
        grad = r.adjoint(grad_image)
        return grad.to(ctx.theta.device, ctx.theta.dtype), None


def render_image(theta: torch.Tensor, renderer: DiffRenderer) -> torch.Tensor:
    return RenderFunction.apply(theta, renderer)


class DecorrelatedRenderFunction(torch.autograd.Function):
    """image = render(theta) with the samples of seed base `seed_base`; the backward replays the
    independent samples of seed base `adjoint_base` (rendered again inside the adjoint). The loss
    residual and the path derivatives then come from independent samples, so the gradient of a
    loss that is nonlinear in the image (e.g. L2) carries no sample covariance term. Costs one more
    render per step."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, renderer: DiffRenderer, seed_base: int, adjoint_base: int) -> torch.Tensor:
        if seed_base == adjoint_base:
            raise ValueError("the forward's and the adjoint's seed bases must differ")
        ctx.renderer = renderer
        ctx.theta = theta.detach().clone()
        ctx.adjoint_base = adjoint_base
        return renderer.forward(theta, seed_base=seed_base)

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        r = ctx.renderer
        if not torch.equal(r.last_theta, ctx.theta.cpu()):
            r.set_theta(ctx.theta)
        grad = r.adjoint(grad_image, reuse_forward=False, seed_base=ctx.adjoint_base)
        return grad.to(ctx.theta.device, ctx.theta.dtype), None, None, None


def render_image_decorrelated(theta: torch.Tensor, renderer: DiffRenderer, seed_base: int,
                              adjoint_base: int) -> torch.Tensor:
    return DecorrelatedRenderFunction.apply(theta, renderer, seed_base, adjoint_base)


class BatchDiffRenderer:
    """The fixed-spp renders and adjoints of several DiffRenderers (one spp for all) as one batch
    wavefront (pathTracerRenderFixedBatch / pathTracerAdjointRenderBatch): images concatenated
    in renderer order, each equal to that renderer's own render; gradients per renderer."""

    def __init__(self, renderers: list[DiffRenderer]):
        if not renderers or len({r.spp for r in renderers}) != 1:
            raise ValueError("a batch needs at least one renderer, all with the same spp")
        if len({isinstance(r, SpectralDiffRenderer) for r in renderers}) != 1:
            raise ValueError("a batch holds RGB renderers or spectral renderers, not both")
        self.renderers = list(renderers)
        self.spp = renderers[0].spp
        self.batch = render.SceneBatch([r.scene for r in renderers], [r.inputs for r in renderers],
                                       [(r.width, r.height) for r in renderers])
        # one group of paths for the whole batch where it fits 64M paths: every launch's slowest
        # rays then cost once per bounce per render (the adjoint uses the reserved size too)
        self.max_paths = int(min(max(self.batch.total_pixels * self.spp, 1 << 24), 1 << 26))
        self.batch.reserve(self.max_paths)

    def pixels(self, i: int) -> slice:
        return self.batch.pixels(i)

    def set_thetas(self, thetas) -> None:
        """Uploads every renderer's parameters that differ from what its scene holds."""
        for r, th in zip(self.renderers, thetas):
            t = th.detach().to("cpu", torch.float32)
            if r.uploaded is None or not torch.equal(r.uploaded, t):
                r.set_theta(t)

    def forward(self, thetas, seed_base: int = 0) -> torch.Tensor:
        """(total pixels, 3) images at parameters thetas; sample k uses seed seed_base + k + 1."""
        self.set_thetas(thetas)
        self.batch.render_fixed(self.spp, first_seed=seed_base + 1, max_paths=self.max_paths)
        return self.batch.acc.acc[:, :3].clone()

    def adjoint(self, grad_image: torch.Tensor, seed_base: int) -> list[torch.Tensor]:
        """Per renderer (materials, 7) float64 dLoss/dtheta at the uploaded parameters for
        dLoss/dImage = grad_image (total pixels, 3), replaying the samples of seed base seed_base
        ((materials, 36) for SpectralDiffRenderers, rendered at hero wavelengths)."""
        g = grad_image.detach().to("cuda", torch.float32).contiguous()
        cols = SPECTRAL_PARAMS if isinstance(self.renderers[0], SpectralDiffRenderer) else 7
        out = torch.zeros((self.batch.grad_rows, cols), dtype=torch.float64, device="cuda")
        abi.check(abi.lib().pathTracerAdjointRenderBatch(g.data_ptr(), self.batch.handle, self.spp, seed_base + 1,
                                                         out.data_ptr(), None), "pathTracerAdjointRenderBatch")
        return [out[o:o + r.material_count] for o, r in zip(self.batch.grad_row_offsets, self.renderers)]


class BatchDecorrelatedRenderFunction(torch.autograd.Function):
    """DecorrelatedRenderFunction over a batch: images = render(thetas) with the samples of
    seed base `seed_base`; the backward replays the independent samples of `adjoint_base`."""

    @staticmethod
    def forward(ctx, br: BatchDiffRenderer, seed_base: int, adjoint_base: int, *thetas):
        if seed_base == adjoint_base:
            raise ValueError("the forward's and the adjoint's seed bases must differ")
        ctx.br = br
        ctx.thetas = [t.detach().clone() for t in thetas]
        ctx.adjoint_base = adjoint_base
        return br.forward(thetas, seed_base=seed_base)

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        br = ctx.br
        br.set_thetas(ctx.thetas)
        grads = br.adjoint(grad_image, ctx.adjoint_base)
        return (None, None, None, *[g.to(t.device, t.dtype) for g, t in zip(grads, ctx.thetas)])


def render_batch_decorrelated(thetas, br: BatchDiffRenderer, seed_base: int, adjoint_base: int) -> torch.Tensor:
    return BatchDecorrelatedRenderFunction.apply(br, seed_base, adjoint_base, *thetas)
