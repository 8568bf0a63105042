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

    @property
    def material_count(self) -> int:
        return self.base.shape[0]

    def theta0(self) -> torch.Tensor:
        """Current parameters as a (materials, 7) float32 tensor."""
        return torch.from_numpy(self.base[:, list(PARAM_WORDS)].copy())

    def set_theta(self, theta: torch.Tensor) -> None:
        t = theta.detach().to("cpu", torch.float32).numpy()
        mats = self.base.copy()
        mats[:, list(PARAM_WORDS)] = t
        abi.check(abi.lib().ptSceneUpdateMaterials(self.scene.handle, mats.ctypes.data, mats.shape[0],
                                                   render._stream_handle(self.stream)), "ptSceneUpdateMaterials")

    def forward(self, theta: torch.Tensor) -> torch.Tensor:
        """(height * width, 3) float32 image of the fixed-spp render at parameters theta."""
        self.set_theta(theta)
        render.render_wavefront(self.scene, self.acc, self.inputs, passes=self.spp, stream=self.stream)
        return self.acc.acc[:, :3].clone()

    def adjoint(self, grad_image: torch.Tensor) -> torch.Tensor:
        """dLoss/dtheta (materials, 7) float64 for dLoss/dImage = grad_image (height * width, 3)."""
        g = grad_image.detach().to("cuda", torch.float32).contiguous()
        out = torch.zeros((self.material_count, 7), dtype=torch.float64, device="cuda")
        abi.check(abi.lib().pathTracerAdjointRender(
            g.data_ptr(), self.width, self.height, self.spp, self.inputs.camera.data_ptr(), self.scene.device,
            self.inputs.env_ptr, self.inputs.params.data_ptr(), self.material_count, out.data_ptr(),
            render._stream_handle(self.stream)), "pathTracerAdjointRender")
        return out


class RenderFunction(torch.autograd.Function):
    """image = render(theta); backward through the adjoint kernel."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, renderer: DiffRenderer) -> torch.Tensor:
        ctx.renderer = renderer
        ctx.theta = theta.detach().clone()
        return renderer.forward(theta)

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        # the adjoint replays the paths of the primal parameters
        ctx.renderer.set_theta(ctx.theta)
        grad = ctx.renderer.adjoint(grad_image)
        return grad.to(ctx.theta.device, ctx.theta.dtype), None


def render_image(theta: torch.Tensor, renderer: DiffRenderer) -> torch.Tensor:
    return RenderFunction.apply(theta, renderer)
