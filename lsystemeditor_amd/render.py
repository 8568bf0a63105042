"""Host driver for the GPU renderer: accumulator buffers and the render-loop schedule.

Mirrors PathTracer::renderLoop (PathTracerLib/src/PathTracer.cpp:2220-2543) without the GUI
parts: globalSeed = pass + 1, one adaptive pass per iteration over the active list, active-list
compaction every 4 full-resolution passes (kAdaptiveCompactActiveListInterval) with a 4-byte
count readback, stop when no pixel is active. Device memory and streams come from torch; all
compute goes through libpt_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import abi

COMPACT_INTERVAL = 4  # PathTracerAdaptiveSampling.h:7


def _stream_handle(stream: torch.cuda.Stream | None) -> int:
    s = stream or torch.cuda.current_stream()
    return int(s.cuda_stream)


@dataclass
class Accumulator:
    """Per-pixel device state (PathTracer.cpp:323-419): acc float4, counts u32, lumMean, m2,
    converged u8, two active lists and the device active count."""
    width: int
    height: int
    acc: torch.Tensor
    counts: torch.Tensor
    mean: torch.Tensor
    m2: torch.Tensor
    conv: torch.Tensor
    active: torch.Tensor
    scratch: torch.Tensor
    active_count: torch.Tensor

    @classmethod
    def allocate(cls, width: int, height: int, device: str = "cuda") -> "Accumulator":
        n = width * height
        return cls(width, height,
                   torch.zeros((n, 4), dtype=torch.float32, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(n, dtype=torch.float32, device=device),
                   torch.zeros(n, dtype=torch.float32, device=device),
                   torch.zeros(n, dtype=torch.uint8, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(1, dtype=torch.int32, device=device))

    def clear(self, stream=None) -> None:
        abi.check(abi.lib().pathTracerClearAccumulator(
            self.acc.data_ptr(), self.counts.data_ptr(), self.mean.data_ptr(), self.m2.data_ptr(),
            self.conv.data_ptr(), self.active.data_ptr(), self.scratch.data_ptr(), self.active_count.data_ptr(),
            self.width, self.height, _stream_handle(stream)), "pathTracerClearAccumulator")

    def host(self) -> dict[str, np.ndarray]:
        torch.cuda.synchronize()
        return {"acc": self.acc.cpu().numpy(), "counts": self.counts.cpu().numpy().view(np.uint32),
                "mean": self.mean.cpu().numpy(), "m2": self.m2.cpu().numpy(), "conv": self.conv.cpu().numpy()}


class DeviceInputs:
    """Camera, params and (optional) environment map in device memory."""

    def __init__(self, camera: np.ndarray, params: np.ndarray, env: abi.Environment | None = None,
                 device: str = "cuda"):
        self.camera = torch.from_numpy(np.ascontiguousarray(camera, dtype=np.float32)).to(device)
        self.params = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(device)
        self.env = env  # keeps the device tables alive
        self.env_ptr = env.device if env is not None else None
        self.params_host = np.ascontiguousarray(params, dtype=np.float32)

    @property
    def max_spp(self) -> int:
        return int(self.params_host[7:8].view(np.int32)[0])


def adaptive_pass(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, active_count: int, seed: int,
                  stream=None) -> None:
    abi.check(abi.lib().pathTracerAdaptiveSample(
        acc.acc.data_ptr(), acc.counts.data_ptr(), acc.mean.data_ptr(), acc.m2.data_ptr(), acc.conv.data_ptr(),
        acc.active.data_ptr(), active_count, acc.width, acc.height, inputs.camera.data_ptr(), scene.device,
        inputs.env_ptr, inputs.params.data_ptr(), seed, _stream_handle(stream)), "pathTracerAdaptiveSample")


def compact(acc: Accumulator, inputs: DeviceInputs, active_count: int, region: tuple | None = None,
            stream=None) -> int:
    """Compacts the active list into scratch (optionally keeping only pixels inside
    region = (minX, minY, maxX, maxY)), swaps the lists and returns the new count."""
    r = region or (0, 0, 0, 0)
    abi.check(abi.lib().pathTracerCompactActiveList(
        acc.active.data_ptr(), acc.scratch.data_ptr(), acc.active_count.data_ptr(), acc.conv.data_ptr(),
        acc.counts.data_ptr(), inputs.max_spp, active_count, acc.width, acc.height, region is not None,
        r[0], r[1], r[2], r[3], _stream_handle(stream)), "pathTracerCompactActiveList")
    count = int(acc.active_count.item())
    acc.active, acc.scratch = acc.scratch, acc.active
    return count


def clear_region(acc: Accumulator, region: tuple, stream=None) -> int:
    """Resets region = (minX, minY, maxX, maxY) and makes it the active list; returns its size."""
    abi.check(abi.lib().pathTracerClearRegionAccumulator(
        acc.acc.data_ptr(), acc.counts.data_ptr(), acc.mean.data_ptr(), acc.m2.data_ptr(), acc.conv.data_ptr(),
        acc.active.data_ptr(), acc.active_count.data_ptr(), acc.width, acc.height, region[0], region[1], region[2],
        region[3], _stream_handle(stream)), "pathTracerClearRegionAccumulator")
    return int(acc.active_count.item())


def rebuild_active(acc: Accumulator, inputs: DeviceInputs, stream=None) -> int:
    """Lists every pixel still sampling (pathTracerRebuildActiveList); returns the count."""
    abi.check(abi.lib().pathTracerRebuildActiveList(
        acc.conv.data_ptr(), acc.counts.data_ptr(), acc.active.data_ptr(), acc.active_count.data_ptr(), acc.width,
        acc.height, inputs.max_spp, _stream_handle(stream)), "pathTracerRebuildActiveList")
    return int(acc.active_count.item())


def preview(scene: abi.Scene, inputs: DeviceInputs, width: int, height: int, seed: int,
            stream=None) -> torch.Tensor:
    """One unjittered sample per pixel into a fresh (width*height, 4) buffer (pathTracerPreviewSample)."""
    buf = torch.zeros((width * height, 4), dtype=torch.float32, device="cuda")
    abi.check(abi.lib().pathTracerPreviewSample(buf.data_ptr(), width, height, inputs.camera.data_ptr(), scene.device,
                                                inputs.env_ptr, inputs.params.data_ptr(), seed,
                                                _stream_handle(stream)), "pathTracerPreviewSample")
    return buf


def display(acc: Accumulator, inputs: DeviceInputs, exposure: float = 1.0, stream=None) -> torch.Tensor:
    """Tone-mapped RGBA8 of the accumulator (pathTracerCopyToPbo): (height*width, 4) uint8."""
    pbo = torch.zeros((acc.width * acc.height, 4), dtype=torch.uint8, device="cuda")
    abi.check(abi.lib().pathTracerCopyToPbo(acc.acc.data_ptr(), acc.counts.data_ptr(), acc.conv.data_ptr(),
                                            pbo.data_ptr(), acc.width, acc.height, inputs.params.data_ptr(),
                                            exposure, _stream_handle(stream)), "pathTracerCopyToPbo")
    return pbo


def upsample_preview(preview_buf: torch.Tensor, preview_w: int, preview_h: int, downscale: int, full_w: int,
                     full_h: int, inputs: DeviceInputs, exposure: float = 1.0, stream=None) -> torch.Tensor:
    """Nearest-neighbour upsampled, tone-mapped preview (pathTracerUpsamplePreviewToPbo)."""
    pbo = torch.zeros((full_w * full_h, 4), dtype=torch.uint8, device="cuda")
    abi.check(abi.lib().pathTracerUpsamplePreviewToPbo(preview_buf.data_ptr(), preview_w, preview_h, downscale,
                                                       pbo.data_ptr(), full_w, full_h, inputs.params.data_ptr(),
                                                       exposure, _stream_handle(stream)),
              "pathTracerUpsamplePreviewToPbo")
    return pbo


def uv_debug(scene: abi.Scene, inputs: DeviceInputs, width: int, height: int, stream=None) -> torch.Tensor:
    """Fractional hit UV as red/green (pathTracerWriteUvDebugToPbo): (height*width, 4) uint8."""
    pbo = torch.zeros((width * height, 4), dtype=torch.uint8, device="cuda")
    abi.check(abi.lib().pathTracerWriteUvDebugToPbo(inputs.camera.data_ptr(), scene.device, pbo.data_ptr(), width,
                                                    height, inputs.params.data_ptr(), _stream_handle(stream)),
              "pathTracerWriteUvDebugToPbo")
    return pbo


def render_loop(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, first_iter: int = 0,
                max_passes: int = 1 << 30, stream=None) -> int:
    """The reference schedule, one launch per pass. Returns the number of passes run."""
    acc.clear(stream)
    active = acc.width * acc.height
    passes = 0
    it = first_iter
    while active > 0 and passes < max_passes:
        adaptive_pass(scene, acc, inputs, active, it + 1, stream)
        passes += 1
        if (it - first_iter + 1) % COMPACT_INTERVAL == 0:
            active = compact(acc, inputs, active, stream)
        it += 1
    return passes


def render_persistent(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, first_iter: int = 0,
                      passes: int | None = None, clear: bool = True, stream=None) -> None:
    """Same per-pixel result as render_loop in one launch (pathTracerAdaptiveRenderPasses)."""
    if clear:
        acc.clear(stream)
    if passes is None:
        passes = inputs.max_spp
        if passes <= 0:
            raise ValueError("render_persistent needs a finite pass count when maxSamplesPerPixel is 0")
    abi.check(abi.lib().pathTracerAdaptiveRenderPasses(
        acc.acc.data_ptr(), acc.counts.data_ptr(), acc.mean.data_ptr(), acc.m2.data_ptr(), acc.conv.data_ptr(),
        acc.active.data_ptr(), acc.width * acc.height, acc.width, acc.height, inputs.camera.data_ptr(),
        scene.device, inputs.env_ptr, inputs.params.data_ptr(), first_iter + 1, passes, _stream_handle(stream)),
        "pathTracerAdaptiveRenderPasses")


def resolved_depth(params_host: np.ndarray) -> int:
    """maxPathDepth as the integrator resolves it (PathIntegratorRandCore.h:219-227)."""
    d = int(params_host[4:5].view(np.int32)[0])
    return d if d > 0 else (256 if d == 0 else 32)


def render_wavefront(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, first_iter: int = 0,
                     passes: int | None = None, clear: bool = True, max_paths: int = 0, stream=None) -> None:
    """Same per-pixel result as render_loop, wavefront-scheduled (pathTracerAdaptiveRenderWavefront)."""
    if clear:
        acc.clear(stream)
    if passes is None:
        passes = inputs.max_spp
        if passes <= 0:
            raise ValueError("render_wavefront needs a finite pass count when maxSamplesPerPixel is 0")
    abi.check(abi.lib().pathTracerAdaptiveRenderWavefront(
        acc.acc.data_ptr(), acc.counts.data_ptr(), acc.mean.data_ptr(), acc.m2.data_ptr(), acc.conv.data_ptr(),
        acc.active.data_ptr(), acc.width * acc.height, acc.width, acc.height, inputs.camera.data_ptr(),
        scene.device, inputs.env_ptr, inputs.params.data_ptr(), first_iter + 1, passes,
        resolved_depth(inputs.params_host), max_paths, _stream_handle(stream)), "pathTracerAdaptiveRenderWavefront")


def trace_rays(scene: abi.Scene, origins: np.ndarray, dirs: np.ndarray, tmin: float = 0.0,
               tmax: float = 1.0e6) -> np.ndarray:
    """Closest hits through the renderer's traversal: (N, 8) float32 (hit, t, normal, uv, tri bits)."""
    o = torch.from_numpy(np.ascontiguousarray(origins, dtype=np.float32)).cuda()
    d = torch.from_numpy(np.ascontiguousarray(dirs, dtype=np.float32)).cuda()
    out = torch.zeros((o.shape[0], 8), dtype=torch.float32, device="cuda")
    abi.check(abi.lib().pathTracerTraceRays(scene.device, o.data_ptr(), d.data_ptr(), o.shape[0], tmin, tmax,
                                            out.data_ptr(), _stream_handle(None)), "pathTracerTraceRays")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def count_traversal(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, active_count: int,
                    seed: int) -> np.ndarray:
    """Per-sample (internal nodes fetched, triangles tested) for one pass: (active_count, 4) uint32."""
    out = torch.zeros((active_count, 4), dtype=torch.int32, device="cuda")
    abi.check(abi.lib().pathTracerCountTraversal(
        acc.counts.data_ptr(), acc.active.data_ptr(), active_count, acc.width, acc.height, inputs.camera.data_ptr(),
        scene.device, inputs.env_ptr, inputs.params.data_ptr(), seed, out.data_ptr(), _stream_handle(None)),
        "pathTracerCountTraversal")
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def workspace_pins() -> int:
    """Pins on the shared wavefront workspace: one per graph-captured fixed-spp render; while any
    is held the workspace does not move and calls that need a larger one fail."""
    return int(abi.lib().pathTracerWavefrontPins())


def release_workspace_pins(count: int | None = None) -> None:
    """Releases `count` pins (all when None); call it once the graphs that took them are gone."""
    n = workspace_pins() if count is None else count
    abi.check(abi.lib().pathTracerWavefrontUnpin(n), "pathTracerWavefrontUnpin")


def render_fixed(scene: abi.Scene, acc: Accumulator, inputs: DeviceInputs, spp: int, max_paths: int = 0,
                 clear: bool = True, stream=None) -> None:
    """Fixed-spp wavefront render without host synchronisation (pathTracerRenderFixedWavefront);
    capturable into a CUDA/HIP graph once the workspace is reserved."""
    if clear:
        acc.clear(stream)
    abi.check(abi.lib().pathTracerRenderFixedWavefront(
        acc.acc.data_ptr(), acc.counts.data_ptr(), acc.mean.data_ptr(), acc.m2.data_ptr(), acc.conv.data_ptr(),
        acc.width, acc.height, inputs.camera.data_ptr(), scene.device, inputs.env_ptr, inputs.params.data_ptr(),
        spp, resolved_depth(inputs.params_host), max_paths, _stream_handle(stream)), "pathTracerRenderFixedWavefront")


class SceneBatch:
    """Scenes rendered in one wavefront (ptBatchCreate): their pixels concatenated in one
    accumulator, each scene's result equal to that of its own fixed-spp render."""

    def __init__(self, scenes: list, inputs: list, sizes: list[tuple[int, int]]):
        if not (len(scenes) == len(inputs) == len(sizes)) or not scenes:
            raise ValueError("one input set and size per scene, at least one scene")
        n = len(scenes)
        entries = (abi.PtBatchSceneGpu * n)()
        for e, sc, inp, (w, h) in zip(entries, scenes, inputs, sizes):
            e.scene = sc.device
            e.camera = inp.camera.data_ptr()
            e.params = inp.params.data_ptr()
            e.env = inp.env_ptr
            e.width = w
            e.height = h
        L = abi.lib()
        self.handle = L.ptBatchCreate(entries, n)
        if not self.handle:
            raise RuntimeError(f"ptBatchCreate failed: {abi.last_error()}")
        self.scenes, self.inputs, self.sizes = list(scenes), list(inputs), list(sizes)
        tp, rows, depth = C.c_int(), C.c_int(), C.c_int()
        po = (C.c_int * n)()
        ro = (C.c_int * n)()
        abi.check(L.ptBatchInfo(self.handle, C.byref(tp), C.byref(rows), C.byref(depth), po, ro), "ptBatchInfo")
        self.total_pixels, self.grad_rows, self.max_depth = tp.value, rows.value, depth.value
        self.pixel_offsets = list(po)
        self.grad_row_offsets = list(ro)
        self.acc = Accumulator.allocate(self.total_pixels, 1)

    def __len__(self) -> int:
        return len(self.scenes)

    def reserve(self, max_paths: int = 0) -> None:
        """Sizes the shared workspace (needed before graph capture)."""
        paths = max(max_paths if max_paths > 0 else (1 << 26), self.total_pixels)
        abi.check(abi.lib().pathTracerWavefrontReserve(paths, self.total_pixels, self.max_depth),
                  "pathTracerWavefrontReserve")

    def render_fixed(self, spp: int, first_seed: int = 1, max_paths: int = 0, clear: bool = True,
                     stream=None) -> None:
        """Every scene's fixed-spp render (sample k with seed first_seed + k) in one wavefront."""
        if clear:
            self.acc.clear(stream)
        a = self.acc
        abi.check(abi.lib().pathTracerRenderFixedBatch(
            a.acc.data_ptr(), a.counts.data_ptr(), a.mean.data_ptr(), a.m2.data_ptr(), a.conv.data_ptr(), self.handle,
            spp, max_paths, first_seed, _stream_handle(stream)), "pathTracerRenderFixedBatch")

    def pixels(self, i: int) -> slice:
        w, h = self.sizes[i]
        return slice(self.pixel_offsets[i], self.pixel_offsets[i] + w * h)

    def close(self) -> None:
        if self.handle:
            abi.lib().ptBatchDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
