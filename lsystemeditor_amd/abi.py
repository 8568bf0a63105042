"""ctypes binding of libpt_hip.so (include/pt_hip.h) and packers for its POD records.

The library is the product: every call here goes to the gfx950 kernels. Loading fails
loudly when the library has not been built (python -m lsystemeditor_amd.build).
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libpt_hip.so"
HEADER = PKG.parent / "include" / "pt_hip.h"
PROCEDURAL_HEADER = PKG.parent / "include" / "pt_procedural.h"
TRACER_HEADER = PKG.parent / "include" / "pt_tracer.h"

# Record sizes in 32-bit words (reference layouts, include/pt_hip.h).
HOST_TRIANGLE_WORDS = 25   # HostTriangle, 100 bytes
TRIANGLE_GPU_WORDS = 28    # reference TriangleGpu, 112 bytes (oracle input)
MATERIAL_WORDS = 24        # MaterialGpu, 96 bytes
TEXTURE_WORDS = 16         # TextureDescGpu, 64 bytes
CAMERA_WORDS = 20          # CameraGpu, 80 bytes
PARAMS_WORDS = 11          # RenderParamsGpu, 44 bytes
BVH_NODE_WORDS = 12        # MeshBvhNode, 48 bytes
SCENE_DESC_BYTES = 144     # MeshAccelSceneGpu
SPECTRUM_BINS = 16         # kPtSpectrumBins
SPECTRUM_WORDS = 32        # kPtSpectrumWords: reflectance bins, then emission bins


class LibraryMissing(ImportError):
    pass


_lib = None

_b = C.c_bool
_i = C.c_int
_u = C.c_uint
_f = C.c_float
_p = C.c_void_p


class Vec3C(C.Structure):
    """Vec3 passed by value (three floats)."""
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

SIGNATURES = {
    "pathTracerLastError": (C.c_char_p, []),
    "pathTracerClearAccumulator": (_b, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p]),
    "pathTracerClearRegionAccumulator": (_b, [_p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _p]),
    "pathTracerRebuildActiveList": (_b, [_p, _p, _p, _p, _i, _i, _i, _p]),
    "pathTracerAdaptiveSample": (_b, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _u, _p]),
    "pathTracerCompactActiveList": (_b, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _b, _i, _i, _i, _i, _p]),
    "pathTracerCompactActiveListNoAlloc": (_b, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _b, _i, _i, _i, _i, _p]),
    "pathTracerPreviewSample": (_b, [_p, _i, _i, _p, _p, _p, _p, _u, _p]),
    "pathTracerCopyToPbo": (_b, [_p, _p, _p, _p, _i, _i, _p, _f, _p]),
    "pathTracerUpsamplePreviewToPbo": (_b, [_p, _i, _i, _i, _p, _i, _i, _p, _f, _p]),
    "pathTracerWriteUvDebugToPbo": (_b, [_p, _p, _p, _i, _i, _p, _p]),
    "spectralUploadStateToSymbol": (_i, [_p]),
    "ptSpectralWhiteNorm": (_b, [_p, _p, _i, _p]),
    "ptSpectralCreate": (_p, [_p, _p, _i]),
    "ptSpectralLoadCoeffFile": (_p, [C.c_char_p]),
    "ptSpectralGetState": (_b, [_p, _p]),
    "ptSpectralDestroy": (None, [_p]),
    "pathTracerSetLightSampling": (_b, [_i]),
    "pathTracerLightSampling": (_i, []),
    "pathTracerSetSpectralSampling": (_b, [_i, _f, _f]),
    "pathTracerSpectralSampling": (_i, []),
    "pathTracerAdaptiveRenderPasses": (_b, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _u, _i, _p]),
    "pathTracerAdaptiveRenderPassesCounted": (_b, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _u, _i, _p,
                                                   _p]),
    "pathTracerAdaptiveRenderWavefront": (_b, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _u, _i, _i, _i,
                                               _p]),
    "pathTracerWavefrontProfile": (_b, [_i, _i]),
    "pathTracerWavefrontProfileRead": (_b, [_p, _i]),
    "pathTracerWavefrontProfileBounces": (_b, [_p, _i]),
    "pathTracerWavefrontSchedule": (_b, [_i]),
    "pathTracerTraceRays": (_b, [_p, _p, _p, _i, _f, _f, _p, _p]),
    "pathTracerCountTraversal": (_b, [_p, _p, _i, _i, _i, _p, _p, _p, _p, _u, _p, _p]),
    "ptSceneCreate": (_p, []),
    "ptSceneDestroy": (None, [_p]),
    "ptSceneBuild": (_b, [_p, _p, _u, _p, _u, _p, _u, _p]),
    "ptSceneDevice": (_p, [_p]),
    "ptSceneDescription": (_b, [_p, _p]),
    "ptSceneReadBvh": (_b, [_p, _p, _u, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "ptSceneLastBuildMs": (_f, [_p]),
    "ptSceneReadEmissive": (_b, [_p, _p, _p, _p, _u, C.POINTER(C.c_uint32)]),
    "ptSceneSetBuildMode": (_b, [_p, _i]),
    "ptSceneUpdateMaterials": (_b, [_p, _p, _u, _p]),
    "ptSceneSetSpectra": (_b, [_p, _p, _u, _p]),
    "pathTracerAdjointRender": (_b, [_p, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _p]),
    "pathTracerAdjointRenderSeeded": (_b, [_p, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _u, _p]),
    "pathTracerAdjointRenderSpectral": (_b, [_p, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _u, _p]),
    "ptBatchCreate": (_p, [_p, _i]),
    "ptBatchDestroy": (None, [_p]),
    "ptBatchInfo": (_b, [_p, _p, _p, _p, _p, _p]),
    "pathTracerRenderFixedBatch": (_b, [_p, _p, _p, _p, _p, _p, _i, _i, _u, _p]),
    "pathTracerAdjointRenderBatch": (_b, [_p, _p, _i, _u, _p, _p]),
    "pathTracerWavefrontCaptureRadiance": (_b, [_p, _i]),
    "pathTracerRenderFixedWavefront": (_b, [_p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p, _i, _i, _i, _p]),
    "pathTracerWavefrontReserve": (_b, [_i, _i, _i]),
    "pathTracerDetachedRender": (_b, [_p, _i, _i, _i, _p, _p, _p, _p, _p, _p]),
    "ptSceneReadLayout": (_b, [_p, _p, _p, _p]),
    "ptSceneReadWide": (_b, [_p, _p, _p, _p]),
    "ptBuildWorkspaceCreate": (_p, [_i]),
    "ptBuildWorkspaceDestroy": (None, [_p]),
    "ptSceneRebuild": (_b, [_p, _p, _p]),
    "ptSceneRebuildCheck": (_b, [_p, _p, _p]),
    "ptSceneRebuildCold": (_b, [_p, _p, _p]),
    "ptEnvironmentCreate": (_p, [_p, _i, _i, _p]),
    "ptEnvironmentDevice": (_p, [_p]),
    "ptEnvironmentReadTables": (_b, [_p, _p, _p]),
    "ptEnvironmentDestroy": (None, [_p]),
    # pt_tracer.h
    "ptTracerCreate": (_p, []),
    "ptTracerDestroy": (None, [_p]),
    "ptTracerSetFrameReadyCallback": (None, [_p, _p, _p]),
    "ptTracerConfigure": (_b, [_p, _i, _i, _p, _u, _p, _u, _p, _u]),
    "ptTracerRebuildMeshScene": (_b, [_p, _p, _u, _p, _u, _p, _u]),
    "ptTracerConfigureProcedural": (_b, [_p, _i, _i, _p, _u, _p]),
    "ptEnvironmentLoadHdr": (_p, [C.c_char_p, _p, _p]),
    "pathTracerWavefrontPins": (_i, []),
    "pathTracerWavefrontPin": (_b, []),
    "pathTracerWavefrontUnpin": (_b, [_i]),
    "ptEnvironmentFreeImage": (None, [_p]),
    "ptTracerRebuildMeshSceneProcedural": (_b, [_p, _p, _u, _p]),
    "ptTracerStart": (None, [_p]),
    "ptTracerStop": (None, [_p]),
    "ptTracerResetAccumulation": (None, [_p]),
    "ptTracerResetRegionAccumulation": (None, [_p]),
    "ptTracerIsRunning": (_b, [_p]),
    "ptTracerReleaseOutputSurfaces": (None, [_p]),
    "ptTracerBeginDisplayPublish": (_b, [_p, _i]),
    "ptTracerIsDisplayPublishReady": (_b, [_p]),
    "ptTracerHasDisplayPublishInFlight": (_b, [_p]),
    "ptTracerFinishDisplayPublish": (_i, [_p]),
    "ptTracerPublishDisplayFrame": (_b, [_p, _i]),
    "ptTracerReadDisplay": (_b, [_p, _i, _p]),
    "ptTracerSetMaxSamplesPerPixel": (None, [_p, _i]),
    "ptTracerMaxSamplesPerPixel": (_i, [_p]),
    "ptTracerSetMinSamples": (None, [_p, _i]),
    "ptTracerMinSamples": (_i, [_p]),
    "ptTracerSetRelativeErrorThreshold": (None, [_p, _f]),
    "ptTracerRelativeErrorThreshold": (_f, [_p]),
    "ptTracerSetDebugOverlayMode": (None, [_p, _i]),
    "ptTracerDebugOverlayMode": (_i, [_p]),
    "ptTracerSetBrdfDebugFlags": (None, [_p, _i]),
    "ptTracerBrdfDebugFlags": (_i, [_p]),
    "ptTracerSetPassesPerLaunch": (None, [_p, _i]),
    "ptTracerPassesPerLaunch": (_i, [_p]),
    "ptTracerLastWorkerError": (_b, [_p, _p, _i]),
    "ptTracerSetBvhBuildMode": (None, [_p, _i]),
    "ptTracerBvhBuildMode": (_i, [_p]),
    "ptTracerSetPreviewStepsPerLevel": (None, [_p, _i]),
    "ptTracerPreviewStepsPerLevel": (_i, [_p]),
    "ptTracerCurrentSampleCount": (_i, [_p]),
    "ptTracerCurrentActivePixelCount": (_i, [_p]),
    "ptTracerSampleBudgetTotalIterations": (_i, [_p]),
    "ptTracerIsSampleBudgetExhausted": (_b, [_p]),
    "ptTracerSetCamera": (None, [_p, _p]),
    "ptTracerLastSampleCamera": (None, [_p, _p]),
    "ptTracerSetClearColor": (None, [_p, _f, _f, _f]),
    "ptTracerSetPhysicalCamera": (None, [_p, _f, _f, _f]),
    "ptTracerSetFocusPoint": (None, [_p, Vec3C]),
    "ptTracerClearFocusPoint": (None, [_p]),
    "ptTracerPickSurface": (_b, [_p, Vec3C, Vec3C, _p]),
    "ptTracerSetEnvironmentImage": (_b, [_p, _p, _i, _i]),
    "ptTracerSetEnvironmentIntensity": (None, [_p, _f]),
    "ptTracerEnvironmentIntensity": (_f, [_p]),
    "ptTracerSetRussianRouletteMinDepth": (None, [_p, _i]),
    "ptTracerRussianRouletteMinDepth": (_i, [_p]),
    "ptTracerSetRegionRenderEnabled": (None, [_p, _b]),
    "ptTracerRegionRenderEnabled": (_b, [_p]),
    "ptTracerSetRenderRegion": (None, [_p, _i, _i, _i, _i]),
    "ptTracerRenderRegion": (None, [_p, _p, _p, _p, _p]),
    "ptTracerWaitUntilIdle": (_b, [_p, _i]),
    "ptTracerReadAccumulator": (_b, [_p, _p, _p, _p, _p, _p]),
    "ptTracerReadPreviewLevel": (_b, [_p, _i, _p, _p, _p, _p]),
    "ptTracerExposureMultiplier": (_f, [_p]),
    "ptTracerSetMaxPathDepth": (None, [_p, _i]),
    "ptTracerMaxPathDepth": (_i, [_p]),
}


def header_symbols() -> list[str]:
    """Function names declared with C linkage in include/pt_hip.h, include/pt_tracer.h and
    include/pt_procedural.h."""
    text = HEADER.read_text()
    tracer = TRACER_HEADER.read_text()
    text += tracer[: tracer.index("} /* extern \"C\" */")]
    text += PROCEDURAL_HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) - {
        "if", "defined", "alignas", "_Alignas", "sizeof", "PT_ALIGN16", "void"})


def lib():
    """The loaded library. Raises LibraryMissing when libpt_hip.so is absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m lsystemeditor_amd.build` (hipcc, gfx950)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class LightSampling:
    """Context manager for pathTracerSetLightSampling (0: the reference's light sampling, 1: MIS on
    area lights); restores the previous mode on exit."""

    def __init__(self, mode: int):
        self.mode = mode
        self.prev = None

    def __enter__(self):
        self.prev = lib().pathTracerLightSampling()
        check(lib().pathTracerSetLightSampling(self.mode), "pathTracerSetLightSampling")
        return self

    def __exit__(self, *exc):
        lib().pathTracerSetLightSampling(self.prev)


def load_hdr(path) -> np.ndarray:
    """ptEnvironmentLoadHdr: a Radiance .hdr picture as a (height, width, 3) float32 array."""
    L = lib()
    w, h = C.c_int(0), C.c_int(0)
    ptr = L.ptEnvironmentLoadHdr(str(path).encode(), C.byref(w), C.byref(h))
    if not ptr:
        raise RuntimeError(f"ptEnvironmentLoadHdr failed: {last_error()}")
    try:
        out = np.zeros((h.value, w.value, 3), np.float32)
        C.memmove(out.ctypes.data, ptr, out.nbytes)
        return out
    finally:
        L.ptEnvironmentFreeImage(ptr)


def last_error() -> str:
    return lib().pathTracerLastError().decode()


def check(ok: bool, what: str) -> None:
    if not ok:
        raise RuntimeError(f"{what} failed: {last_error()}")


# ---- record packers ------------------------------------------------------------------------

def pack_host_triangles(v0, v1, v2, n0=None, n1=None, n2=None, uv0=None, uv1=None, uv2=None,
                        material=None) -> np.ndarray:
    """(N, 25) float32 words in the HostTriangle layout (HostMesh.h:9-21)."""
    v0 = np.asarray(v0, dtype=np.float32)
    n = v0.shape[0]
    out = np.zeros((n, HOST_TRIANGLE_WORDS), dtype=np.float32)
    out[:, 0:3] = v0
    out[:, 3:6] = np.asarray(v1, dtype=np.float32)
    out[:, 6:9] = np.asarray(v2, dtype=np.float32)
    if n0 is None:
        e1 = out[:, 3:6].astype(np.float64) - out[:, 0:3]
        e2 = out[:, 6:9].astype(np.float64) - out[:, 0:3]
        nn = np.cross(e1, e2)
        nn /= np.maximum(np.linalg.norm(nn, axis=1, keepdims=True), 1e-30)
        n0 = n1 = n2 = nn.astype(np.float32)
    out[:, 9:12] = n0
    out[:, 12:15] = n1
    out[:, 15:18] = n2
    if uv0 is not None:
        out[:, 18:20] = uv0
        out[:, 20:22] = uv1
        out[:, 22:24] = uv2
    mats = np.zeros(n, dtype=np.uint32) if material is None else np.asarray(material, dtype=np.uint32)
    out[:, 24] = mats.view(np.float32) if mats.ndim else np.full(n, mats, dtype=np.uint32).view(np.float32)
    return out


def host_to_gpu_triangles(host: np.ndarray) -> np.ndarray:
    """(N, 28) float32 words in the reference TriangleGpu layout (MeshAccelTypes.h:90-102)."""
    out = np.zeros((host.shape[0], TRIANGLE_GPU_WORDS), dtype=np.float32)
    out[:, :HOST_TRIANGLE_WORDS] = host
    return out


def pack_material(r=0.8, g=0.8, b=0.8, roughness=0.5, metallic=0.0, transmission=0.0, thin=0.0, ior=1.5,
                  subsurface=0.0, emission=0.0, diffuseRoughness=-1.0, scatterRadius=(0.0, 0.0, 0.0),
                  specular=1.0, textures=(0, 0, 0, 0, 0, 0, 0, 0)) -> np.ndarray:
    """(24,) float32 words in the MaterialGpu layout (MeshAccelTypes.h:49-86)."""
    w = np.zeros(MATERIAL_WORDS, dtype=np.float32)
    w[:15] = [r, g, b, roughness, metallic, transmission, thin, ior, subsurface, emission, diffuseRoughness,
              scatterRadius[0], scatterRadius[1], scatterRadius[2], specular]
    w[15:23] = np.asarray(textures, dtype=np.uint32).view(np.float32)
    return w


def quat_look(forward, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Quaternion (w, x, y, z) rotating camera space (-Z forward, +Y up) onto `forward`."""
    f = np.asarray(forward, dtype=np.float64)
    f = f / np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    r = r / np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.stack([r, u, -f], axis=1)  # columns: camera x, y, z axes in world space
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
    elif m[1, 1] > m[2, 2]:
        s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
    else:
        s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
    return np.asarray(q, dtype=np.float64)


def pack_camera(position, orientation_wxyz=(1.0, 0.0, 0.0, 0.0), fov_y=0.7853981633974483, aspect=1.0, near=0.1,
                far=1.0e5, aperture_radius=0.0, focus_distance=0.0, focus_point=(0.0, 0.0, 0.0),
                focus_valid=0) -> np.ndarray:
    """(20,) float32 words in the CameraGpu layout (CameraGpu.h:5-16, 80 bytes)."""
    w = np.zeros(CAMERA_WORDS, dtype=np.float32)
    w[0:3] = position
    w[4:8] = orientation_wxyz
    w[8:14] = [fov_y, aspect, near, far, aperture_radius, focus_distance]
    w[14:17] = focus_point
    w[17:18] = np.array([focus_valid], dtype=np.int32).view(np.float32)
    return w


def pack_params(background=(10.0 / 255.0, 10.0 / 255.0, 10.0 / 255.0), environment_intensity=1.0,
                max_path_depth=32, russian_roulette_min_depth=3, min_samples=16, max_samples_per_pixel=1024,
                debug_overlay_mode=0, brdf_debug_flags=0, relative_error_threshold=0.02) -> np.ndarray:
    """(11,) float32 words in the RenderParamsGpu layout (RenderTypes.h:15-29); defaults as the reference."""
    w = np.zeros(PARAMS_WORDS, dtype=np.float32)
    w[0:4] = [background[0], background[1], background[2], environment_intensity]
    w[4:10] = np.array([max_path_depth, russian_roulette_min_depth, min_samples, max_samples_per_pixel,
                        debug_overlay_mode, brdf_debug_flags], dtype=np.int32).view(np.float32)
    w[10] = relative_error_threshold
    return w


class Scene:
    """A scene built on the GPU (ptSceneBuild): BVH, device layout, emissive list."""

    BUILD_MODES = {"median": 0, "sah": 1, "sah-binary": 2}

    def __init__(self, host_triangles: np.ndarray, materials: np.ndarray, textures: np.ndarray | None = None,
                 stream: int = 0, bvh: str = "median"):
        L = lib()
        self.handle = L.ptSceneCreate()
        if not self.handle:
            raise RuntimeError("ptSceneCreate failed")
        self.bvh = bvh
        check(L.ptSceneSetBuildMode(self.handle, self.BUILD_MODES[bvh]), "ptSceneSetBuildMode")
        tris = np.ascontiguousarray(host_triangles, dtype=np.float32)
        mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, MATERIAL_WORDS)
        tex = None if textures is None else np.ascontiguousarray(textures, dtype=np.float32).reshape(-1, TEXTURE_WORDS)
        self.triangle_count = tris.shape[0]
        ok = L.ptSceneBuild(self.handle, tris.ctypes.data if tris.size else None, tris.shape[0],
                            mats.ctypes.data if mats.size else None, mats.shape[0],
                            None if tex is None else tex.ctypes.data, 0 if tex is None else tex.shape[0], stream)
        if not ok:
            err = last_error()
            L.ptSceneDestroy(self.handle)
            self.handle = None
            raise RuntimeError(f"ptSceneBuild failed: {err}")

    @property
    def device(self) -> int:
        return lib().ptSceneDevice(self.handle)

    def set_spectra(self, spectra: np.ndarray | None, stream: int = 0) -> None:
        """Binned spectra for hero-wavelength rendering (ptSceneSetSpectra): (materials, 32)
        float32, 16 reflectance then 16 emission values per material (a negative first value of
        either half keeps the material's rgb2spec colour for it); None removes them."""
        if spectra is None:
            check(lib().ptSceneSetSpectra(self.handle, None, 0, stream), "ptSceneSetSpectra")
            self._spectra = None
            return
        self._spectra = np.ascontiguousarray(spectra, dtype=np.float32).reshape(-1, SPECTRUM_WORDS)
        check(lib().ptSceneSetSpectra(self.handle, self._spectra.ctypes.data, self._spectra.shape[0], stream),
              "ptSceneSetSpectra")

    @property
    def build_ms(self) -> float:
        return lib().ptSceneLastBuildMs(self.handle)

    def rebuild(self, workspace: "BuildWorkspace", stream: int = 0, cold: bool = False) -> None:
        """Enqueues the scene's binned-SAH build again on its device triangles (capturable): the
        levels its first build needed, or with cold every level the depth cap allows."""
        fn = lib().ptSceneRebuildCold if cold else lib().ptSceneRebuild
        check(fn(self.handle, workspace.handle, stream), "ptSceneRebuildCold" if cold else "ptSceneRebuild")

    def rebuild_check(self, workspace: "BuildWorkspace", stream: int = 0) -> None:
        """Waits for the stream; raises if the last rebuild with `workspace` did not reproduce the tree."""
        check(lib().ptSceneRebuildCheck(self.handle, workspace.handle, stream), "ptSceneRebuildCheck")

    def read_layout(self) -> tuple[np.ndarray, np.ndarray, int]:
        """Device layout: internal nodes (n-1, 16) float32, leaf slots (n, 12) float32, SAH levels."""
        n = self.triangle_count
        nodes = np.zeros((max(n - 1, 0), 16), np.float32)
        slots = np.zeros((n, 12), np.float32)
        levels = C.c_int(0)
        check(lib().ptSceneReadLayout(self.handle, nodes.ctypes.data if nodes.size else None,
                                      slots.ctypes.data if slots.size else None, C.byref(levels)), "ptSceneReadLayout")
        return nodes, slots, levels.value

    def read_wide(self) -> tuple[np.ndarray, int]:
        """4-wide tree (build mode "sah"): nodes (count, 32) float32 and the root reference."""
        count, root = C.c_int(0), C.c_int(0)
        check(lib().ptSceneReadWide(self.handle, None, C.byref(count), C.byref(root)), "ptSceneReadWide")
        nodes = np.zeros((count.value, 32), np.float32)
        if count.value:
            check(lib().ptSceneReadWide(self.handle, nodes.ctypes.data, C.byref(count), C.byref(root)),
                  "ptSceneReadWide")
        return nodes, root.value

    def description(self) -> bytes:
        buf = (C.c_uint8 * SCENE_DESC_BYTES)()
        check(lib().ptSceneDescription(self.handle, C.addressof(buf)), "ptSceneDescription")
        return bytes(buf)

    def read_emissive(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Device emissive list: original triangle indices, prefix CDF (count + 1) and the packed
        (count, 12) float32 records the sampler reads."""
        cnt = C.c_uint32(0)
        check(lib().ptSceneReadEmissive(self.handle, None, None, None, 0, C.byref(cnt)), "ptSceneReadEmissive")
        n = cnt.value
        idx = np.zeros(n, np.uint32)
        cdf = np.zeros(n + 1, np.float32)
        rec = np.zeros((n, 12), np.float32)
        if n:
            check(lib().ptSceneReadEmissive(self.handle, idx.ctypes.data, cdf.ctypes.data, rec.ctypes.data, n,
                                            C.byref(cnt)), "ptSceneReadEmissive")
        else:
            check(lib().ptSceneReadEmissive(self.handle, None, cdf.ctypes.data, None, 1, C.byref(cnt)),
                  "ptSceneReadEmissive")
        return idx, cdf, rec

    def read_bvh(self) -> tuple[np.ndarray, int]:
        """Nodes in the reference's 48-byte pre-order layout, as (count, 12) float32 words."""
        L = lib()
        cnt = C.c_uint32(0)
        root = C.c_uint32(0)
        check(L.ptSceneReadBvh(self.handle, None, 0, C.byref(cnt), C.byref(root)), "ptSceneReadBvh")
        nodes = np.zeros((cnt.value, BVH_NODE_WORDS), dtype=np.float32)
        if cnt.value:
            check(L.ptSceneReadBvh(self.handle, nodes.ctypes.data, cnt.value, C.byref(cnt), C.byref(root)),
                  "ptSceneReadBvh")
        return nodes, root.value

    def close(self):
        if self.handle:
            lib().ptSceneDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BuildWorkspace:
    """Temporaries of the binned-SAH build for scenes of up to max_triangles triangles
    (ptBuildWorkspaceCreate); rebuilds that share one must be ordered on one stream."""

    def __init__(self, max_triangles: int):
        self.max_triangles = int(max_triangles)
        self.handle = lib().ptBuildWorkspaceCreate(self.max_triangles)
        if not self.handle:
            raise RuntimeError(f"ptBuildWorkspaceCreate failed: {last_error()}")

    def close(self) -> None:
        if self.handle:
            lib().ptBuildWorkspaceDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Environment:
    """A lat-long HDR environment with importance tables on the GPU (ptEnvironmentCreate)."""

    def __init__(self, rgb: np.ndarray, stream: int = 0):
        img = np.ascontiguousarray(rgb, dtype=np.float32)
        if img.ndim != 3 or img.shape[2] != 3:
            raise ValueError("environment image must be (height, width, 3) float32")
        self.height, self.width = img.shape[0], img.shape[1]
        self.handle = lib().ptEnvironmentCreate(img.ctypes.data, self.width, self.height, stream)
        if not self.handle:
            raise RuntimeError(f"ptEnvironmentCreate failed: {last_error()}")

    @property
    def device(self) -> int:
        return lib().ptEnvironmentDevice(self.handle)

    def tables(self) -> tuple[np.ndarray, np.ndarray]:
        marg = np.zeros(self.height + 1, dtype=np.float32)
        rows = np.zeros(self.height * (self.width + 1), dtype=np.float32)
        check(lib().ptEnvironmentReadTables(self.handle, marg.ctypes.data, rows.ctypes.data), "ptEnvironmentReadTables")
        return marg, rows

    def close(self):
        if self.handle:
            lib().ptEnvironmentDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PtBatchSceneGpu(C.Structure):
    """One scene of a batch (ptBatchCreate): device pointers and image size."""
    _fields_ = [("scene", C.c_void_p), ("camera", C.c_void_p), ("params", C.c_void_p), ("env", C.c_void_p),
                ("width", C.c_int), ("height", C.c_int)]


class SpectralStateGpu(C.Structure):
    """Spectral/SpectralState.h:11-19 (device table pointers and white normalisation)."""
    _fields_ = [("rgb2specScale", C.c_void_p), ("rgb2specData", C.c_void_p), ("rgb2specRes", C.c_int),
                ("whiteNormR", C.c_float), ("whiteNormG", C.c_float), ("whiteNormB", C.c_float)]


class Spectral:
    """An rgb2spec coefficient table on the GPU (ptSpectralCreate), the role of SpectralStateHost."""

    def __init__(self, scale: np.ndarray, data: np.ndarray):
        self.scale = np.ascontiguousarray(scale, dtype=np.float32)
        self.data = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        self.res = int(self.scale.shape[0])
        if self.data.size != 9 * self.res ** 3:
            raise ValueError("data must hold 3 * 3 * res^3 coefficients")
        self.handle = lib().ptSpectralCreate(self.scale.ctypes.data, self.data.ctypes.data, self.res)
        if not self.handle:
            raise RuntimeError(f"ptSpectralCreate failed: {last_error()}")

    def state(self) -> SpectralStateGpu:
        st = SpectralStateGpu()
        check(lib().ptSpectralGetState(self.handle, C.byref(st)), "ptSpectralGetState")
        return st

    def white_norm(self) -> tuple[float, float, float]:
        st = self.state()
        return st.whiteNormR, st.whiteNormG, st.whiteNormB

    def enable(self, lambda_min: float = 360.0, lambda_max: float = 830.0) -> None:
        """Stores this table (spectralUploadStateToSymbol) and turns hero-wavelength sampling on."""
        st = self.state()
        if lib().spectralUploadStateToSymbol(C.byref(st)) != 0:
            raise RuntimeError("spectralUploadStateToSymbol failed")
        check(lib().pathTracerSetSpectralSampling(4, lambda_min, lambda_max), "pathTracerSetSpectralSampling")

    @staticmethod
    def disable() -> None:
        check(lib().pathTracerSetSpectralSampling(0, 0.0, 0.0), "pathTracerSetSpectralSampling")

    def close(self):
        if self.handle:
            lib().ptSpectralDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

