"""ctypes binding of the oracle (CPU restatement of PathTracerLib's hot path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the shipped package. Builds liboracle.so with `make` on
first use when it is missing.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
# The same sources with FP contraction allowed (FMA, AVX2 target): a CPU-baseline timing
# variant only; its results are not bit-equal to the reference order of operations.
LIB_CONTRACT_PATH = HERE / "liboracle_contract.so"
REF_PATH = HERE / "_ref" / "libref_policy.so"
REF_PROC_PATH = HERE / "_ref" / "libref_procedural.so"


def build(force: bool = False, path: Path = LIB_PATH) -> Path:
    if force or not path.exists():
        subprocess.run(["make", "-C", str(HERE), path.name], check=True, capture_output=True)
    return path


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def np(self):
        return np.array([self.x, self.y, self.z], dtype=np.float32)


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Material(C.Structure):
    _pack_ = 16
    _fields_ = [(n, C.c_float) for n in (
        "r", "g", "b", "roughness", "metallic", "transmission", "thin", "ior", "subsurface", "emission",
        "diffuseRoughness", "scatterRadiusR", "scatterRadiusG", "scatterRadiusB", "specular")] + [
        (n, C.c_uint32) for n in ("albedoTex", "roughnessTex", "metallicTex", "transmissionTex", "thinTex",
                                  "iorTex", "subsurfaceTex", "emissionTex")] + [("_pad", C.c_uint32)]


class BrdfCtx(C.Structure):
    _fields_ = [("normal", Vec3), ("wo", Vec3), ("_pad0", C.c_uint32 * 2), ("material", Material),
                ("etaMedium", C.c_float), ("wavelengthNm", C.c_float), ("debugFlags", C.c_int),
                ("_pad1", C.c_uint32)]


class Sample(C.Structure):
    _fields_ = [("direction", Vec3), ("pdf", C.c_float), ("valid", C.c_int), ("transmitted", C.c_int),
                ("nextMediumEta", C.c_float), ("subsurfaceScatter", C.c_int), ("subsurfaceExitOffset", Vec3),
                ("subsurfaceInternalSteps", C.c_int)]


class LobeWeights(C.Structure):
    _fields_ = [("diffuse", C.c_float), ("specular", C.c_float), ("transmit", C.c_float),
                ("subsurface", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("hit", C.c_int), ("t", C.c_float), ("normal", Vec3), ("uv", Vec2), ("triangleIndex", C.c_uint32)]


class Scene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("triangles", C.c_void_p), ("materials", C.c_void_p),
                ("emissiveIndices", C.c_void_p), ("emissiveCdf", C.c_void_p), ("textures", C.c_void_p),
                ("nodeCount", C.c_uint32), ("triangleCount", C.c_uint32), ("rootIndex", C.c_uint32),
                ("materialCount", C.c_uint32), ("emissiveCount", C.c_uint32), ("textureCount", C.c_uint32),
                ("spectra", C.c_void_p)]


class Env(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("marginalCdf", C.c_void_p), ("rowCdf", C.c_void_p),
                ("width", C.c_int), ("height", C.c_int), ("valid", C.c_int)]


class Rng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("ctr", C.c_uint32 * 4), ("buf", C.c_uint32 * 4), ("avail", C.c_int)]


class TraceStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "closestRays", "shadowRays", "nodeFetches", "triTests", "hits")]


assert C.sizeof(Material) == 96
assert C.sizeof(Scene) == 80

_libs: dict = {}


def lib(variant: str = "parity"):
    """The oracle library: "parity" (FP contraction off, the checker) or "contract" (contraction
    allowed; used only to time the CPU baseline)."""
    if variant not in _libs:
        path = LIB_PATH if variant == "parity" else LIB_CONTRACT_PATH
        build(path=path)
        L = C.CDLL(str(path))
        f = C.c_float
        fp = C.POINTER(C.c_float)
        vp = C.c_void_p
        sig = {
            "orc_sinf": (f, [f]), "orc_cosf": (f, [f]), "orc_tanf": (f, [f]), "orc_expf": (f, [f]),
            "orc_logf": (f, [f]), "orc_atan2f": (f, [f, f]), "orc_acosf": (f, [f]),
            "orc_vec_normalize": (Vec3, [Vec3]),
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_seed_hash": (C.c_uint32, [C.c_int, C.c_int, C.c_uint32]),
            "orc_u32_to_unit": (f, [C.c_uint32]),
            "orc_bvh_build": (C.c_int, [vp, C.c_uint32, vp, C.POINTER(C.c_uint32)]),
            "orc_ray_triangle": (C.c_int, [Vec3, Vec3, vp, f, f, fp, C.POINTER(Vec3), C.POINTER(Vec2)]),
            "orc_trace_ray": (Hit, [Vec3, Vec3, C.POINTER(Scene), f, f]),
            "orc_trace_ray_counted": (Hit, [Vec3, Vec3, C.POINTER(Scene), f, f, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint64)]),
            "orc_build_emissive": (C.c_uint32, [vp, C.c_uint32, vp, C.c_uint32, vp, vp]),
            "orc_lobe_weights": (LobeWeights, [C.POINTER(BrdfCtx)]),
            "orc_brdf_eval": (Vec3, [C.POINTER(BrdfCtx), Vec3]),
            "orc_brdf_pdf": (f, [C.POINTER(BrdfCtx), Vec3]),
            "orc_brdf_sample": (Sample, [C.POINTER(BrdfCtx), f, f]),
            "orc_brdf_sample_transmit": (Sample, [C.POINTER(BrdfCtx), f, f]),
            "orc_oren_nayar_ctx": (f, [C.POINTER(BrdfCtx), Vec3]),
            "orc_refract_direction": (C.c_int, [C.POINTER(BrdfCtx), C.POINTER(Vec3), fp, fp, fp]),
            "orc_skips_env_nee": (C.c_int, [C.POINTER(Material)]),
            "orc_apply_throughput": (Vec3, [Vec3, C.POINTER(BrdfCtx), C.POINTER(Sample), Vec3]),
            "orc_ggx_d": (f, [f, f]), "orc_smith_g1": (f, [f, f]),
            "orc_dielectric_fresnel": (f, [f, f, f]), "orc_schlick_f": (Vec3, [f, Vec3]),
            "orc_energy_compensation": (f, [f, f]), "orc_oren_nayar_factor": (f, [f, f, f, f]),
            "orc_sample_cosine_local": (Vec3, [f, f]),
            "orc_sample_ggx_half_local": (Vec3, [f, f, f, fp]),
            "orc_mis_balance": (f, [f, f]),
            "orc_lat_long_to_dir": (Vec3, [f, f]), "orc_dir_to_lat_long": (Vec3, [Vec3]),
            "orc_env_eval": (Vec3, [C.POINTER(Env), vp, Vec3]),
            "orc_env_pdf": (f, [C.POINTER(Env), vp, Vec3]),
            "orc_env_sample": (Vec3, [C.POINTER(Env), vp, f, f, fp]),
            "orc_env_build_tables": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
            "orc_light_sample_emissive": (C.c_int, [C.POINTER(Scene), f, f, f, C.POINTER(Vec3), C.POINTER(Vec3),
                                                    C.POINTER(Vec3), fp]),
            "orc_camera_ray": (None, [vp, f, f, f, f, C.POINTER(Vec3), C.POINTER(Vec3)]),
            "orc_clamp_firefly": (Vec3, [Vec3]),
            "orc_pixel_radiance": (Vec3, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                          C.POINTER(Scene), C.POINTER(Env), vp, C.c_uint32]),
            "orc_stats_begin": (None, [C.POINTER(TraceStats)]), "orc_stats_end": (None, []),
            "orc_is_converged": (C.c_int, [C.c_uint32, f, f, C.c_int, f]),
            "orc_adaptive_sample": (None, [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, C.POINTER(Scene),
                                           C.POINTER(Env), vp, C.c_uint32, C.c_int]),
            "orc_preview_sample": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), C.POINTER(Env), vp,
                                          C.c_uint32, C.c_int]),
            "orc_compact_active": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int]),
            "orc_clear_accumulator": (None, [vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int]),
            "orc_render_adaptive": (C.c_int, [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp, C.POINTER(Scene),
                                              C.POINTER(Env), vp, C.c_int, C.c_int, C.c_int]),
            "orc_copy_to_pbo": (None, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, f]),
            "orc_uv_debug": (None, [vp, C.POINTER(Scene), vp, C.c_int, C.c_int, vp]),
            "orc_trace_path_ad": (Vec3, [Vec3, Vec3, C.POINTER(Scene), vp, vp, C.POINTER(Env), vp, vp, Vec3, vp]),
            "orc_trace_path": (Vec3, [Vec3, Vec3, C.POINTER(Scene), vp, C.POINTER(Env), vp]),
            "orc_brdf_grad": (None, [vp, Vec3, vp, vp, vp]),
            "orc_clamp_vjp": (None, [Vec3, vp, vp]),
            "orc_brdf_eval_f64": (None, [vp, Vec3, vp, vp]),
            "orc_render_detached_spectral": (None, [vp, vp, vp, C.c_int, C.c_int, vp, C.POINTER(Scene), vp, vp,
                                                    C.POINTER(Env), vp, vp, C.c_int, C.c_uint32, C.c_int]),
            "orc_render_adjoint_spectral": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), C.POINTER(Env), vp,
                                                   vp, C.c_int, C.c_uint32, vp, C.c_int]),
            "orc_render_detached_f64_seeded": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), vp, C.POINTER(Env),
                                                      vp, C.c_int, C.c_uint32, vp, C.c_int]),
            "orc_render_detached": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), vp, C.POINTER(Env), vp, C.c_int,
                                           C.c_int]),
            "orc_render_adjoint": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), C.POINTER(Env), vp, C.c_int, vp,
                                          C.c_int]),
            "orc_render_detached_seeded": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), vp, C.POINTER(Env), vp,
                                                  C.c_int, C.c_uint32, C.c_int]),
            "orc_render_adjoint_seeded": (None, [vp, C.c_int, C.c_int, vp, C.POINTER(Scene), C.POINTER(Env), vp,
                                                 C.c_int, C.c_uint32, vp, C.c_int]),
            "orc_clear_region": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_rebuild_active": (C.c_int, [vp, vp, vp, C.c_int, C.c_int]),
            "orc_cie_interp": (f, [C.c_int, f]),
            "orc_cie_rgb_response": (None, [f, vp]),
            "orc_rgb2spec_coeffs": (None, [C.POINTER(SpecModel), f, f, f, vp]),
            "orc_rgb2spec_sigmoid": (f, [vp, f]),
            "orc_rgb2spec_reflectance": (f, [C.POINTER(SpecModel), f, f, f, f]),
            "orc_rgb2spec_radiance": (f, [C.POINTER(SpecModel), f, f, f, f]),
            "orc_spectral_white_norm": (None, [C.POINTER(SpecModel), vp]),
            "orc_spectral_to_rgb": (None, [C.POINTER(SpecModel), f, f, f, vp]),
            "orc_hero_lambdas": (None, [C.POINTER(SpecModel), f, vp]),
            "orc_hero_to_rgb": (Vec3, [C.POINTER(SpecModel), vp, vp]),
            "orc_clamp_firefly_spectral": (Vec3, [Vec3]),
            "orc_trace_path_spectral": (None, [Vec3, Vec3, C.POINTER(Scene), vp, C.POINTER(Env), vp,
                                               C.POINTER(SpecModel), vp, vp]),
            "orc_set_spectral": (None, [C.POINTER(SpecModel)]),
            "orc_set_light_sampling": (None, [C.c_int]),
            "orc_emissive_pdf_hit": (f, [C.POINTER(Scene), C.c_uint32, Vec3, f]),
            "orc_loft_triangles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
            "orc_root_transform": (None, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
            "orc_rng_init": (None, [C.POINTER(Rng), C.c_int, C.c_int, C.c_uint32]),
            "orc_burley_sample_radius": (f, [f, f]), "orc_burley_profile": (f, [f, f]),
            "orc_burley_pdf_radius": (f, [f, f]),
            "orc_subsurface_effective_radius": (f, [C.POINTER(Material), f]),
            "orc_resolve_material": (Material, [C.POINTER(Material), Vec2, C.POINTER(Scene)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[variant] = L
    return _libs[variant]


def ref_policy():
    """The reference's own header-only policies compiled from its sources, or None."""
    if not REF_PATH.exists():
        return None
    L = C.CDLL(str(REF_PATH))
    L.ref_is_converged.restype = C.c_int
    L.ref_is_converged.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_int, C.c_float]
    L.ref_sample_luminance.restype = C.c_float
    L.ref_sample_luminance.argtypes = [C.c_float, C.c_float, C.c_float]
    for n in ("ref_budget_total", "ref_can_take_sample", "ref_budget_exhausted", "ref_adaptive_exhausted"):
        getattr(L, n).restype = C.c_int
    L.ref_budget_total.argtypes = [C.c_int, C.c_int]
    L.ref_can_take_sample.argtypes = [C.c_int, C.c_int, C.c_int]
    L.ref_budget_exhausted.argtypes = [C.c_int, C.c_int, C.c_int]
    L.ref_adaptive_exhausted.argtypes = [C.c_int, C.c_int, C.c_int]
    L.ref_preview_level.restype = None
    L.ref_preview_level.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ref_mis_balance.restype = C.c_float
    L.ref_mis_balance.argtypes = [C.c_float, C.c_float]
    L.ref_normalize.restype = None
    L.ref_normalize.argtypes = [C.c_void_p, C.c_void_p]
    L.ref_rotate_ypr.restype = None
    L.ref_rotate_ypr.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]
    L.ref_rgb2spec_coeffs.restype = None
    L.ref_rgb2spec_coeffs.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    for n in ("ref_rgb2spec_reflectance", "ref_rgb2spec_radiance"):
        getattr(L, n).restype = C.c_float
        getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_int] + [C.c_float] * 4
    return L


def loft_triangles(frames: np.ndarray, circular_segments: int) -> np.ndarray:
    """orc_loft_triangles: (K, 13) frames -> (T, 25) HostTriangle words (material word 0)."""
    f = np.ascontiguousarray(frames, np.float32).reshape(-1, 13)
    L = lib()
    n = L.orc_loft_triangles(f.ctypes.data, f.shape[0], circular_segments, None)
    out = np.zeros((n, 25), np.float32)
    if n:
        L.orc_loft_triangles(f.ctypes.data, f.shape[0], circular_segments, out.ctypes.data)
    return out


def root_transform(tris: np.ndarray, translation, rotation_deg) -> np.ndarray:
    """orc_root_transform on a copy of (T, 25) HostTriangle words."""
    out = np.ascontiguousarray(tris, np.float32).copy()
    t = np.asarray(translation, np.float32)
    r = np.asarray(rotation_deg, np.float32)
    lib().orc_root_transform(out.ctypes.data, out.shape[0], t.ctypes.data, r.ctypes.data)
    return out


def ref_procedural():
    """The reference's L-system library, turtle and spline compiled from its sources, or None."""
    if not REF_PROC_PATH.exists():
        return None
    L = C.CDLL(str(REF_PROC_PATH))
    p, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    L.ref_last_error.restype = C.c_char_p
    L.ref_last_error.argtypes = []
    L.ref_lsystem_parse_dump.restype = C.c_int64
    L.ref_lsystem_parse_dump.argtypes = [C.c_int, C.c_char_p, p, u64]
    L.ref_lsystem_evaluate_text.restype = C.c_int64
    L.ref_lsystem_evaluate_text.argtypes = [C.c_char_p, u64, u64, C.c_int, p, u64]
    L.ref_lsystem_materials_text.restype = C.c_int64
    L.ref_lsystem_materials_text.argtypes = [C.c_char_p, p, u64]
    L.ref_turtle_execute.restype = C.c_int
    L.ref_turtle_execute.argtypes = [C.c_char_p, u64, C.c_float, C.c_float, p, p, u32, p, u32, p, u32, p, u32]
    L.ref_spline_frames.restype = C.c_int
    L.ref_spline_frames.argtypes = [p, u32, C.c_float, C.c_int, C.c_int, p, u32, p, p]
    return L


# ---- helpers ---------------------------------------------------------------------------------

def v3(x, y=None, z=None) -> Vec3:
    if y is None:
        x, y, z = (float(c) for c in x)
    return Vec3(float(x), float(y), float(z))


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def material(r=0.8, g=0.8, b=0.8, roughness=0.5, metallic=0.0, transmission=0.0, thin=0.0, ior=1.5,
             subsurface=0.0, emission=0.0, diffuseRoughness=-1.0, scatterRadiusR=0.0, scatterRadiusG=0.0,
             scatterRadiusB=0.0, specular=1.0) -> Material:
    m = Material()
    for k, v in dict(r=r, g=g, b=b, roughness=roughness, metallic=metallic, transmission=transmission, thin=thin,
                     ior=ior, subsurface=subsurface, emission=emission, diffuseRoughness=diffuseRoughness,
                     scatterRadiusR=scatterRadiusR, scatterRadiusG=scatterRadiusG, scatterRadiusB=scatterRadiusB,
                     specular=specular).items():
        setattr(m, k, float(v))
    return m


def ctx(mat: Material, normal=(0.0, 0.0, 1.0), wo=None, eta=1.0, flags=0) -> BrdfCtx:
    c = BrdfCtx()
    c.normal = v3(normal)
    c.wo = wo if isinstance(wo, Vec3) else v3(wo)
    c.material = mat
    c.etaMedium = float(eta)
    c.wavelengthNm = 550.0
    c.debugFlags = int(flags)
    return c


class OracleScene:
    """Holds the arrays of a scene in the reference layout and the OScene view of them."""

    def __init__(self, tris112: np.ndarray, mats: np.ndarray, textures: np.ndarray | None = None):
        L = lib()
        self.tris = np.ascontiguousarray(tris112)  # (N, 28) float32 words, 112-byte records
        self.mats = np.ascontiguousarray(mats)  # (M, 24) float32 words, 96-byte records
        n = self.tris.shape[0]
        self.nodes = np.zeros((max(2 * n - 1, 1), 12), dtype=np.float32)
        root = C.c_uint32(0)
        self.node_count = L.orc_bvh_build(ptr(self.tris), n, ptr(self.nodes), C.byref(root)) if n else 0
        self.root = root.value
        self.em_idx = np.zeros(max(n, 1), dtype=np.uint32)
        self.em_cdf = np.zeros(max(n, 1) + 1, dtype=np.float32)
        self.em_count = L.orc_build_emissive(ptr(self.tris), n, ptr(self.mats), self.mats.shape[0],
                                             ptr(self.em_idx), ptr(self.em_cdf))
        self.textures = None if textures is None else np.ascontiguousarray(textures)
        s = Scene()
        s.nodes = ptr(self.nodes)
        s.triangles = ptr(self.tris)
        s.materials = ptr(self.mats)
        s.emissiveIndices = ptr(self.em_idx)
        s.emissiveCdf = ptr(self.em_cdf)
        s.textures = ptr(self.textures) if self.textures is not None else None
        s.nodeCount = self.node_count
        s.triangleCount = n
        s.rootIndex = self.root
        s.materialCount = self.mats.shape[0]
        s.emissiveCount = self.em_count
        s.textureCount = 0 if self.textures is None else self.textures.shape[0]
        self.scene = s
        self.spectra = None

    def set_spectra(self, spectra: np.ndarray | None) -> None:
        """(materials, 32) float32 binned spectra for the hero-wavelength integrator, or None."""
        if spectra is None:
            self.spectra = None
            self.scene.spectra = None
            return
        self.spectra = np.ascontiguousarray(spectra, dtype=np.float32).reshape(self.mats.shape[0], 32)
        self.scene.spectra = ptr(self.spectra)

    def trace(self, ro, rd, tmin=0.0, tmax=1.0e6) -> Hit:
        return lib().orc_trace_ray(v3(ro), v3(rd), C.byref(self.scene), tmin, tmax)


class SpecModel(C.Structure):
    _fields_ = [("scale", C.c_void_p), ("data", C.c_void_p), ("res", C.c_int), ("white", C.c_float * 3),
                ("lambdaMin", C.c_float), ("lambdaRange", C.c_float)]


class OracleSpectral:
    """An rgb2spec table as the oracle's OSpecModel (white normalisation computed by the oracle,
    SpectralState.cpp:15-46), with the hero-wavelength range. Keeps the arrays alive."""

    def __init__(self, scale: np.ndarray, data: np.ndarray, lambda_min: float = 360.0, lambda_max: float = 830.0):
        self.scale = np.ascontiguousarray(scale, dtype=np.float32)
        self.data = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        rng = float(np.float32(np.float32(lambda_max) - np.float32(lambda_min)))
        self.model = SpecModel(ptr(self.scale), ptr(self.data), int(self.scale.shape[0]), (C.c_float * 3)(1, 1, 1),
                               float(lambda_min), rng)
        w = np.zeros(3, np.float32)
        lib().orc_spectral_white_norm(C.byref(self.model), ptr(w))
        self.model.white = (C.c_float * 3)(*w.tolist())

    def __enter__(self):
        lib().orc_set_spectral(C.byref(self.model))
        return self

    def __exit__(self, *exc):
        lib().orc_set_spectral(None)


class OracleEnv:
    """A lat-long RGB environment with its importance tables (EnvironmentMap.cpp:76-127), as the
    OEnv view the integrator reads. Keeps the arrays alive."""

    def __init__(self, rgb: np.ndarray):
        self.pixels = np.ascontiguousarray(rgb, dtype=np.float32)
        self.height, self.width = self.pixels.shape[0], self.pixels.shape[1]
        self.marginal = np.zeros(self.height + 1, np.float32)
        self.rows = np.zeros(self.height * (self.width + 1), np.float32)
        if not lib().orc_env_build_tables(ptr(self.pixels), self.width, self.height, ptr(self.marginal),
                                          ptr(self.rows)):
            raise ValueError("orc_env_build_tables rejected the image")
        self.env = Env(ptr(self.pixels), ptr(self.marginal), ptr(self.rows), self.width, self.height, 1)


class Accumulator:
    """Per-pixel adaptive state in host memory (PathTracer.cpp:323-419 buffers)."""

    def __init__(self, width: int, height: int):
        n = width * height
        self.width, self.height = width, height
        self.acc = np.zeros((n, 4), dtype=np.float32)
        self.counts = np.zeros(n, dtype=np.uint32)
        self.mean = np.zeros(n, dtype=np.float32)
        self.m2 = np.zeros(n, dtype=np.float32)
        self.conv = np.zeros(n, dtype=np.uint8)
        self.active = np.zeros(n, dtype=np.int32)
        self.active_count = C.c_int(0)

    def clear(self):
        lib().orc_clear_accumulator(ptr(self.acc), ptr(self.counts), ptr(self.mean), ptr(self.m2), ptr(self.conv),
                                    ptr(self.active), C.addressof(self.active_count), self.width, self.height)


def render_adaptive(scene: OracleScene, camera: np.ndarray, params: np.ndarray, width: int, height: int,
                    first_iter: int = 0, max_passes: int = 1 << 30, threads: int = 1, env: Env | None = None):
    """Runs the reference render-loop schedule on the CPU; returns (Accumulator, passes)."""
    acc = Accumulator(width, height)
    env = env or Env()
    passes = lib().orc_render_adaptive(ptr(acc.acc), ptr(acc.counts), ptr(acc.mean), ptr(acc.m2), ptr(acc.conv),
                                       width, height, ptr(camera), C.byref(scene.scene), C.byref(env), ptr(params),
                                       first_iter, max_passes, threads)
    return acc, passes
