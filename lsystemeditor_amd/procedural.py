"""Procedural scene input through the C ABI of include/pt_procedural.h: L-system definitions ->
HostMesh -> scene (the reference's ProceduralMeshBuilder::buildHostMesh and meshSceneBuild,
Application/src/Procedural/ProceduralMeshBuilder.cpp:266-330, PathTracerLib/src/MeshAccel/
MeshSceneContent.cpp:5-52), plus the inspection hooks (parse structure, generations, turtle states,
spline frames) that the parity tests compare with the reference's own library.

All of it runs in C++ inside libpt_hip.so on the host; nothing here needs a GPU except
mesh_scene_build, which uploads into a Scene.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi

TURN_SYMBOLS = 1
LEAVES = 2
DEFAULT_RULE_SEED = 0xC0FFEE


class LSystemError(ValueError):
    """A definition that does not parse (the parser's message names line and column)."""


class BuildParams(C.Structure):
    """PtProceduralBuildParams: ProceduralBuildParams (ProceduralTypes.h:67-76) + turtle extensions."""
    _fields_ = [("circularSegments", C.c_int), ("samplesPerSpan", C.c_int), ("hermiteTension", C.c_float),
                ("segmentRefineTolerance", C.c_float), ("globalRefineTolerance", C.c_float),
                ("creaseAngleDeg", C.c_float), ("defaultStepLength", C.c_float), ("defaultRadius", C.c_float),
                ("turtleFlags", C.c_uint32), ("turnAngleDeg", C.c_float), ("rollAngleDeg", C.c_float),
                ("radiusDecay", C.c_float), ("jitterDeg", C.c_float), ("jitterSeed", C.c_uint64),
                ("leafLength", C.c_float), ("leafWidth", C.c_float), ("ruleSeed", C.c_uint64)]


class Instance(C.Structure):
    """PtProceduralInstance (ProceduralTypes.h:80-86)."""
    _fields_ = [("commandString", C.c_char_p), ("iterations", C.c_uint64), ("translation", abi.Vec3C),
                ("rotationDeg", abi.Vec3C)]


class MeshSceneParams(C.Structure):
    """PtMeshSceneBuildParams (MeshSceneContent.h:9-13)."""
    _fields_ = [("circularSegments", C.c_int), ("creaseAngleDeg", C.c_float)]


class TurtleCounts(C.Structure):
    _fields_ = [("segments", C.c_uint32), ("states", C.c_uint32), ("groups", C.c_uint32), ("leaves", C.c_uint32),
                ("materialIdChars", C.c_uint32)]


_p = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_SIGS = {
    "ptProceduralDefaultParams": (None, [_p]),
    "ptHostMeshDestroy": (None, [_p]),
    "ptHostMeshCounts": (None, [_p, _p, _p, _p]),
    "ptHostMeshTriangles": (_p, [_p]),
    "ptHostMeshMaterials": (_p, [_p]),
    "ptHostMeshTextures": (_p, [_p]),
    "ptProceduralBuildHostMesh": (_p, [C.c_char_p, _u64, abi.Vec3C, abi.Vec3C, _p]),
    "ptProceduralBuildScene": (_p, [_p, _u32, _p]),
    "ptMeshSceneBuild": (C.c_bool, [_p, _u32, _p, _p, _p]),
    "ptLSystemEvaluateText": (C.c_int64, [C.c_char_p, _u64, _u64, _u32, C.c_int, _p, _u64]),
    "ptLSystemParseDump": (C.c_int64, [C.c_int, C.c_char_p, _u32, _p, _u64]),
    "ptLSystemMaterialsText": (C.c_int64, [C.c_char_p, _p, _u64]),
    "ptTurtleExecute": (C.c_bool, [C.c_char_p, _u64, _p, _p, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _p, _p, _u32]),
    "ptSplineFrames": (C.c_bool, [_p, _u32, C.c_float, C.c_int, C.c_int, _p, _u32, _p, _p]),
}
_bound = None


def lib():
    global _bound
    L = abi.lib()
    if _bound is not L:
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = L
    return L


def default_params(**overrides) -> BuildParams:
    """The reference's ProceduralBuildParams defaults (circularSegments 32, samplesPerSpan 4,
    tension 1, step 500, radius 100), extensions off; keyword arguments override fields."""
    p = BuildParams()
    lib().ptProceduralDefaultParams(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(f"BuildParams has no field {k!r}")
        setattr(p, k, v)
    return p


def _text(call) -> str:
    n = call(None, 0)
    if n < 0:
        raise LSystemError(abi.last_error())
    buf = C.create_string_buffer(int(n) + 1)
    call(buf, n + 1)
    return buf.value.decode()


def evaluate_text(definition: str, iterations: int = 0, seed: int = DEFAULT_RULE_SEED, turns: bool = False,
                  all_steps: bool = False) -> str:
    """Generations of LSystemEvaluator::evaluate_steps as canonical text (modules separated by
    spaces, numbers %.17g); with all_steps one line per generation, generation 0 first."""
    d = definition.encode()
    flags = TURN_SYMBOLS if turns else 0
    return _text(lambda b, n: lib().ptLSystemEvaluateText(d, iterations, seed, flags, int(all_steps), b, n))


def parse_dump(kind: int, text: str, turns: bool = False) -> str:
    """Structural dump of a parse: 0 module string, 1 rule, 2 rules block, 3 definition."""
    d = text.encode()
    flags = TURN_SYMBOLS if turns else 0
    return _text(lambda b, n: lib().ptLSystemParseDump(kind, d, flags, b, n))


def materials_text(definition: str) -> str:
    d = definition.encode()
    return _text(lambda b, n: lib().ptLSystemMaterialsText(d, b, n))


@dataclass
class TurtleOutput:
    states: np.ndarray        # (S, 7): position, tangent, radius, every segment's states in order
    state_counts: np.ndarray  # (segments,)
    material_ids: list[str]
    groups: np.ndarray        # branch-group sizes
    leaves: np.ndarray        # (L, 9): position, tangent, right
    leaf_sizes: np.ndarray    # (L, 2): length, width

    def segments(self) -> list[np.ndarray]:
        ends = np.cumsum(self.state_counts)
        return np.split(self.states, ends[:-1]) if len(ends) else []


def turtle(definition: str, iterations: int = 0, params: BuildParams | None = None) -> TurtleOutput:
    p = params if params is not None else default_params()
    L = lib()
    d = definition.encode()
    c = TurtleCounts()
    if not L.ptTurtleExecute(d, iterations, C.byref(p), C.byref(c), None, 0, None, 0, None, 0, None, 0, None, None,
                             0):
        raise LSystemError(abi.last_error())
    st = np.zeros((c.states, 7), np.float32)
    sc = np.zeros(c.segments, np.uint32)
    ids = C.create_string_buffer(c.materialIdChars + 1)
    gr = np.zeros(c.groups, np.uint32)
    lv = np.zeros((c.leaves, 9), np.float32)
    ls = np.zeros((c.leaves, 2), np.float32)
    ok = L.ptTurtleExecute(d, iterations, C.byref(p), C.byref(c), st.ctypes.data, c.states, sc.ctypes.data,
                           c.segments, ids, c.materialIdChars + 1, gr.ctypes.data, c.groups, lv.ctypes.data,
                           ls.ctypes.data, c.leaves)
    abi.check(ok, "ptTurtleExecute")
    names = ids.value.decode().split("\n") if c.segments else []
    return TurtleOutput(st, sc, names, gr, lv, ls)


def spline_frames(states: np.ndarray, tension: float = 1.0, samples_per_span: int = 4, uniform: bool = False):
    """SplinePath over one segment: (frames (K, 13), arc length) from computeLoftFrames, or
    (samples (K, 7), arc length) from sampleUniform."""
    s = np.ascontiguousarray(states, np.float32).reshape(-1, 7)
    L = lib()
    cnt = C.c_uint32(0)
    arc = C.c_float(0.0)
    mode = 1 if uniform else 0
    abi.check(L.ptSplineFrames(s.ctypes.data, s.shape[0], tension, samples_per_span, mode, None, 0, C.byref(cnt),
                               C.byref(arc)), "ptSplineFrames")
    out = np.zeros((cnt.value, 7 if uniform else 13), np.float32)
    abi.check(L.ptSplineFrames(s.ctypes.data, s.shape[0], tension, samples_per_span, mode, out.ctypes.data,
                               cnt.value, C.byref(cnt), C.byref(arc)), "ptSplineFrames")
    return out, arc.value


@dataclass
class HostMesh:
    triangles: np.ndarray   # (N, 25) HostTriangle words
    materials: np.ndarray   # (M, 24) MaterialGpu words
    textures: np.ndarray    # (T, 16) TextureDescGpu words


def _take(h) -> HostMesh:
    L = lib()
    try:
        nt, nm, nx = C.c_uint32(), C.c_uint32(), C.c_uint32()
        L.ptHostMeshCounts(h, C.byref(nt), C.byref(nm), C.byref(nx))

        def grab(ptr, n, words):
            out = np.zeros((n, words), np.float32)
            if n:
                C.memmove(out.ctypes.data, ptr, out.nbytes)
            return out

        return HostMesh(grab(L.ptHostMeshTriangles(h), nt.value, abi.HOST_TRIANGLE_WORDS),
                        grab(L.ptHostMeshMaterials(h), nm.value, abi.MATERIAL_WORDS),
                        grab(L.ptHostMeshTextures(h), nx.value, abi.TEXTURE_WORDS))
    finally:
        L.ptHostMeshDestroy(h)


def build_host_mesh(definition: str, iterations: int = 0, translation=(0.0, 0.0, 0.0), rotation_deg=(0.0, 0.0, 0.0),
                    params: BuildParams | None = None) -> HostMesh:
    """ProceduralMeshBuilder::buildHostMesh; raises LSystemError when it fails."""
    p = params if params is not None else default_params()
    h = lib().ptProceduralBuildHostMesh(definition.encode(), iterations, abi.Vec3C(*map(float, translation)),
                                        abi.Vec3C(*map(float, rotation_deg)), C.byref(p))
    if not h:
        raise LSystemError(abi.last_error())
    return _take(h)


@dataclass
class ProceduralInstance:
    command_string: str
    iterations: int = 0
    translation: tuple = (0.0, 0.0, 0.0)
    rotation_deg: tuple = (0.0, 0.0, 0.0)


def _instances(instances):
    keep = [it.command_string.encode() for it in instances]
    arr = (Instance * max(len(instances), 1))()
    for i, it in enumerate(instances):
        arr[i].commandString = keep[i]
        arr[i].iterations = it.iterations
        arr[i].translation = abi.Vec3C(*map(float, it.translation))
        arr[i].rotationDeg = abi.Vec3C(*map(float, it.rotation_deg))
    return arr, keep


def build_scene(instances: list[ProceduralInstance], params: BuildParams | None = None) -> HostMesh:
    """The HostMesh half of meshSceneBuild with full build parameters."""
    p = params if params is not None else default_params()
    arr, _keep = _instances(instances)
    h = lib().ptProceduralBuildScene(arr, len(instances), C.byref(p))
    if not h:
        raise LSystemError(abi.last_error())
    return _take(h)


def mesh_scene_build(instances: list[ProceduralInstance], scene, circular_segments: int = 64,
                     crease_angle_deg: float = 50.0) -> None:
    """meshSceneBuild into an abi.Scene (uploads; needs the GPU)."""
    arr, _keep = _instances(instances)
    mp = MeshSceneParams(circular_segments, crease_angle_deg)
    abi.check(lib().ptMeshSceneBuild(arr, len(instances), scene.handle, C.byref(mp), None), "ptMeshSceneBuild")
    # triangleCount of MeshAccelSceneGpu: after eight pointers and nodeCount
    scene.triangle_count = int(np.frombuffer(scene.description()[68:72], np.uint32)[0])
