"""The benchmark and test configurations, built by the procedural mesher.

Every plant is an L-system definition run through the library's C++ pipeline (include/
pt_procedural.h: evaluator, turtle, Hermite loft frames, capped tubes, Mat(id) tables, root
transform -- the reference's ProceduralMeshBuilder::buildHostMesh, Application/src/Procedural/
ProceduralMeshBuilder.cpp:266-330), with the turtle extensions the workloads need (turn symbols
+ - & ^ / \\ with a fixed angle, radius decay per bracket depth, seeded yaw/pitch/roll jitter, leaf
quads). The turtle grows along -Z; a root pitch of +90 degrees stands the plant up along +Y. The
ground quad and rectangular area lights are appended at the HostMesh boundary
(MeshAccelScene.cpp:141), after the plant's materials.

Rule choice uses the evaluator's std::mt19937_64 stream and the jitter a SplitMix64 stream, both
keyed by the scene seed, so the triangles are identical on every machine (IEEE single precision,
the C library's cosf/sinf/asinf).
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass

import numpy as np

from . import procedural
from .abi import pack_camera, pack_host_triangles, pack_material, pack_params, quat_look

MASK64 = (1 << 64) - 1
STAND_UP = (0.0, 90.0, 0.0)   # root rotation: pitch +90 about +X takes the turtle's -Z to +Y


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform() returns (x >> 11) * 2^-53 in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


# F -> F[+F]F[-F]F in the definition language (turn symbols need PT_TURTLE_TURN_SYMBOLS)
PLANAR_PLANT = """Mat(0) = {0.4, 0.3, 0.2, 0.6, 0}
F
F -> F [ + F ] F [ - F ] F
"""

# three stochastic rules (0.4 / 0.3 / 0.3); every branch ends in a leaf of material 1
BUSH = """Mat(0) = {0.4, 0.3, 0.2, 0.6}
Mat(1) = {0.2, 0.5, 0.1, 0.5, 0, 0, 0, 1.5, 0, 0, -1, 0, 0, 0, 0.5}
F
F :: 0.4 -> F [ + F Mat(1) Leaf Mat(0) ] [ - F Mat(1) Leaf Mat(0) ] F [ & F Mat(1) Leaf Mat(0) ]
F :: 0.3 -> F [ ^ F Mat(1) Leaf Mat(0) ] F [ / F Mat(1) Leaf Mat(0) ] F
F :: 0.3 -> F [ & + F Mat(1) Leaf Mat(0) ] F [ ^ - F Mat(1) Leaf Mat(0) ] F
"""


def planar_params(sides: int = 8, radius: float = 0.05) -> procedural.BuildParams:
    return procedural.default_params(circularSegments=sides, turtleFlags=procedural.TURN_SYMBOLS, turnAngleDeg=25.7,
                                     defaultStepLength=1.0, defaultRadius=radius, radiusDecay=0.9)


def bush_params(seed: int, sides: int = 12) -> procedural.BuildParams:
    return procedural.default_params(
        circularSegments=sides, turtleFlags=procedural.TURN_SYMBOLS | procedural.LEAVES, turnAngleDeg=25.7,
        rollAngleDeg=137.5, defaultStepLength=1.0, defaultRadius=0.6, radiusDecay=0.9, jitterDeg=10.0,
        jitterSeed=0x5EED0000 + seed, leafLength=2.0, leafWidth=1.2, ruleSeed=0x5EED0000 + seed)


def plant(definition: str, iterations: int, params: procedural.BuildParams):
    """(triangles, materials, textures, lo, hi) of a standing plant."""
    m = procedural.build_host_mesh(definition, iterations, (0.0, 0.0, 0.0), STAND_UP, params)
    v = m.triangles[:, 0:9].reshape(-1, 3).astype(np.float64)
    return m.triangles, m.materials, m.textures, v.min(0), v.max(0)


def quad(center, ex, ey, material: int, flip: bool = False) -> np.ndarray:
    """Rectangle center +- ex +- ey as two triangles; the winding normal is ex x ey (or its opposite)."""
    c = np.asarray(center, dtype=np.float64)
    ex = np.asarray(ex, dtype=np.float64)
    ey = np.asarray(ey, dtype=np.float64)
    p = [c - ex - ey, c + ex - ey, c + ex + ey, c - ex + ey]
    if flip:
        p = [p[0], p[3], p[2], p[1]]
    n = np.cross(p[1] - p[0], p[2] - p[0])
    n = n / np.linalg.norm(n)
    v0 = np.stack([p[0], p[0]])
    v1 = np.stack([p[1], p[2]])
    v2 = np.stack([p[2], p[3]])
    nn = np.stack([n, n])
    uv = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    return pack_host_triangles(v0, v1, v2, nn, nn, nn, np.stack([uv[0], uv[0]]), np.stack([uv[1], uv[2]]),
                               np.stack([uv[2], uv[3]]), np.full(2, material, dtype=np.uint32))


@dataclass
class SceneSpec:
    name: str
    triangles: np.ndarray   # (N, 25) HostTriangle words
    materials: np.ndarray   # (M, 24) MaterialGpu words
    camera: np.ndarray      # (20,) CameraGpu words
    params: np.ndarray      # (11,) RenderParamsGpu words
    width: int
    height: int
    spp: int | None         # fixed samples per pixel (None = adaptive)
    info: dict

    def sha256(self) -> str:
        return hashlib.sha256(np.ascontiguousarray(self.triangles).tobytes()).hexdigest()


def _frame_camera(lo, hi, width, height, fov_deg=45.0, direction=(0.0, -0.15, -1.0)):
    center = 0.5 * (lo + hi)
    radius = 0.5 * float(np.linalg.norm(hi - lo))
    fov = math.radians(fov_deg)
    dist = radius / math.sin(0.5 * fov) * 0.85
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    pos = center - d * dist
    q = quat_look(d)
    return pack_camera(pos, q, fov_y=fov, aspect=width / height, near=0.1, far=1.0e5)


def config0(width: int = 256, height: int = 256, spp: int = 16, iterations: int = 4) -> SceneSpec:
    """Plain method, small: F -> F[+F]F[-F]F, 25.7 deg, 4 iterations (625 segments), 8-sided capped
    tubes (length 1, radius 0.05, x0.9 per bracket depth; 32 triangles each, 20,000 in all), 20x20
    ground quad, one 2x2 area light (emission 10, white), Principled base (0.4, 0.3, 0.2) roughness
    0.6 metallic 0, pinhole camera, max depth 8, fixed spp."""
    plant_t, plant_m, _, lo, hi = plant(PLANAR_PLANT, iterations, planar_params())
    ground = quad((0.0, 0.0, 0.0), (0.0, 0.0, 10.0), (10.0, 0.0, 0.0), 1)
    light = quad((0.0, hi[1] + 4.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2)  # winding normal -Y (downward)
    tris = np.concatenate([plant_t, ground, light])
    mats = np.concatenate([plant_m, np.stack([pack_material(0.4, 0.3, 0.2, roughness=0.6, metallic=0.0),
                                              pack_material(1.0, 1.0, 1.0, roughness=0.6, emission=10.0)])])
    cam = _frame_camera(lo, hi, width, height)
    params = pack_params(max_path_depth=8, russian_roulette_min_depth=3, min_samples=spp,
                         max_samples_per_pixel=spp, relative_error_threshold=0.001)
    return SceneSpec("C0", tris, mats, cam, params, width, height, spp,
                     {"segments": 5 ** iterations, "iterations": iterations, "definition": PLANAR_PLANT})


def config1(width: int = 1024, height: int = 1024, iterations: int = 7, seed: int = 1,
            min_spp: int = 16, max_spp: int = 256, threshold: float = 0.01) -> SceneSpec:
    """Plain method, large: stochastic 3-D bush (BUSH: 3 rules, 0.4/0.3/0.3), yaw/pitch/roll
    jitter uniform in +-10 deg before every F, 7 iterations, 12-sided capped tubes, a 2 x 1.2 leaf
    quad at every branch tip (about 3.8M triangles at 7 iterations), 2 area lights + uniform
    environment, adaptive 16..256 spp at relative error 0.01, max depth 12."""
    plant_t, plant_m, _, lo, hi = plant(BUSH, iterations, bush_params(seed))
    span = hi - lo
    gsize = 0.75 * float(max(span[0], span[2])) + 10.0
    gc = (0.5 * (lo[0] + hi[0]), lo[1], 0.5 * (lo[2] + hi[2]))
    ground = quad(gc, (0.0, 0.0, gsize), (gsize, 0.0, 0.0), 2)
    ls_ = 0.1 * float(max(span[0], span[2])) + 2.0
    l1 = quad((lo[0] + 0.3 * span[0], hi[1] + 0.1 * span[1], gc[2]), (ls_, 0.0, 0.0), (0.0, 0.0, ls_), 3)
    l2 = quad((hi[0] - 0.2 * span[0], hi[1] + 0.2 * span[1], gc[2] + 0.2 * span[2]), (ls_, 0.0, 0.0),
              (0.0, 0.0, ls_), 4)
    tris = np.concatenate([plant_t, ground, l1, l2])
    mats = np.concatenate([plant_m, np.stack([pack_material(0.35, 0.35, 0.3, roughness=0.8),
                                              pack_material(1.0, 0.95, 0.9, emission=10.0),
                                              pack_material(0.9, 0.95, 1.0, emission=5.0)])])
    cam = _frame_camera(lo, hi, width, height, fov_deg=40.0, direction=(-0.3, -0.25, -1.0))
    params = pack_params(background=(0.6, 0.7, 0.9), environment_intensity=1.0, max_path_depth=12,
                         russian_roulette_min_depth=3, min_samples=min_spp, max_samples_per_pixel=max_spp,
                         relative_error_threshold=threshold)
    leaves = int(np.count_nonzero(plant_t[:, 24].view(np.uint32) == 1) // 2)
    return SceneSpec("C1", tris, mats, cam, params, width, height, None,
                     {"segments": 5 ** iterations, "leaves": leaves, "iterations": iterations, "seed": seed,
                      "definition": BUSH})


def config2(width: int = 512, height: int = 512, spp: int = 32, iterations: int = 7, seed: int = 1,
            c1: SceneSpec | None = None) -> SceneSpec:
    """Differentiable render of the C1 bush at 512x512, 32 spp, max depth 8. Material groups bark,
    leaf, ground and the two lights; every group's base colour, roughness, metallic, specular and
    (lights) emission are parameters, so their values sit strictly inside the clamp ranges."""
    if c1 is None:
        c1 = config1(width, height, iterations=iterations, seed=seed)
    camera = c1.camera.copy()
    camera[9] = np.float32(width / height)  # aspect
    mats = np.stack([pack_material(0.4, 0.3, 0.2, roughness=0.6, metallic=0.05, specular=0.5),
                     pack_material(0.2, 0.5, 0.1, roughness=0.5, metallic=0.05, specular=0.5),
                     pack_material(0.35, 0.35, 0.3, roughness=0.8, metallic=0.05, specular=0.5),
                     pack_material(1.0, 0.95, 0.9, roughness=0.5, metallic=0.05, specular=0.5, emission=10.0),
                     pack_material(0.9, 0.95, 1.0, roughness=0.5, metallic=0.05, specular=0.5, emission=5.0)])
    params = pack_params(background=(0.6, 0.7, 0.9), environment_intensity=1.0, max_path_depth=8,
                         russian_roulette_min_depth=3, min_samples=spp, max_samples_per_pixel=spp,
                         relative_error_threshold=0.01)
    return SceneSpec("C2", c1.triangles, mats, camera, params, width, height, spp, dict(c1.info))


def perturbed_target(theta: np.ndarray, seed: int = 7, spread: float = 0.2) -> np.ndarray:
    """Target parameters: each one scaled by an independent uniform factor in [1 - spread,
    1 + spread] (SplitMix64-seeded), clamped back inside the differentiable ranges."""
    rng = SplitMix64(0xC2000000 + seed)
    out = np.array(theta, dtype=np.float64, copy=True)
    for i in range(out.shape[0]):
        for k in range(out.shape[1]):
            u = rng.uniform()
            out[i, k] *= 1.0 - spread + 2.0 * spread * u
    out[:, 0:3] = np.clip(out[:, 0:3], 0.0, 1.0)
    out[:, 3:6] = np.clip(out[:, 3:6], 0.01, 0.99)
    return out.astype(np.float32)
