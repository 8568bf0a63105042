"""Deterministic L-system scene generator for the benchmark and test configurations.

It plays the role of the reference's procedural mesher (Application/src/Procedural/
Turtle.cpp:168-283, Loft.cpp:108-155) for the workloads used here: string rewriting,
a 3-D turtle with branch stack, lofted open cylinders per segment, leaf quads at branch
tips, a ground quad and rectangular area lights. Output enters the renderer at the
HostMesh boundary (MeshAccelScene.cpp:141).

Turtle symbols: F draw a segment; + - yaw about U; & ^ pitch about L; / \\ roll about H;
[ ] push/pop (branch depth scales the radius). Stochastic rule choice and per-segment
yaw/pitch/roll jitter come from a SplitMix64 stream keyed by the scene seed, so the
triangles are identical on every machine (all geometry is computed in IEEE double with
separate multiplies and adds, then rounded to float32).
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

from .abi import pack_camera, pack_host_triangles, pack_material, pack_params, quat_look

MASK64 = (1 << 64) - 1


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


@dataclass
class LSystem:
    axiom: str
    rules: dict[str, list[tuple[float, str]]]  # predecessor -> [(probability, successor)]
    iterations: int

    def expand(self, rng: SplitMix64 | None = None) -> str:
        s = self.axiom
        for _ in range(self.iterations):
            out = []
            for ch in s:
                options = self.rules.get(ch)
                if not options:
                    out.append(ch)
                elif len(options) == 1:
                    out.append(options[0][1])
                else:
                    u = rng.uniform()
                    acc = 0.0
                    pick = options[-1][1]
                    for p, succ in options:
                        acc += p
                        if u < acc:
                            pick = succ
                            break
                    out.append(pick)
            s = "".join(out)
        return s


@dataclass
class TurtleParams:
    yaw_deg: float = 25.7
    pitch_deg: float = 25.7
    roll_deg: float = 137.5
    length: float = 1.0
    radius: float = 0.05
    radius_decay: float = 0.9
    jitter_deg: float = 0.0
    leaves: bool = False
    leaf_length: float = 1.0
    leaf_width: float = 0.6


@dataclass
class TurtleOutput:
    p0: np.ndarray
    p1: np.ndarray
    basis_l: np.ndarray
    basis_u: np.ndarray
    radius: np.ndarray
    leaf_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    leaf_h: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    leaf_l: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))


def _rot(a, b, ang):
    c = math.cos(ang)
    s = math.sin(ang)
    return ((a[0] * c + b[0] * s, a[1] * c + b[1] * s, a[2] * c + b[2] * s),
            (b[0] * c - a[0] * s, b[1] * c - a[1] * s, b[2] * c - a[2] * s))


def run_turtle(commands: str, tp: TurtleParams, rng: SplitMix64 | None = None) -> TurtleOutput:
    """Interprets the command string. Initial frame: H = +Y (heading), L = -X, U = +Z."""
    pos = (0.0, 0.0, 0.0)
    H = (0.0, 1.0, 0.0)
    Lv = (-1.0, 0.0, 0.0)
    U = (0.0, 0.0, 1.0)
    depth = 0
    stack = []
    yaw = math.radians(tp.yaw_deg)
    pitch = math.radians(tp.pitch_deg)
    roll = math.radians(tp.roll_deg)
    jit = math.radians(tp.jitter_deg)
    p0, p1, bl, bu, rad = [], [], [], [], []
    lp, lh, ll = [], [], []
    for ch in commands:
        if ch == "F":
            if jit > 0.0 and rng is not None:
                H, Lv = _rot(H, Lv, (2.0 * rng.uniform() - 1.0) * jit)
                H, U = _rot(H, U, (2.0 * rng.uniform() - 1.0) * jit)
                Lv, U = _rot(Lv, U, (2.0 * rng.uniform() - 1.0) * jit)
            end = (pos[0] + H[0] * tp.length, pos[1] + H[1] * tp.length, pos[2] + H[2] * tp.length)
            p0.append(pos)
            p1.append(end)
            bl.append(Lv)
            bu.append(U)
            rad.append(tp.radius * tp.radius_decay ** depth)
            pos = end
        elif ch == "+":
            H, Lv = _rot(H, Lv, yaw)
        elif ch == "-":
            H, Lv = _rot(H, Lv, -yaw)
        elif ch == "&":
            H, U = _rot(H, U, pitch)
        elif ch == "^":
            H, U = _rot(H, U, -pitch)
        elif ch == "/":
            Lv, U = _rot(Lv, U, roll)
        elif ch == "\\":
            Lv, U = _rot(Lv, U, -roll)
        elif ch == "[":
            stack.append((pos, H, Lv, U, depth))
            depth += 1
        elif ch == "]":
            if tp.leaves:
                lp.append(pos)
                lh.append(H)
                ll.append(Lv)
            pos, H, Lv, U, depth = stack.pop()
    f64 = lambda a: np.asarray(a, dtype=np.float64).reshape(-1, 3)  # noqa: E731
    return TurtleOutput(f64(p0), f64(p1), f64(bl), f64(bu), np.asarray(rad, dtype=np.float64),
                        f64(lp), f64(lh), f64(ll))


def cylinders(t: TurtleOutput, sides: int, material: int) -> np.ndarray:
    """Open cylinders (2*sides triangles per segment), radial normals, UV = (j/sides, 0|1)."""
    s = t.p0.shape[0]
    ang = [2.0 * math.pi * j / sides for j in range(sides + 1)]
    cs = np.array([math.cos(a) for a in ang])
    sn = np.array([math.sin(a) for a in ang])
    off = cs[None, :, None] * t.basis_l[:, None, :] + sn[None, :, None] * t.basis_u[:, None, :]  # (S, k+1, 3)
    r = t.radius[:, None, None]
    bot = t.p0[:, None, :] + r * off
    top = t.p1[:, None, :] + r * off
    j0 = np.arange(sides)
    j1 = j0 + 1
    u0 = (j0 / sides)[None, :].repeat(s, 0)
    u1 = (j1 / sides)[None, :].repeat(s, 0)
    zero = np.zeros_like(u0)
    one = np.ones_like(u0)

    def tri(a, b, c, na, nb, nc, ua, ub, uc):
        return (a.reshape(-1, 3), b.reshape(-1, 3), c.reshape(-1, 3), na.reshape(-1, 3), nb.reshape(-1, 3),
                nc.reshape(-1, 3), np.stack(ua, -1).reshape(-1, 2), np.stack(ub, -1).reshape(-1, 2),
                np.stack(uc, -1).reshape(-1, 2))

    A = tri(bot[:, j0], bot[:, j1], top[:, j1], off[:, j0], off[:, j1], off[:, j1], (u0, zero), (u1, zero),
            (u1, one))
    B = tri(bot[:, j0], top[:, j1], top[:, j0], off[:, j0], off[:, j1], off[:, j0], (u0, zero), (u1, one),
            (u0, one))
    # interleave so that the two triangles of each quad are adjacent
    parts = [np.stack([a, b], axis=1).reshape(-1, a.shape[-1]) for a, b in zip(A, B)]
    return pack_host_triangles(*parts, material=np.full(parts[0].shape[0], material, dtype=np.uint32))


def leaves(t: TurtleOutput, length: float, width: float, material: int) -> np.ndarray:
    """One quad (two triangles) per branch tip, spanned by the heading and the left axis."""
    if t.leaf_pos.shape[0] == 0:
        return np.zeros((0, 25), dtype=np.float32)
    p = t.leaf_pos
    h = t.leaf_h * length
    w = t.leaf_l * (0.5 * width)
    a, b, c, d = p - w, p + w, p + w + h, p - w + h
    n = np.cross(t.leaf_l, t.leaf_h)
    uv = [np.array([0.0, 0.0]), np.array([1.0, 0.0]), np.array([1.0, 1.0]), np.array([0.0, 1.0])]
    k = p.shape[0]
    UV = [np.broadcast_to(x, (k, 2)) for x in uv]
    v0 = np.stack([a, a], 1).reshape(-1, 3)
    v1 = np.stack([b, c], 1).reshape(-1, 3)
    v2 = np.stack([c, d], 1).reshape(-1, 3)
    nn = np.stack([n, n], 1).reshape(-1, 3)
    t0 = np.stack([UV[0], UV[0]], 1).reshape(-1, 2)
    t1 = np.stack([UV[1], UV[2]], 1).reshape(-1, 2)
    t2 = np.stack([UV[2], UV[3]], 1).reshape(-1, 2)
    return pack_host_triangles(v0, v1, v2, nn, nn, nn, t0, t1, t2, np.full(2 * k, material, dtype=np.uint32))


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
    """Plain method, small: F -> F[+F]F[-F]F, 25.7 deg, 4 iterations (625 segments), 8-sided open
    cylinders (length 1, radius 0.05, x0.9 per bracket depth), 20x20 ground quad, one 2x2 area light
    (emission 10, white), Principled base (0.4, 0.3, 0.2) roughness 0.6 metallic 0, pinhole camera,
    max depth 8, fixed spp."""
    ls = LSystem("F", {"F": [(1.0, "F[+F]F[-F]F")]}, iterations)
    t = run_turtle(ls.expand(), TurtleParams(yaw_deg=25.7, length=1.0, radius=0.05, radius_decay=0.9))
    plant = cylinders(t, 8, 0)
    lo = np.minimum(t.p0.min(0), t.p1.min(0))
    hi = np.maximum(t.p0.max(0), t.p1.max(0))
    ground = quad((0.0, 0.0, 0.0), (0.0, 0.0, 10.0), (10.0, 0.0, 0.0), 1)
    light = quad((0.0, hi[1] + 4.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2)  # winding normal -Y (downward)
    tris = np.concatenate([plant, ground, light])
    mats = np.stack([pack_material(0.4, 0.3, 0.2, roughness=0.6, metallic=0.0),
                     pack_material(0.4, 0.3, 0.2, roughness=0.6, metallic=0.0),
                     pack_material(1.0, 1.0, 1.0, roughness=0.6, emission=10.0)])
    cam = _frame_camera(lo, hi, width, height)
    params = pack_params(max_path_depth=8, russian_roulette_min_depth=3, min_samples=spp,
                         max_samples_per_pixel=spp, relative_error_threshold=0.001)
    return SceneSpec("C0", tris, mats, cam, params, width, height, spp,
                     {"segments": int(t.p0.shape[0]), "iterations": iterations})


def config1(width: int = 1024, height: int = 1024, iterations: int = 7, seed: int = 1,
            min_spp: int = 16, max_spp: int = 256, threshold: float = 0.01) -> SceneSpec:
    """Plain method, large: stochastic 3-D bush, 3 rules (0.4/0.3/0.3), yaw/pitch/roll jitter
    uniform in +-10 deg, 7 iterations, 12-sided branches, leaf quads at branch tips (~2M triangles),
    2 area lights + uniform environment, adaptive 16..256 spp at relative error 0.01, max depth 12."""
    rules = {"F": [(0.4, "F[+F][-F]F[&F]"), (0.3, "F[^F]F[/F]F"), (0.3, "F[&+F]F[^-F]F")]}
    rng = SplitMix64(0x5EED0000 + seed)
    ls = LSystem("F", rules, iterations)
    t = run_turtle(ls.expand(rng), TurtleParams(yaw_deg=25.7, pitch_deg=25.7, roll_deg=137.5, length=1.0,
                                                radius=0.6, radius_decay=0.9, jitter_deg=10.0, leaves=True),
                   rng)
    wood = cylinders(t, 12, 0)
    leaf = leaves(t, 2.0, 1.2, 1)
    lo = np.minimum(t.p0.min(0), t.p1.min(0))
    hi = np.maximum(t.p0.max(0), t.p1.max(0))
    span = hi - lo
    gsize = 0.75 * float(max(span[0], span[2])) + 10.0
    gc = (0.5 * (lo[0] + hi[0]), lo[1], 0.5 * (lo[2] + hi[2]))
    ground = quad(gc, (0.0, 0.0, gsize), (gsize, 0.0, 0.0), 2)
    ls_ = 0.1 * float(max(span[0], span[2])) + 2.0
    l1 = quad((lo[0] + 0.3 * span[0], hi[1] + 0.1 * span[1], gc[2]), (ls_, 0.0, 0.0), (0.0, 0.0, ls_), 3)
    l2 = quad((hi[0] - 0.2 * span[0], hi[1] + 0.2 * span[1], gc[2] + 0.2 * span[2]), (ls_, 0.0, 0.0),
              (0.0, 0.0, ls_), 4)
    tris = np.concatenate([wood, leaf, ground, l1, l2])
    mats = np.stack([pack_material(0.4, 0.3, 0.2, roughness=0.6),
                     pack_material(0.2, 0.5, 0.1, roughness=0.5, specular=0.5),
                     pack_material(0.35, 0.35, 0.3, roughness=0.8),
                     pack_material(1.0, 0.95, 0.9, emission=10.0),
                     pack_material(0.9, 0.95, 1.0, emission=5.0)])
    cam = _frame_camera(lo, hi, width, height, fov_deg=40.0, direction=(-0.3, -0.25, -1.0))
    params = pack_params(background=(0.6, 0.7, 0.9), environment_intensity=1.0, max_path_depth=12,
                         russian_roulette_min_depth=3, min_samples=min_spp, max_samples_per_pixel=max_spp,
                         relative_error_threshold=threshold)
    return SceneSpec("C1", tris, mats, cam, params, width, height, None,
                     {"segments": int(t.p0.shape[0]), "leaves": int(t.leaf_pos.shape[0]),
                      "iterations": iterations, "seed": seed})


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
