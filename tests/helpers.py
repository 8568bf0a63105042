"""Test scenes and comparison helpers (inputs only; both the oracle and the GPU consume them)."""
from __future__ import annotations

import math

import numpy as np

from lsystemeditor_amd import abi, scenes


def dense_plant(width=64, height=64, spp=8, iterations=3, radius=0.35, max_depth=8, seed=0, min_spp=None,
                threshold=0.001, materials="zoo", thin_lens=False, env_intensity=1.0):
    """A small plant with thick branches inside a box, seen up close, so most paths bounce.

    materials="zoo" cycles the branch material over diffuse, rough metal, glass, thin glass,
    subsurface, textured (grid albedo, stripe roughness) and a diffuse-roughness override.
    """
    body, _, _, lo, hi = scenes.plant(scenes.PLANAR_PLANT, iterations, scenes.planar_params(8, radius))
    per = 8 * 4     # one capped 8-sided tube per segment: 16 wall + 16 cap triangles
    seg = np.arange(body.shape[0]) // per
    parts = []
    nmat = 0
    mats = []
    if materials == "zoo":
        tex = [0, 0, 0, 0, 0, 0, 0, 0]
        mats = [
            abi.pack_material(0.4, 0.3, 0.2, roughness=0.6),
            abi.pack_material(0.9, 0.6, 0.3, roughness=0.3, metallic=1.0),
            abi.pack_material(1.0, 1.0, 1.0, roughness=0.0, transmission=1.0, ior=1.5),
            abi.pack_material(0.2, 0.8, 0.1, roughness=0.5, transmission=1.0, thin=1.0),
            abi.pack_material(0.9, 0.85, 0.7, roughness=0.85, subsurface=0.8, diffuseRoughness=0.85,
                              scatterRadius=(0.05, 0.04, 0.02)),
            abi.pack_material(0.5, 0.5, 0.5, roughness=0.5, textures=[1, 2, 0, 0, 0, 0, 0, 0]),
            abi.pack_material(0.6, 0.6, 0.6, roughness=0.2, diffuseRoughness=0.9, specular=0.5),
            abi.pack_material(1.0, 1.0, 1.0, roughness=0.35, transmission=1.0, ior=1.33),
        ]
        nmat = len(mats)
        body[:, 24] = (seg % nmat).astype(np.uint32).view(np.float32)
        parts.append(body)
    elif materials == "diff":
        # opaque materials with every differentiable parameter strictly inside its clamp range
        mats = [abi.pack_material(0.4, 0.3, 0.2, roughness=0.6, metallic=0.2, specular=0.5),
                abi.pack_material(0.2, 0.6, 0.15, roughness=0.35, metallic=0.1, specular=0.7)]
        nmat = 2
        body[:, 24] = (seg % nmat).astype(np.uint32).view(np.float32)
        parts.append(body)
    else:
        mats = [abi.pack_material(0.4, 0.3, 0.2, roughness=0.6)]
        nmat = 1
        body[:, 24] = np.zeros(body.shape[0], np.uint32).view(np.float32)
        parts.append(body)
    c = 0.5 * (lo + hi)
    ext = 0.5 * (hi - lo) + 2.0
    box_m = nmat
    light_m = nmat + 1
    if materials == "diff":
        mats.append(abi.pack_material(0.7, 0.7, 0.7, roughness=0.9, metallic=0.05, specular=0.5))
        mats.append(abi.pack_material(1.0, 0.9, 0.8, emission=8.0, roughness=0.5, metallic=0.1, specular=0.5))
    else:
        mats.append(abi.pack_material(0.7, 0.7, 0.7, roughness=0.9))
        mats.append(abi.pack_material(1.0, 0.9, 0.8, emission=8.0))
    # floor, back wall, left wall (open front and right side so the environment is visible)
    parts.append(scenes.quad((c[0], lo[1] - 0.01, c[2]), (0.0, 0.0, ext[2]), (ext[0], 0.0, 0.0), box_m))
    parts.append(scenes.quad((c[0], c[1], c[2] - ext[2]), (ext[0], 0.0, 0.0), (0.0, ext[1], 0.0), box_m))
    parts.append(scenes.quad((c[0] - ext[0], c[1], c[2]), (0.0, ext[1], 0.0), (0.0, 0.0, ext[2]), box_m))
    parts.append(scenes.quad((c[0], hi[1] + 1.0, c[2]), (0.3 * ext[0], 0.0, 0.0), (0.0, 0.0, 0.3 * ext[2]), light_m))
    tris = np.concatenate(parts)
    mats = np.stack(mats)
    textures = np.zeros((3, abi.TEXTURE_WORDS), dtype=np.float32)
    textures[1, 0:1] = np.array([2], dtype=np.uint32).view(np.float32)   # Grid2D albedo
    textures[1, 4:8] = [4.0, 2.0, 0.2, 0.0]
    textures[1, 8:12] = [0.9, 0.2, 0.2, 0.0]
    textures[1, 12:16] = [0.1, 0.2, 0.9, 0.0]
    textures[2, 0:1] = np.array([3], dtype=np.uint32).view(np.float32)   # Stripe1D roughness
    textures[2, 4:8] = [6.0, 0.5, 0.1, 0.9]
    d = np.array([-0.2, -0.1, -1.0])
    d /= np.linalg.norm(d)
    dist = 0.9 * float(np.linalg.norm(hi - lo))
    pos = c - d * dist
    q = abi.quat_look(d)
    if thin_lens:
        cam = abi.pack_camera(pos, q, fov_y=math.radians(50.0), aspect=width / height, near=0.05,
                              aperture_radius=0.3, focus_distance=dist, focus_valid=1)
    else:
        cam = abi.pack_camera(pos, q, fov_y=math.radians(50.0), aspect=width / height, near=0.05)
    params = abi.pack_params(background=(0.5, 0.6, 0.8), environment_intensity=env_intensity,
                             max_path_depth=max_depth, min_samples=spp if min_spp is None else min_spp,
                             max_samples_per_pixel=spp, relative_error_threshold=threshold)
    return tris, mats, textures, cam, params


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint32)


def mismatch_report(name, got, want):
    g = bits(np.asarray(got, dtype=np.float32))
    w = bits(np.asarray(want, dtype=np.float32))
    bad = np.nonzero(g != w)
    return f"{name}: {len(bad[0])} of {g.size} words differ (first at {tuple(x[0] for x in bad) if len(bad[0]) else None})"
