"""The binned-SAH build's topology, node for node, against its CPU restatement (tests/sah_model.py):
the same Morton order, bins, lowest-cost boundaries, middle splits, depth cap, stable partitions
and level-by-level node numbering. A miscounted bin, a wrong cost or split choice, or a partition
that loses the Morton order gives a valid tree with identical renders; this comparison catches
it. The soups reach every path of the GPU build: segments filling whole blocks (the uniform LDS
path), blocks spanning 2-32 segments (local histograms), blocks spanning more (global bins),
segments inside one wave (plain stores), all centroids of a segment in one bin (middle splits),
a flat axis, and unbalanced chains. (The depth cap, level + ceil(log2 count) >= 48, needs trees
deeper than float32 coordinates allow for these soups; it is not reached here.)"""
import numpy as np
import pytest

from lsystemeditor_amd import abi, scenes
from tests.sah_model import build
from tests.test_gpu_parity import random_soup

pytestmark = pytest.mark.gpu


def clustered_soup(n, clusters, seed):
    """Triangles whose centroids repeat exactly within each cluster (bins cannot separate them)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-10, 10, (clusters, 3))
    offs = rng.normal(0, 0.3, (n, 3, 3))
    offs -= offs.mean(1, keepdims=True)
    v = centres[rng.integers(0, clusters, n)][:, None, :] + offs
    return abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2], material=np.zeros(n, np.uint32))


def flat_soup(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-4, 4, (n, 3, 3))
    v[..., 2] = 0.5
    return abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2], material=np.zeros(n, np.uint32))


def geometric_chain(n):
    """Small triangles at x = 1.1^k: the bins of every segment put all but its outermost few
    triangles in the first bin, so the SAH peels a few off per level (unbalanced, deep trees)."""
    x = 1.1 ** np.arange(n, dtype=np.float64)
    v = np.zeros((n, 3, 3))
    v[:, :, 0] = x[:, None] + np.array([0.0, 1e-3, 0.0]) * x[:, None]
    v[:, :, 1] = np.array([0.0, 0.0, 1e-3]) * x[:, None]
    return abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2], material=np.zeros(n, np.uint32))


CASES = [(f"random{n}", lambda n=n: random_soup(n, 500 + n)) for n in (2, 3, 5, 64, 257, 1000, 4097, 20000)] + [
    ("clustered", lambda: clustered_soup(3000, 40, 11)),
    ("flat", lambda: flat_soup(2500, 12)),
    ("chain", lambda: geometric_chain(400)),
    ("c0", lambda: scenes.config0().triangles),
]


@pytest.mark.parametrize("bvh", ["sah", "sah-binary"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_sah_tree_equals_cpu_restatement(cuda, bvh, name, make):
    tris = make()
    n = tris.shape[0]
    sc = abi.Scene(tris, abi.pack_material()[None], bvh=bvh)
    nodes, slots, levels = sc.read_layout()
    desc = np.frombuffer(sc.description(), np.uint8)
    root_ref = int(np.frombuffer(desc[84:88].tobytes(), np.int32)[0])
    root_box = np.frombuffer(desc[88:112].tobytes(), np.float32)
    want_nodes, want_perm, want_box, want_root, want_levels = build(tris)
    assert levels == want_levels, (name, levels, want_levels)
    assert root_ref == want_root
    assert np.array_equal(root_box.view(np.uint32), want_box.view(np.uint32))
    assert np.array_equal(slots[:, 3].view(np.int32), want_perm), f"{name}: leaf order differs"
    got = nodes[:, :14].view(np.uint32)
    want = want_nodes.view(np.uint32)
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, f"{name}/{bvh}: {bad.size} of {n - 1} nodes differ (first {bad[:5].tolist()})"
