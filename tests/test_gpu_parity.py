"""GPU parity: the gfx950 kernels against the oracle on identical inputs.

Bar: bit-exact for the BVH, traversal results, sample counts and convergence flags; the
accumulated radiance, Welford mean and M2 are compared bit for bit as well, with a reported
fallback bound of 1e-4 relative for the rare transcendental-free divergences (none expected:
both sides use the same correctly rounded operations in the same order).
"""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi, scenes
from oracle import oracle as O
from tests.helpers import dense_plant, mismatch_report

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


@pytest.fixture(scope="module")
def render_mod(cuda):
    from lsystemeditor_amd import render
    return render


def oracle_scene(tris, mats, textures=None):
    return O.OracleScene(abi.host_to_gpu_triangles(tris), mats, textures)


def random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    v = c + rng.normal(0, 0.6, (n, 3, 3))
    # duplicate some centroids exactly to exercise the tie-break
    v[: n // 8] = v[n // 8: 2 * (n // 8)]
    return abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2], material=np.zeros(n, np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 1000, 4097])
def test_bvh_matches_oracle(cuda, n):
    tris = random_soup(n, 7 + n)
    mats = abi.pack_material()[None]
    sc = abi.Scene(tris, mats)
    nodes, root = sc.read_bvh()
    ref = oracle_scene(tris, mats)
    assert root == ref.root == 0
    assert nodes.shape[0] == ref.node_count == 2 * n - 1
    assert np.array_equal(nodes.view(np.uint32), ref.nodes[: 2 * n - 1].view(np.uint32)), \
        mismatch_report("nodes", nodes, ref.nodes[: 2 * n - 1])


def test_bvh_c0_matches_oracle(cuda):
    s = scenes.config0()
    sc = abi.Scene(s.triangles, s.materials)
    nodes, _ = sc.read_bvh()
    ref = oracle_scene(s.triangles, s.materials)
    assert np.array_equal(nodes.view(np.uint32), ref.nodes.view(np.uint32)), mismatch_report("nodes", nodes, ref.nodes)


def test_trace_rays_match_oracle(cuda, render_mod):
    tris, mats, tex, cam, params = dense_plant()
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    rng = np.random.default_rng(3)
    lo = tris[:, 0:3].min(0)
    hi = tris[:, 0:3].max(0)
    n = 4000
    o = rng.uniform(lo - 2, hi + 2, (n, 3)).astype(np.float32)
    target = rng.uniform(lo, hi, (n, 3))
    d = (target - o).astype(np.float32)
    got = render_mod.trace_rays(sc, o, d)
    want = np.zeros_like(got)
    for i in range(n):
        h = ref.trace(o[i], d[i])
        want[i] = [np.float32(0).view(np.float32), h.t, h.normal.x, h.normal.y, h.normal.z, h.uv.x, h.uv.y, 0.0]
        want[i, 0] = np.array([h.hit], np.int32).view(np.float32)[0]
        want[i, 7] = np.array([h.triangleIndex if h.hit else 0], np.uint32).view(np.float32)[0]
        if not h.hit:
            want[i, 1:7] = 0.0
    hit = got[:, 0].view(np.int32)
    assert hit.sum() > n // 4, "test rays should mostly hit"
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), mismatch_report("hits", got, want)


def compare_accumulators(gpu, acc, label):
    counts = gpu["counts"]
    assert np.array_equal(counts, acc.counts), f"{label}: sample counts differ at {np.nonzero(counts != acc.counts)[0][:8]}"
    assert np.array_equal(gpu["conv"], acc.conv), f"{label}: converged flags differ"
    for name, g, w in (("acc", gpu["acc"], acc.acc), ("mean", gpu["mean"], acc.mean), ("m2", gpu["m2"], acc.m2)):
        exact = np.array_equal(np.asarray(g).view(np.uint32), np.asarray(w).view(np.uint32))
        if not exact:
            rel = np.abs(g - w) / np.maximum(np.abs(w), 1e-6)
            frac = float((np.asarray(g).view(np.uint32) != np.asarray(w).view(np.uint32)).mean())
            assert rel.max() <= REL_TOL, f"{label}/{name}: max rel err {rel.max():.3g}, {frac:.4%} words differ; " \
                                         + mismatch_report(name, g, w)


@pytest.mark.parametrize("kind", ["zoo", "plain"])
def test_render_fixed_spp_matches_oracle(cuda, render_mod, kind):
    w = h = 48
    tris, mats, tex, cam, params = dense_plant(w, h, spp=8, materials=kind)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    acc_o, passes_o = O.render_adaptive(ref, cam, params, w, h, threads=8)
    dev = render_mod.Accumulator.allocate(w, h)
    inputs = render_mod.DeviceInputs(cam, params)
    passes = render_mod.render_loop(sc, dev, inputs)
    assert passes == passes_o
    compare_accumulators(dev.host(), acc_o, f"loop/{kind}")


def test_render_adaptive_matches_oracle(cuda, render_mod):
    w = h = 40
    tris, mats, tex, cam, params = dense_plant(w, h, spp=64, min_spp=4, threshold=0.08)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    acc_o, passes_o = O.render_adaptive(ref, cam, params, w, h, threads=8)
    assert len(np.unique(acc_o.counts)) >= 2, "threshold should make the counts adaptive"
    dev = render_mod.Accumulator.allocate(w, h)
    inputs = render_mod.DeviceInputs(cam, params)
    passes = render_mod.render_loop(sc, dev, inputs)
    assert passes == passes_o
    compare_accumulators(dev.host(), acc_o, "adaptive")


def test_persistent_passes_equal_loop(cuda, render_mod):
    w = h = 40
    tris, mats, tex, cam, params = dense_plant(w, h, spp=32, min_spp=4, threshold=0.08)
    sc = abi.Scene(tris, mats, tex)
    inputs = render_mod.DeviceInputs(cam, params)
    a = render_mod.Accumulator.allocate(w, h)
    render_mod.render_loop(sc, a, inputs)
    b = render_mod.Accumulator.allocate(w, h)
    render_mod.render_persistent(sc, b, inputs)
    ha, hb = a.host(), b.host()
    for k in ("acc", "counts", "mean", "m2", "conv"):
        assert np.array_equal(np.asarray(ha[k]).view(np.uint8), np.asarray(hb[k]).view(np.uint8)), k


@pytest.mark.parametrize("kind,spp,min_spp,thr,max_paths", [
    ("zoo", 8, None, 0.001, 0), ("plain", 8, None, 0.001, 0), ("zoo", 64, 4, 0.08, 0),
    ("zoo", 37, 5, 0.08, 0), ("zoo", 64, 4, 0.08, 3000)])
def test_wavefront_matches_oracle(cuda, render_mod, kind, spp, min_spp, thr, max_paths):
    """Wavefront schedule (grouped samples, speculative groups when few pixels remain, maxSpp
    not a multiple of 4, tiny path budgets) against the oracle's pass-by-pass schedule."""
    w = h = 40
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, min_spp=min_spp, threshold=thr, materials=kind)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    acc_o, _ = O.render_adaptive(ref, cam, params, w, h, threads=8)
    dev = render_mod.Accumulator.allocate(w, h)
    render_mod.render_wavefront(sc, dev, render_mod.DeviceInputs(cam, params), max_paths=max_paths)
    compare_accumulators(dev.host(), acc_o, f"wavefront/{kind}/{spp}/{max_paths}")


def test_wavefront_c0_matches_oracle(cuda, render_mod):
    s = scenes.config0()
    sc = abi.Scene(s.triangles, s.materials)
    ref = oracle_scene(s.triangles, s.materials)
    acc_o, _ = O.render_adaptive(ref, s.camera, s.params, s.width, s.height, threads=8)
    dev = render_mod.Accumulator.allocate(s.width, s.height)
    render_mod.render_wavefront(sc, dev, render_mod.DeviceInputs(s.camera, s.params))
    compare_accumulators(dev.host(), acc_o, "wavefront/C0")


def test_thin_lens_matches_oracle(cuda, render_mod):
    w = h = 32
    tris, mats, tex, cam, params = dense_plant(w, h, spp=4, thin_lens=True)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    acc_o, _ = O.render_adaptive(ref, cam, params, w, h, threads=8)
    dev = render_mod.Accumulator.allocate(w, h)
    render_mod.render_loop(sc, dev, render_mod.DeviceInputs(cam, params))
    compare_accumulators(dev.host(), acc_o, "thin-lens")


def test_c0_slice_matches_oracle(cuda, render_mod):
    """The full C0 workload: 256x256, 16 spp, depth 8."""
    s = scenes.config0()
    sc = abi.Scene(s.triangles, s.materials)
    ref = oracle_scene(s.triangles, s.materials)
    acc_o, passes_o = O.render_adaptive(ref, s.camera, s.params, s.width, s.height, threads=8)
    dev = render_mod.Accumulator.allocate(s.width, s.height)
    passes = render_mod.render_loop(sc, dev, render_mod.DeviceInputs(s.camera, s.params))
    assert passes == passes_o == 16
    compare_accumulators(dev.host(), acc_o, "C0")
