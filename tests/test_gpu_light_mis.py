"""GPU parity of light sampling mode 1 (multiple importance sampling on area lights): every forward
scheduler (per-pass launcher, persistent megakernel, wavefront, batch wavefront), RGB and
hero-wavelength, on the median and binned-SAH trees, bit for bit against the oracle in the same
mode. The oracle's mode 1 is pinned by tests/test_oracle_light_mis.py (the direct-lighting
integral, the area-to-solid-angle pdf, the path-length boundary, equal expectation with mode 0)."""
import numpy as np
import pytest

from lsystemeditor_amd import abi, rgb2spec, scenes
from oracle import oracle as O
from tests.helpers import dense_plant
from tests.test_gpu_parity import compare_accumulators, oracle_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R(cuda):
    from lsystemeditor_amd import render
    return render


@pytest.fixture()
def mis(cuda):
    O.lib().orc_set_light_sampling(1)
    with abi.LightSampling(1):
        yield
    O.lib().orc_set_light_sampling(0)


@pytest.mark.parametrize("kind,bvh", [("zoo", "median"), ("plain", "median"), ("plain", "sah")])
def test_mis_loop_and_wavefront_match_oracle(R, mis, kind, bvh):
    w = h = 32
    tris, mats, tex, cam, params = dense_plant(w, h, spp=12, min_spp=4, threshold=0.08, materials=kind)
    sc = abi.Scene(tris, mats, tex, bvh=bvh)
    ref = oracle_scene(tris, mats, tex)
    acc_o, passes_o = O.render_adaptive(ref, cam, params, w, h, threads=8)
    inputs = R.DeviceInputs(cam, params)
    dev = R.Accumulator.allocate(w, h)
    assert R.render_loop(sc, dev, inputs) == passes_o
    compare_accumulators(dev.host(), acc_o, f"mis loop/{kind}/{bvh}")
    for max_paths in (0, 3000):
        dev = R.Accumulator.allocate(w, h)
        R.render_wavefront(sc, dev, inputs, max_paths=max_paths)
        compare_accumulators(dev.host(), acc_o, f"mis wavefront/{kind}/{bvh}/{max_paths}")
    dev = R.Accumulator.allocate(w, h)
    R.render_persistent(sc, dev, inputs)
    compare_accumulators(dev.host(), acc_o, f"mis persistent/{kind}/{bvh}")


def test_mis_changes_the_image_but_not_its_counts_structure(R, cuda):
    """Mode 1 differs from mode 0 sample by sample (next-event light is now seen), with the same
    per-pixel schedule under fixed spp."""
    w = h = 24
    tris, mats, tex, cam, params = dense_plant(w, h, spp=8, materials="plain")
    sc = abi.Scene(tris, mats, tex)
    inputs = R.DeviceInputs(cam, params)
    a = R.Accumulator.allocate(w, h)
    R.render_wavefront(sc, a, inputs)
    with abi.LightSampling(1):
        b = R.Accumulator.allocate(w, h)
        R.render_wavefront(sc, b, inputs)
    ha, hb = a.host(), b.host()
    assert np.array_equal(ha["counts"], hb["counts"])
    assert not np.array_equal(ha["acc"], hb["acc"])


def test_mis_spectral_wavefront_matches_oracle(R, mis):
    table = rgb2spec.table(16)
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    try:
        w = h = 28
        tris, mats, tex, cam, params = dense_plant(w, h, spp=8, materials="zoo")
        sc = abi.Scene(tris, mats, tex, bvh="sah")
        ref = oracle_scene(tris, mats, tex)
        with O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0):
            acc_o, _ = O.render_adaptive(ref, cam, params, w, h, threads=8)
        dev = R.Accumulator.allocate(w, h)
        R.render_wavefront(sc, dev, R.DeviceInputs(cam, params))
        compare_accumulators(dev.host(), acc_o, "mis spectral wavefront")
        dev = R.Accumulator.allocate(w, h)
        R.render_loop(sc, dev, R.DeviceInputs(cam, params))
        compare_accumulators(dev.host(), acc_o, "mis spectral loop")
    finally:
        abi.Spectral.disable()
        s.close()


def test_mis_batch_equals_single_renders(R, mis):
    from tests.test_gpu_scene_batch import _batch, _part, _same, _scenes, _single
    items = _scenes()
    b = _batch(items)
    b.render_fixed(4)
    got = b.acc.host()
    for i, it in enumerate(items):
        _same(_part(got, b.pixels(i)), _single(it, 4), f"mis batch scene {i}")
    tris, mats, tex, cam, params = items[0][3]
    acc_o, _ = O.render_adaptive(oracle_scene(tris, mats, tex), cam, params, 24, 20, threads=8)
    compare_accumulators(_part(got, b.pixels(0)), acc_o, "mis batch scene 0 / oracle")
    b.close()


def test_mis_c0_matches_oracle(R, mis):
    c0 = scenes.config0()
    sc = abi.Scene(c0.triangles, c0.materials, bvh="sah")
    ref = oracle_scene(c0.triangles, c0.materials)
    acc_o, _ = O.render_adaptive(ref, c0.camera, c0.params, c0.width, c0.height, threads=8)
    dev = R.Accumulator.allocate(c0.width, c0.height)
    R.render_wavefront(sc, dev, R.DeviceInputs(c0.camera, c0.params))
    compare_accumulators(dev.host(), acc_o, "mis C0")


def test_adjoint_rejects_mis(R, mis, cuda):
    from lsystemeditor_amd import diff
    w = h = 8
    tris, mats, tex, cam, params = dense_plant(w, h, spp=2, materials="diff")
    sc = abi.Scene(tris, mats, tex)
    r = diff.DiffRenderer(sc, mats, cam, params, w, h, 2)
    with pytest.raises(RuntimeError, match="reference light sampling only"):
        r.adjoint(cuda.zeros((w * h, 3), device="cuda"))
