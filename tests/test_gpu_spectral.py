"""GPU parity of hero-wavelength rendering: the launchers (per-pass megakernel, preview) and the
wavefront scheduler against the oracle's spectral integrator on identical inputs, bit for bit.

The rgb2spec table is the project's own (lsystemeditor_amd.rgb2spec, the reference's srgb.coeff
is missing); its values are parity-unpinned against the reference's file, but the fetch, the
sigmoid, the white normalisation and the integrator are checked exactly here and against the
reference's Rgb2Spec.h in tests/test_oracle_spectral.py.
"""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi, rgb2spec, scenes
from oracle import oracle as O
from tests.helpers import dense_plant, mismatch_report
from tests.test_gpu_parity import compare_accumulators, oracle_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R(cuda):
    from lsystemeditor_amd import render
    return render


@pytest.fixture(scope="module")
def table():
    return rgb2spec.table(16)


@pytest.fixture()
def spectral(cuda, table):
    """Hero-wavelength sampling on (full 360..830 nm range) for one test, RGB again afterwards."""
    s = abi.Spectral(*table)
    s.enable()
    try:
        yield s
    finally:
        abi.Spectral.disable()
        s.close()


def test_device_white_norm_equals_oracle(spectral, table):
    ospec = O.OracleSpectral(*table)
    got = np.array(spectral.white_norm(), np.float32)
    assert np.array_equal(got.view(np.uint32), np.array(ospec.model.white, np.float32).view(np.uint32))
    assert abi.lib().pathTracerSpectralSampling() == 4


def test_coeff_file_loader(cuda, table, tmp_path):
    p = tmp_path / "table.coeff"
    rgb2spec.write_coeff_file(p, *table)
    h = abi.lib().ptSpectralLoadCoeffFile(str(p).encode())
    assert h
    try:
        st = abi.SpectralStateGpu()
        assert abi.lib().ptSpectralGetState(h, C.byref(st))
        assert st.rgb2specRes == 16
        ospec = O.OracleSpectral(*table)
        assert np.float32(st.whiteNormG) == np.float32(ospec.model.white[1])
    finally:
        abi.lib().ptSpectralDestroy(h)
    assert not abi.lib().ptSpectralLoadCoeffFile(str(tmp_path / "missing.coeff").encode())


def test_spectral_sampling_validation(cuda):
    lib = abi.lib()
    abi.Spectral.disable()
    assert not lib.pathTracerSetSpectralSampling(3, 360.0, 830.0)
    assert not lib.pathTracerSetSpectralSampling(4, 350.0, 830.0)
    assert not lib.pathTracerSetSpectralSampling(4, 500.0, 500.0)
    assert lib.pathTracerSpectralSampling() == 0


@pytest.mark.parametrize("kind", ["zoo", "plain"])
def test_spectral_loop_matches_oracle(R, spectral, table, kind):
    w = h = 32
    tris, mats, tex, cam, params = dense_plant(w, h, spp=4, materials=kind)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    with O.OracleSpectral(*table):
        acc_o, passes_o = O.render_adaptive(ref, cam, params, w, h, threads=8)
    dev = R.Accumulator.allocate(w, h)
    passes = R.render_loop(sc, dev, R.DeviceInputs(cam, params))
    assert passes == passes_o
    compare_accumulators(dev.host(), acc_o, f"spectral loop/{kind}")
    assert float(np.abs(dev.host()["acc"][:, :3]).sum()) > 0.0


@pytest.mark.parametrize("kind,spp,min_spp,thr", [("zoo", 8, 8, 0.001), ("plain", 24, 4, 0.08)])
def test_spectral_wavefront_matches_oracle(R, cuda, table, kind, spp, min_spp, thr):
    """380..720 nm hero range; fixed and adaptive counts."""
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    try:
        w = h = 36
        tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, min_spp=min_spp, threshold=thr, materials=kind)
        sc = abi.Scene(tris, mats, tex)
        ref = oracle_scene(tris, mats, tex)
        with O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0):
            acc_o, _ = O.render_adaptive(ref, cam, params, w, h, threads=8)
        dev = R.Accumulator.allocate(w, h)
        R.render_wavefront(sc, dev, R.DeviceInputs(cam, params))
        compare_accumulators(dev.host(), acc_o, f"spectral wavefront/{kind}/{spp}")
    finally:
        abi.Spectral.disable()
        s.close()


def test_spectral_preview_with_hdr_env_matches_oracle(R, spectral, table):
    from tests.test_gpu_launchers import hdr_image
    w, h = 33, 27
    tris, mats, tex, cam, params = dense_plant(w, h, spp=4)
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    img = hdr_image()
    oenv = O.OracleEnv(img)
    inputs = R.DeviceInputs(cam, params, env=abi.Environment(img))
    got = R.preview(sc, inputs, w, h, seed=5).cpu().numpy()
    want = np.zeros((w * h, 4), np.float32)
    with O.OracleSpectral(*table):
        O.lib().orc_preview_sample(O.ptr(want), w, h, O.ptr(cam), C.byref(ref.scene), C.byref(oenv.env),
                                   O.ptr(params), 5, 8)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), mismatch_report("spectral preview", got, want)


def test_spectral_c0_matches_oracle(R, cuda, table):
    """C0 as configured: 256x256, 16 spp, depth 8, four hero wavelengths in 380..720 nm."""
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    try:
        c0 = scenes.config0()
        sc = abi.Scene(c0.triangles, c0.materials)
        ref = oracle_scene(c0.triangles, c0.materials)
        with O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0):
            acc_o, passes_o = O.render_adaptive(ref, c0.camera, c0.params, c0.width, c0.height, threads=8)
        dev = R.Accumulator.allocate(c0.width, c0.height)
        R.render_wavefront(sc, dev, R.DeviceInputs(c0.camera, c0.params))
        compare_accumulators(dev.host(), acc_o, "spectral C0")
    finally:
        abi.Spectral.disable()
        s.close()


def test_adjoint_rejects_spectral(R, cuda, spectral):
    from lsystemeditor_amd import diff
    w = h = 8
    tris, mats, tex, cam, params = dense_plant(w, h, spp=2, materials="diff")
    sc = abi.Scene(tris, mats, tex)
    r = diff.DiffRenderer(sc, mats, cam, params, w, h, 2)
    with pytest.raises(RuntimeError, match="RGB rendering only"):
        r.adjoint(cuda.zeros((w * h, 3)))


def binned_table(mats: np.ndarray, seed: int) -> np.ndarray:
    """Random binned spectra: every other material keeps its rgb2spec reflectance, emissive
    materials get a binned emission shape (the plain-colour ones keep theirs on odd seeds)."""
    rng = np.random.default_rng(seed)
    n = mats.shape[0]
    sp = np.empty((n, abi.SPECTRUM_WORDS), np.float32)
    sp[:, :16] = rng.uniform(0.05, 0.95, (n, 16))
    sp[:, 16:] = rng.uniform(0.2, 1.5, (n, 16))
    sp[1::2, 0] = -1.0
    emissive = mats[:, 9] > 0.0
    sp[~emissive, 16] = -1.0
    if seed % 2:
        sp[emissive, 16] = -1.0
    return sp


@pytest.mark.parametrize("kind,mode,seed", [("zoo", 0, 0), ("zoo", 1, 2), ("plain", 1, 1), ("diff", 0, 4)])
def test_binned_spectra_match_oracle(R, cuda, table, kind, mode, seed):
    """Materials with 16-bin reflectance and emission spectra (ptSceneSetSpectra): the per-pass
    launcher and the wavefront scheduler against the oracle's spectral integrator, bit for bit,
    in both light sampling modes (mode 1 evaluates next-event light samples' binned emission)."""
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    O.lib().orc_set_light_sampling(mode)
    try:
        with abi.LightSampling(mode):
            w = h = 32
            tris, mats, tex, cam, params = dense_plant(w, h, spp=8, min_spp=4, threshold=0.05, materials=kind)
            sp = binned_table(mats, seed)
            sc = abi.Scene(tris, mats, tex)
            sc.set_spectra(sp)
            ref = oracle_scene(tris, mats, tex)
            ref.set_spectra(sp)
            with O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0):
                acc_o, passes_o = O.render_adaptive(ref, cam, params, w, h, threads=8)
            dev = R.Accumulator.allocate(w, h)
            passes = R.render_loop(sc, dev, R.DeviceInputs(cam, params))
            assert passes == passes_o
            compare_accumulators(dev.host(), acc_o, f"binned loop/{kind}/{mode}")
            dev2 = R.Accumulator.allocate(w, h)
            R.render_wavefront(sc, dev2, R.DeviceInputs(cam, params))
            compare_accumulators(dev2.host(), acc_o, f"binned wavefront/{kind}/{mode}")
            # the spectra change the image: without them it differs
            sc.set_spectra(None)
            dev3 = R.Accumulator.allocate(w, h)
            R.render_wavefront(sc, dev3, R.DeviceInputs(cam, params))
            assert not np.array_equal(dev3.host()["acc"], dev2.host()["acc"])
    finally:
        O.lib().orc_set_light_sampling(0)
        abi.Spectral.disable()
        s.close()


def test_set_spectra_validation(cuda):
    tris, mats, tex, cam, params = dense_plant(8, 8, spp=2)
    sc = abi.Scene(tris, mats, tex)
    lib = abi.lib()
    bad = np.zeros((mats.shape[0] + 1, abi.SPECTRUM_WORDS), np.float32)
    assert not lib.ptSceneSetSpectra(sc.handle, bad.ctypes.data, bad.shape[0], 0)
    assert "one spectrum record per material" in abi.last_error()
    assert not lib.ptSceneSetSpectra(None, bad.ctypes.data, 1, 0)
    sc.set_spectra(np.zeros((mats.shape[0], abi.SPECTRUM_WORDS), np.float32))
    desc = sc.description()
    assert int.from_bytes(desc[128:136], "little") != 0
    sc.set_spectra(None)
    assert int.from_bytes(sc.description()[128:136], "little") == 0
