"""The hero-wavelength adjoint (pathTracerAdjointRenderSpectral) on the GPU: gradients with
respect to binned reflectance and emission spectra, roughness, metallic, specular and the
emission scale, against the oracle's FP64 spectral adjoint of the same samples (which the CPU
tests pin to FP64 central differences); its argument checks; and autograd through it."""
import ctypes as C

import numpy as np
import pytest
import torch

from lsystemeditor_amd import abi, diff, rgb2spec, scenes
from oracle import oracle as O
from tests.helpers import dense_plant

pytestmark = pytest.mark.gpu


@pytest.fixture()
def hero(cuda):
    table = rgb2spec.table(16)
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    try:
        yield O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0)
    finally:
        abi.Spectral.disable()
        s.close()


def spectra_for(mats, seed):
    rng = np.random.default_rng(seed)
    sp = np.empty((mats.shape[0], 32), np.float32)
    sp[:, :16] = rng.uniform(0.1, 0.9, (mats.shape[0], 16))
    sp[:, 16:] = rng.uniform(0.3, 1.4, (mats.shape[0], 16))
    sp[mats[:, 9] <= 0.0, 16] = -1.0
    return sp


def oracle_grad(tris, mats, sp, cam, params, w, h, spp, gi, first_seed, ospec):
    ref = O.OracleScene(abi.host_to_gpu_triangles(tris), mats)
    ref.set_spectra(sp)
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    p[6:8] = np.array([spp, spp], np.int32).view(np.float32)
    g = np.zeros(diff.SPECTRAL_PARAMS * mats.shape[0])
    gi64 = np.ascontiguousarray(gi, np.float64)  # held: ptr() is a bare address
    O.lib().orc_render_adjoint_spectral(O.ptr(gi64), w, h, O.ptr(cam),
                                        C.byref(ref.scene), C.byref(O.Env()), O.ptr(p), C.byref(ospec.model), spp,
                                        first_seed, O.ptr(g), 8)
    return g.reshape(-1, diff.SPECTRAL_PARAMS)


@pytest.mark.parametrize("seed_base", [0, 5])
def test_spectral_adjoint_matches_oracle_fp64(hero, seed_base):
    w = h = 24
    spp = 8
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, max_depth=6, materials="diff")
    sp = spectra_for(mats, 3)
    sc = abi.Scene(tris, mats, tex)
    r = diff.SpectralDiffRenderer(sc, mats, sp, cam, params, w, h, spp)
    th = r.theta0().cuda()
    img = r.forward(th, seed_base=seed_base)
    assert torch.isfinite(img).all() and float(img.abs().sum()) > 0
    gi = torch.randn((w * h, 3), generator=torch.Generator().manual_seed(7)).cuda()
    got = r.adjoint(gi, reuse_forward=seed_base == 0, seed_base=seed_base).cpu().numpy()
    want = oracle_grad(tris, mats, sp, cam, params, w, h, spp, gi.cpu().numpy(), seed_base + 1, hero)
    live = r.live_mask()
    assert not got[~live].any() and not want[~live].any()
    scale = np.abs(want).max()
    assert scale > 0 and (want[:, 16:32] != 0).sum() >= 4
    err = np.abs(got - want)
    rel = err / np.maximum(np.abs(want), 1e-30)
    print(f"spectral adjoint vs FP64 oracle: max rel err {rel[want != 0].max():.3g}", flush=True)
    assert np.all(err <= 1e-4 * np.abs(want) + 1e-6 * scale), f"max rel err {rel[want != 0].max():.3g}"


def test_spectral_adjoint_reuses_the_forward_radiance(hero):
    w = h = 16
    spp = 4
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, max_depth=5, materials="diff")
    sp = spectra_for(mats, 1)
    sc = abi.Scene(tris, mats, tex)
    r = diff.SpectralDiffRenderer(sc, mats, sp, cam, params, w, h, spp)
    th = r.theta0().cuda()
    r.forward(th)
    gi = torch.randn((w * h, 3), generator=torch.Generator().manual_seed(3)).cuda()
    a = r.adjoint(gi, reuse_forward=True).cpu().numpy()
    b = r.adjoint(gi, reuse_forward=False).cpu().numpy()
    assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(a).max())


def test_spectral_adjoint_argument_checks(cuda):
    lib = abi.lib()
    tris, mats, tex, cam, params = dense_plant(8, 8, spp=2, materials="diff")
    sc = abi.Scene(tris, mats, tex)
    r = diff.DiffRenderer(sc, mats, cam, params, 8, 8, 2)
    g = torch.zeros((64, 3), device="cuda")
    out = torch.zeros((mats.shape[0], diff.SPECTRAL_PARAMS), dtype=torch.float64, device="cuda")
    args = (g.data_ptr(), 8, 8, 2, r.inputs.camera.data_ptr(), sc.device, None, r.inputs.params.data_ptr(),
            mats.shape[0], None, out.data_ptr(), 1, 0)
    abi.Spectral.disable()
    assert not lib.pathTracerAdjointRenderSpectral(*args)
    assert "hero-wavelength rendering only" in abi.last_error()
    bad = list(args)
    bad[8] = 0
    assert not lib.pathTracerAdjointRenderSpectral(*bad)
    assert "invalid arguments" in abi.last_error()


def test_batch_spectral_adjoint_equals_per_scene_adjoints(hero):
    """pathTracerAdjointRenderBatch at hero wavelengths (36 values per row) equals every scene's
    own pathTracerAdjointRenderSpectral over the same seeds, and the batch render equals the
    per-scene renders."""
    from lsystemeditor_amd import inverse
    specs = inverse.c4_specs(count=3, res=24, spp=4, iterations=(3, 4))
    rs = []
    for k, s in enumerate(specs):
        sc = abi.Scene(s.triangles, s.materials, bvh="sah")
        rs.append(diff.SpectralDiffRenderer(sc, s.materials, spectra_for(s.materials, 10 + k), s.camera, s.params,
                                            s.width, s.height, s.spp))
    br = diff.BatchDiffRenderer(rs)
    thetas = [r.theta0() for r in rs]
    img = br.forward(thetas, seed_base=6)
    g = torch.randn((img.shape[0], 3), generator=torch.Generator().manual_seed(2)).cuda()
    got = br.adjoint(g, seed_base=6)
    for i, r in enumerate(rs):
        own = r.forward(thetas[i].cuda(), seed_base=6)
        assert torch.equal(own, img[br.pixels(i)]), i
        want = r.adjoint(g[br.pixels(i)].contiguous(), reuse_forward=False, seed_base=6)
        scale = float(want.abs().max())
        assert scale > 0, i
        assert float((got[i] - want).abs().max()) <= 1e-9 * scale, (i, float((got[i] - want).abs().max()) / scale)
