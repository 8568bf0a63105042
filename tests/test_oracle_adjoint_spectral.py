"""CPU pins of the oracle's differentiable hero-wavelength renderer (binned spectra): the spectral
detached estimator reduces to the hero-wavelength forward at the primal parameters (bit for bit),
its double evaluation equals its float one, and the FP64 adjoint of every parameter -- the 16
reflectance and 16 emission bins, roughness, metallic, specular and the emission scale -- matches
central differences (h = 1e-4) of the double evaluation, which carry no float rounding."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi, rgb2spec
from oracle import oracle as O
from tests.helpers import dense_plant

WORD = {"roughness": 3, "metallic": 4, "specular": 14, "emission": 9}
SCALARS = ["roughness", "metallic", "specular", "emission"]
NP = 36


def setup(kind="diff", w=12, h=12, spp=4, seed=0):
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, max_depth=5, materials=kind)
    sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats, tex)
    rng = np.random.default_rng(seed)
    sp = np.empty((mats.shape[0], 32), np.float32)
    sp[:, :16] = rng.uniform(0.1, 0.9, (mats.shape[0], 16))
    sp[:, 16:] = rng.uniform(0.3, 1.4, (mats.shape[0], 16))
    sp[mats[:, 9] <= 0.0, 16] = -1.0
    sc.set_spectra(sp)
    spec = O.OracleSpectral(*rgb2spec.table(16), lambda_min=380.0, lambda_max=720.0)
    return sc, sp, cam, params, spec


def detached(sc, cam, params, spec, w, h, spp, mats=None, spectra=None, f64=False, mask=None, seed=1):
    if f64:
        img = np.zeros((w * h, 3))
        O.lib().orc_render_detached_spectral(None, O.ptr(img), None if mask is None else O.ptr(mask), w, h,
                                             O.ptr(cam), C.byref(sc.scene), None if mats is None else O.ptr(mats),
                                             None if spectra is None else O.ptr(spectra), C.byref(O.Env()),
                                             O.ptr(params), C.byref(spec.model), spp, seed, 8)
        return img
    img = np.zeros((w * h, 3), np.float32)
    O.lib().orc_render_detached_spectral(O.ptr(img), None, None, w, h, O.ptr(cam), C.byref(sc.scene),
                                         None if mats is None else O.ptr(mats),
                                         None if spectra is None else O.ptr(spectra), C.byref(O.Env()),
                                         O.ptr(params), C.byref(spec.model), spp, seed, 8)
    return img


@pytest.mark.parametrize("kind", ["diff", "zoo"])
def test_spectral_detached_estimator_is_the_hero_forward(kind):
    w = h = 14
    spp = 4
    sc, sp, cam, params, spec = setup(kind, w, h, spp)
    p = params.copy()
    p[6:8] = np.array([spp, spp], np.int32).view(np.float32)
    with spec:
        acc, _ = O.render_adaptive(sc, cam, p, w, h, threads=8)
    img = detached(sc, cam, params, spec, w, h, spp)
    assert np.abs(img).max() > 0
    assert np.array_equal(img.view(np.uint32), acc.acc[:, :3].copy().view(np.uint32))


def test_spectral_f64_estimator_is_the_float_estimator():
    w = h = 12
    spp = 4
    sc, sp, cam, params, spec = setup("diff", w, h, spp)
    em = np.ascontiguousarray(sc.mats.copy())
    em[:, 3] *= np.float32(1.05)
    es = np.ascontiguousarray(sp * np.float32(0.95))
    for mats, spectra in ((None, None), (em, es)):
        a = detached(sc, cam, params, spec, w, h, spp, mats, spectra)
        b = detached(sc, cam, params, spec, w, h, spp, mats, spectra, f64=True)
        assert np.all(np.abs(b - a) <= 2e-5 * np.maximum(np.abs(b), 1e-3 * np.abs(b).max())), \
            float(np.max(np.abs(b - a)))


@pytest.mark.parametrize("seed", [1, 29])
def test_spectral_adjoint_matches_fp64_differences(seed):
    # 16 spp so that BSDF-sampled paths reach the light (the only way emission counts in light
    # sampling mode 0) at enough wavelengths to give most emission bins a gradient
    w = h = 16
    spp = 16
    sc, sp, cam, params, spec = setup("diff", w, h, spp, seed=seed)
    rng = np.random.default_rng(seed)
    gi = np.ascontiguousarray(rng.normal(size=(w * h, 3)))
    grad = np.zeros(NP * sc.mats.shape[0])
    O.lib().orc_render_adjoint_spectral(O.ptr(gi), w, h, O.ptr(cam), C.byref(sc.scene), C.byref(O.Env()),
                                        O.ptr(params), C.byref(spec.model), spp, seed, O.ptr(grad), 8)
    grad = grad.reshape(-1, NP)
    scale = np.abs(grad).max()
    assert scale > 0

    def loss(mats=None, spectra=None):
        return float((gi * detached(sc, cam, params, spec, w, h, spp, mats, spectra, f64=True, seed=seed)).sum())

    checked = {"refl": 0, "emis": 0, "scalar": 0}
    for mi in range(sc.mats.shape[0]):
        cols = list(range(16)) + (list(range(16, 32)) if sp[mi, 16] >= 0 else [])
        for c in cols:
            up, dn = sp.copy(), sp.copy()
            up[mi, c] += np.float32(1e-4)
            dn[mi, c] -= np.float32(1e-4)
            fd = (loss(spectra=up) - loss(spectra=dn)) / (float(up[mi, c]) - float(dn[mi, c]))
            assert abs(grad[mi, c] - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale), (mi, c, grad[mi, c], fd)
            checked["refl" if c < 16 else "emis"] += abs(fd) > 0
        for q, name in enumerate(SCALARS):
            v = float(sc.mats[mi, WORD[name]])
            if name == "emission" and (v <= 0.0 or sp[mi, 16] < 0):
                assert grad[mi, 32 + q] == 0.0
                continue
            up, dn = sc.mats.copy(), sc.mats.copy()
            up[mi, WORD[name]] += np.float32(1e-4)
            dn[mi, WORD[name]] -= np.float32(1e-4)
            fd = (loss(mats=np.ascontiguousarray(up)) - loss(mats=np.ascontiguousarray(dn))) / \
                (float(up[mi, WORD[name]]) - float(dn[mi, WORD[name]]))
            assert abs(grad[mi, 32 + q] - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale), (mi, name, grad[mi, 32 + q], fd)
            checked["scalar"] += abs(fd) > 0
    assert checked["refl"] >= 40 and checked["emis"] >= 8 and checked["scalar"] >= 10, checked
