"""Batched inverse rendering at the C4 scene size (256x256, 16 spp, depth 8; two scenes, 50 Adam
steps at lr 1e-2), with 16-bin base-colour spectra and with RGB base colours: the base colours
(for spectra, their projected colours), which dominate the images, move toward their targets; the
finite-difference spot checks agree with the adjoint along the trajectory; the loss of 256-spp
images falls most of the way to its value at the target parameters. The step losses are not
asserted: each step renders fresh samples, so they sit on the noise floor of two 16-spp images.

Roughness and metallic are nearly degenerate with the base colours in these images (metallic
starts within 0.01 of its target), so their errors are reported by the optimiser but not
asserted."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binned_colour", [True, False])
def test_inverse_rendering_recovers_base_colours(cuda, binned_colour):
    from lsystemeditor_amd import inverse
    specs = inverse.c4_specs(count=2, res=256, spp=16, iterations=(3, 4))
    prob = inverse.InverseProblem(specs, binned_colour=binned_colour)
    res = inverse.optimise(prob, steps=50, lr=1e-2, check_every=25)
    e0, e1 = res["param_mae_initial"], res["param_mae_final"]
    print(res)
    assert e1["base_color"] < 0.6 * e0["base_color"], res
    img = res["image_loss"]
    assert img["final"] < img["initial"], img
    assert img["final_excess_fraction"] < 0.3, img
    assert res["target_spp"] == 256
    assert len(res["fd_checks"]) == 2
    for c in res["fd_checks"]:
        assert c["median_rel_err"] < 0.05, c
    assert res["grad_samples_per_s"] > 0.0
    assert prob.parameter_count() == 2 * (54 if binned_colour else 15)


def hero_window_check(problem, h=1e-4, window=12):
    """Scene 0's GPU spectral adjoint of its L2 loss on a centred window against central
    differences of the oracle's FP64 spectral detached estimator, sampled at the current
    parameters (the FD checker of the hero-wavelength problem)."""
    import ctypes as C
    import numpy as np
    import torch
    from lsystemeditor_amd import abi, diff, rgb2spec
    from lsystemeditor_amd.inverse import HERO_RANGE
    from oracle import oracle as O
    r, p, spec = problem.renderers[0], problem.params[0], problem.specs[0]
    w, hh, spp = r.width, r.height, r.spp
    th = p.theta().detach()
    img = r.forward(th.cuda())
    win = torch.zeros_like(img).view(hh, w, 3)
    y0, x0 = hh // 2 - window // 2, w // 2 - window // 2
    win[y0:y0 + window, x0:x0 + window] = (2.0 * (img - problem.targets[0])).view(hh, w, 3)[y0:y0 + window,
                                                                                               x0:x0 + window]
    win = win.reshape(-1, 3).contiguous()
    g = r.adjoint(win).cpu().numpy()
    ref = O.OracleScene(abi.host_to_gpu_triangles(spec.triangles), spec.materials)
    ospec = O.OracleSpectral(*rgb2spec.table(16), lambda_min=HERO_RANGE[0], lambda_max=HERO_RANGE[1])
    mats, sp = r._tables(th)
    np.copyto(ref.mats, mats.reshape(ref.mats.shape))
    ref.set_spectra(sp)
    gi = np.ascontiguousarray(win.cpu().numpy().astype(np.float64))

    def loss(m, s_):
        im = np.zeros((w * hh, 3))
        O.lib().orc_render_detached_spectral(None, O.ptr(im), O.ptr(gi), w, hh, O.ptr(spec.camera),
                                             C.byref(ref.scene), O.ptr(m), O.ptr(s_), C.byref(O.Env()),
                                             O.ptr(r.params_host), C.byref(ospec.model), spp, 1, 8)
        return float((gi * im).sum())

    fds = {}
    for i, k in list(zip(*np.nonzero(p.mask.numpy())))[::3]:
        if k < abi.SPECTRUM_WORDS:
            up, dn = sp.copy(), sp.copy()
            up[i, k] += np.float32(h)
            dn[i, k] -= np.float32(h)
            fds[i, k] = (loss(mats, up) - loss(mats, dn)) / (float(up[i, k]) - float(dn[i, k]))
        else:
            wd = diff.SPECTRAL_SCALAR_WORDS[k - abi.SPECTRUM_WORDS]
            up, dn = mats.copy(), mats.copy()
            up[i, wd] += np.float32(h)
            dn[i, wd] -= np.float32(h)
            fds[i, k] = (loss(up, sp) - loss(dn, sp)) / (float(up[i, wd]) - float(dn[i, wd]))
    top = max(abs(v) for v in fds.values())
    rel = np.array([abs(g[i, k] - v) / max(abs(v), 1e-3 * top) for (i, k), v in fds.items()])
    return {"params": int(rel.size), "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel))}


def test_hero_wavelength_inverse_rendering(cuda):
    """The hero-wavelength problem (16-bin reflectance spectra rendered and differentiated at four
    wavelengths, batched spectral adjoint; 2 scenes at 64x64, 16 spp, 40 steps): the image loss
    falls toward its floor, the projected colours move toward their targets, and the spectral
    gradient agrees with FP64 central differences along the trajectory. The bins themselves are
    not asserted: the images see them mostly through their projections."""
    from lsystemeditor_amd import abi, inverse, rgb2spec
    hero = abi.Spectral(*rgb2spec.table(16))
    hero.enable(*inverse.HERO_RANGE)
    try:
        specs = inverse.c4_specs(count=2, res=64, spp=16, iterations=(3, 4))
        prob = inverse.InverseProblem(specs, transport="hero", fd_checker=hero_window_check)
        res = inverse.optimise(prob, steps=40, lr=1e-2, check_every=20)
        e0, e1 = res["param_mae_initial"], res["param_mae_final"]
        print({k: v for k, v in res.items() if k != "fd_checks"}, res["fd_checks"])
        assert e1["base_color"] < 0.8 * e0["base_color"], (e0, e1)
        img = res["image_loss"]
        assert img["final"] < img["initial"], img
        assert len(res["fd_checks"]) == 2
        for c in res["fd_checks"]:
            assert c["params"] >= 10 and c["max_rel_err"] < 1e-4, c
        assert prob.parameter_count() == 2 * 3 * 18
    finally:
        abi.Spectral.disable()
        hero.close()
