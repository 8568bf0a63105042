"""CPU pins of the oracle's differentiable renderer: the detached estimator reduces to the
forward integrator at the primal parameters (bit for bit), the analytic BSDF derivative matches
central differences of orc_brdf_eval, the firefly-clamp VJP matches differences of
orc_clamp_firefly, and the FP64 adjoint gradient of an image loss matches central finite
differences of the detached estimator (h = 1e-3) on a small scene, for the default sample set
(seeds 1..spp) and another one. The same checks without float rounding: the detached estimator
and the BSDF value evaluated in double equal their float forms, and the adjoint and
orc_brdf_grad match central differences of them to 1e-5."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi
from oracle import oracle as O
from tests.helpers import dense_plant

PARAMS = ["r", "g", "b", "roughness", "metallic", "specular", "emission"]
WORD = {"r": 0, "g": 1, "b": 2, "roughness": 3, "metallic": 4, "specular": 14, "emission": 9}


def small_scene(w=10, h=10, spp=4, depth=4):
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, max_depth=depth, materials="diff")
    return tris, mats, cam, params


def test_detached_estimator_is_the_forward_at_primal_parameters():
    w = h = 16
    for kind in ("diff", "zoo"):
        tris, mats, tex, cam, params = dense_plant(w, h, spp=4, materials=kind)
        sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats, tex)
        acc, _ = O.render_adaptive(sc, cam, params, w, h, threads=8)
        img = np.zeros((w * h, 3), np.float32)
        O.lib().orc_render_detached(O.ptr(img), w, h, O.ptr(cam), C.byref(sc.scene), None, C.byref(O.Env()),
                                    O.ptr(params), 4, 8)
        assert np.array_equal(img.view(np.uint32), acc.acc[:, :3].copy().view(np.uint32)), kind
        # evaluation table equal to the scene's: also identical
        mats_copy = np.ascontiguousarray(sc.mats.copy())
        img2 = np.zeros_like(img)
        O.lib().orc_render_detached(O.ptr(img2), w, h, O.ptr(cam), C.byref(sc.scene), O.ptr(mats_copy),
                                    C.byref(O.Env()), O.ptr(params), 4, 8)
        if kind == "diff":   # (subsurface paths take the factor form when a table is given)
            assert np.array_equal(img2.view(np.uint32), img.view(np.uint32))


@pytest.mark.parametrize("seed", range(6))
def test_brdf_gradient_matches_central_differences(seed):
    rng = np.random.default_rng(seed)
    L = O.lib()
    base = dict(r=rng.uniform(0.2, 0.8), g=rng.uniform(0.2, 0.8), b=rng.uniform(0.2, 0.8),
                roughness=rng.uniform(0.2, 0.8), metallic=rng.uniform(0.2, 0.8), specular=rng.uniform(0.2, 0.8))
    words = abi.pack_material(base["r"], base["g"], base["b"], roughness=base["roughness"],
                              metallic=base["metallic"], specular=base["specular"])
    for trial in range(20):
        wo = rng.normal(size=3)
        wo[2] = abs(wo[2]) + 0.2
        wi = rng.normal(size=3)
        wi[2] = abs(wi[2]) + 0.2
        wo /= np.linalg.norm(wo)
        wi /= np.linalg.norm(wi)
        a = rng.uniform(0.5, 1.5, 3)
        grad = np.zeros(6)
        m = O.Material.from_buffer_copy(words.tobytes())
        c = O.ctx(m, wo=wo)
        a = np.ascontiguousarray(a)
        L.orc_brdf_grad(C.byref(c), O.v3(wi), O.ptr(a), O.ptr(words), O.ptr(grad))
        for k, name in enumerate(PARAMS[:6]):
            def fd_at(hstep):
                vals = []
                for sgn in (+1, -1):
                    ww = words.copy()
                    ww[WORD[name]] += sgn * hstep
                    f = L.orc_brdf_eval(C.byref(O.ctx(O.Material.from_buffer_copy(ww.tobytes()), wo=wo)), O.v3(wi))
                    vals.append(a[0] * f.x + a[1] * f.y + a[2] * f.z)
                return (vals[0] - vals[1]) / (2 * hstep)
            # float rounding of sharp specular peaks makes the difference quotient itself uncertain:
            # its spread between two step sizes is part of the tolerance
            fd, fd2 = fd_at(1e-3), fd_at(2e-3)
            tol = 2e-3 * max(abs(fd), 1e-2) + 2.0 * abs(fd - fd2)
            assert abs(grad[k] - fd) <= tol, (name, grad[k], fd, fd2)


def test_clamp_vjp_matches_differences():
    L = O.lib()
    rng = np.random.default_rng(3)
    for C3 in ([0.5, 0.2, 0.1], [300.0, 80.0, 40.0], [-0.2, 2.0, 1.0], [150.0, -3.0, 500.0]):
        g = rng.normal(size=3)
        out = np.zeros(3)
        L.orc_clamp_vjp(O.v3(C3), O.ptr(g), O.ptr(out))
        for k in range(3):
            hstep = 1e-2 * max(1.0, abs(C3[k]))
            up, dn = list(C3), list(C3)
            up[k] += hstep
            dn[k] -= hstep
            cu = L.orc_clamp_firefly(O.v3(up))
            cd = L.orc_clamp_firefly(O.v3(dn))
            fd = sum(g[j] * ((cu.x, cu.y, cu.z)[j] - (cd.x, cd.y, cd.z)[j]) for j in range(3)) / (2 * hstep)
            if C3[k] < 0:
                assert out[k] == 0.0
            else:
                assert abs(out[k] - fd) <= 1e-3 * max(1.0, abs(fd)), (C3, k, out[k], fd)


def render(sc, cam, params, w, h, spp, eval_mats=None, first_seed=1):
    img = np.zeros((w * h, 3), np.float32)
    O.lib().orc_render_detached_seeded(O.ptr(img), w, h, O.ptr(cam), C.byref(sc.scene),
                                       None if eval_mats is None else O.ptr(eval_mats), C.byref(O.Env()),
                                       O.ptr(params), spp, first_seed, 8)
    return img


def test_seeded_entry_points_at_seed_one_are_the_plain_ones():
    w = h = 10
    tris, mats, cam, params = small_scene(w, h, 4)
    sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats)
    plain = np.zeros((w * h, 3), np.float32)
    O.lib().orc_render_detached(O.ptr(plain), w, h, O.ptr(cam), C.byref(sc.scene), None, C.byref(O.Env()),
                                O.ptr(params), 4, 8)
    assert np.array_equal(render(sc, cam, params, w, h, 4).view(np.uint32), plain.view(np.uint32))
    other = render(sc, cam, params, w, h, 4, first_seed=5)
    assert not np.array_equal(other, plain)
    gi = np.ascontiguousarray(np.full((w * h, 3), 0.25))
    a = np.zeros(7 * sc.mats.shape[0])
    b = np.zeros_like(a)
    O.lib().orc_render_adjoint(O.ptr(gi), w, h, O.ptr(cam), C.byref(sc.scene), C.byref(O.Env()), O.ptr(params), 4,
                               O.ptr(a), 1)
    O.lib().orc_render_adjoint_seeded(O.ptr(gi), w, h, O.ptr(cam), C.byref(sc.scene), C.byref(O.Env()),
                                      O.ptr(params), 4, 1, O.ptr(b), 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("first_seed", [1, 37])
def test_adjoint_matches_finite_differences_of_the_detached_estimator(first_seed):
    w = h = 12
    spp = 4
    tris, mats, cam, params = small_scene(w, h, spp)
    sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats)
    img = render(sc, cam, params, w, h, spp, first_seed=first_seed)
    target = (img * np.float32(0.8) + np.float32(0.02)).astype(np.float32)
    grad_img = np.ascontiguousarray(2.0 * (img.astype(np.float64) - target))
    grad = np.zeros(7 * sc.mats.shape[0])
    O.lib().orc_render_adjoint_seeded(O.ptr(grad_img), w, h, O.ptr(cam), C.byref(sc.scene), C.byref(O.Env()),
                                      O.ptr(params), spp, first_seed, O.ptr(grad), 8)
    grad = grad.reshape(-1, 7)
    scale = np.abs(grad).max()
    assert scale > 0
    hstep = 1e-3
    checked = 0
    for mi in range(sc.mats.shape[0]):
        for k, name in enumerate(PARAMS):
            v = float(sc.mats[mi, WORD[name]])
            if name == "emission" and v <= 0.0:
                assert grad[mi, k] == 0.0
                continue
            losses = []
            for sgn in (+1, -1):
                em = np.ascontiguousarray(sc.mats.copy())
                em[mi, WORD[name]] = np.float32(v + sgn * hstep)
                im = render(sc, cam, params, w, h, spp, em, first_seed)
                losses.append(float(((im.astype(np.float64) - target) ** 2).sum()))
            fd = (losses[0] - losses[1]) / (2 * hstep)
            assert abs(grad[mi, k] - fd) <= 2e-3 * abs(fd) + 2e-4 * scale, (mi, name, grad[mi, k], fd)
            checked += 1
    assert checked >= 20


def _word_steps(words, name, hstep):
    """Parameter tables at theta0 +- hstep (stored in float32) and the exact step between them."""
    up, dn = words.copy(), words.copy()
    up[..., WORD[name]] += np.float32(hstep)
    dn[..., WORD[name]] -= np.float32(hstep)
    return up, dn


@pytest.mark.parametrize("seed", range(4))
def test_fp64_brdf_value_is_the_float_value_and_brdf_grad_its_derivative(seed):
    """orc_brdf_eval_f64 equals the forward's float BSDF value, and orc_brdf_grad equals central
    differences of it at h = 1e-5, where double rounding leaves the quotient exact to ~1e-10. The
    bar is 1e-5 of the derivative plus 1e-7 of the value: orc_brdf_grad takes the Oren-Nayar factor
    from the float evaluation, which carries 1e-7 of the value into derivatives (metallic's) that
    are small differences of value-sized terms."""
    rng = np.random.default_rng(100 + seed)
    L = O.lib()
    words = abi.pack_material(*rng.uniform(0.2, 0.8, 3), roughness=rng.uniform(0.1, 0.9),
                              metallic=rng.uniform(0.05, 0.9), specular=rng.uniform(0.1, 0.9))

    def value(ww, wo, wi):
        f = np.zeros(3)
        L.orc_brdf_eval_f64(C.byref(O.ctx(O.Material.from_buffer_copy(ww.tobytes()), wo=wo)), O.v3(wi), O.ptr(ww),
                            O.ptr(f))
        return f

    worst_val = 0.0
    for trial in range(40):
        wo = rng.normal(size=3)
        wo[2] = abs(wo[2]) + 0.05
        wi = rng.normal(size=3)
        wi[2] = abs(wi[2]) + 0.05
        wo /= np.linalg.norm(wo)
        wi /= np.linalg.norm(wi)
        f64 = value(words, wo, wi)
        f32 = L.orc_brdf_eval(C.byref(O.ctx(O.Material.from_buffer_copy(words.tobytes()), wo=wo)), O.v3(wi))
        f32 = np.array([f32.x, f32.y, f32.z])
        worst_val = max(worst_val, float(np.max(np.abs(f64 - f32) / np.maximum(np.abs(f64), 1e-12))))
        a = rng.uniform(0.5, 1.5, 3)
        grad = np.zeros(6)
        a = np.ascontiguousarray(a)
        L.orc_brdf_grad(C.byref(O.ctx(O.Material.from_buffer_copy(words.tobytes()), wo=wo)), O.v3(wi),
                        O.ptr(a), O.ptr(words), O.ptr(grad))
        for k, name in enumerate(PARAMS[:6]):
            up, dn = _word_steps(words, name, 1e-5)
            step = float(up[WORD[name]]) - float(dn[WORD[name]])
            fd = float(a @ (value(up, wo, wi) - value(dn, wo, wi))) / step
            assert abs(grad[k] - fd) <= 1e-5 * abs(fd) + 1e-7 * abs(float(a @ f64)), (trial, name, grad[k], fd)
    assert worst_val <= 2e-5, worst_val


def render_f64(sc, cam, params, w, h, spp, eval_mats=None, first_seed=1, mask=None):
    img = np.zeros((w * h, 3))
    mask = None if mask is None else np.ascontiguousarray(mask)  # held: ptr() is a bare address
    O.lib().orc_render_detached_f64_seeded(O.ptr(img), w, h, O.ptr(cam), C.byref(sc.scene),
                                           None if eval_mats is None else O.ptr(eval_mats), C.byref(O.Env()),
                                           O.ptr(params), spp, first_seed,
                                           None if mask is None else O.ptr(mask), 8)
    return img


def test_fp64_detached_estimator_is_the_float_estimator():
    """The double evaluation is the same estimator: equal to the float image to float rounding,
    at the primal parameters and at perturbed ones; a mask skips the other pixels."""
    w = h = 12
    spp = 4
    tris, mats, cam, params = small_scene(w, h, spp)
    sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats)
    em = np.ascontiguousarray(sc.mats.copy())
    em[:, WORD["roughness"]] = np.clip(em[:, WORD["roughness"]] * np.float32(1.1), 0.0, 1.0)
    em[:, WORD["r"]] *= np.float32(0.9)
    for table in (None, em):
        a = render(sc, cam, params, w, h, spp, table)
        b = render_f64(sc, cam, params, w, h, spp, table)
        assert np.abs(a).max() > 0
        assert np.all(np.abs(b - a) <= 1e-5 * np.maximum(np.abs(b), 1e-3 * np.abs(b).max())), \
            float(np.max(np.abs(b - a)))
    mask = np.zeros((w * h, 3))
    mask[5 * w + 3: 5 * w + 9, 1] = 1.0
    c = render_f64(sc, cam, params, w, h, spp, mask=mask)
    full = render_f64(sc, cam, params, w, h, spp)
    on = mask.any(axis=1)
    assert np.array_equal(c[on], full[on]) and not c[~on].any()


@pytest.mark.parametrize("first_seed", [1, 37])
def test_adjoint_matches_fp64_differences_of_the_detached_estimator(first_seed):
    """The FP64 adjoint against central differences (h = 1e-4) of the detached estimator evaluated
    in double, for a random linear loss: no float rounding in the loss difference, so the bar is
    1e-5 relative (200 times below the float-difference test above). What remains is the
    adjoint's own float inputs (throughput, remaining radiance, Oren-Nayar factor): 4e-7 for 49
    of the 50 parameters here, 2.3e-6 for one small roughness gradient."""
    w = h = 12
    spp = 4
    tris, mats, cam, params = small_scene(w, h, spp)
    sc = O.OracleScene(abi.host_to_gpu_triangles(tris), mats)
    rng = np.random.default_rng(first_seed)
    gi = np.ascontiguousarray(rng.normal(size=(w * h, 3)))
    grad = np.zeros(7 * sc.mats.shape[0])
    O.lib().orc_render_adjoint_seeded(O.ptr(gi), w, h, O.ptr(cam), C.byref(sc.scene), C.byref(O.Env()),
                                      O.ptr(params), spp, first_seed, O.ptr(grad), 8)
    grad = grad.reshape(-1, 7)
    scale = np.abs(grad).max()
    checked = 0
    for mi in range(sc.mats.shape[0]):
        for k, name in enumerate(PARAMS):
            v = float(sc.mats[mi, WORD[name]])
            if name == "emission" and v <= 0.0:
                continue
            up, dn = _word_steps(sc.mats, name, 1e-4)
            up[np.arange(len(up)) != mi] = sc.mats[np.arange(len(up)) != mi]
            dn[np.arange(len(dn)) != mi] = sc.mats[np.arange(len(dn)) != mi]
            step = float(up[mi, WORD[name]]) - float(dn[mi, WORD[name]])
            lu = float((gi * render_f64(sc, cam, params, w, h, spp, np.ascontiguousarray(up), first_seed)).sum())
            ld = float((gi * render_f64(sc, cam, params, w, h, spp, np.ascontiguousarray(dn), first_seed)).sum())
            fd = (lu - ld) / step
            err = abs(grad[mi, k] - fd) / max(abs(fd), 1e-6 * scale)
            assert err <= 1e-5, (mi, name, grad[mi, k], fd)
            checked += 1
    assert checked >= 20
