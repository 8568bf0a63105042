"""GPU adjoint (pathTracerAdjointRender) against the oracle's FP64 adjoint on the same loss
gradient image -- the oracle's gradient is itself pinned to central finite differences of the
detached estimator on the CPU -- plus the material-update path and the autograd wrapper."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi
from oracle import oracle as O
from tests.helpers import dense_plant
from tests.test_gpu_parity import oracle_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D(cuda):
    from lsystemeditor_amd import diff
    return diff


def setup(D, w, h, spp, depth=6, bvh="median"):
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, max_depth=depth, materials="diff")
    sc = abi.Scene(tris, mats, tex, bvh=bvh)
    r = D.DiffRenderer(sc, mats, cam, params, w, h, spp)
    ref = oracle_scene(tris, mats, tex)
    return r, ref, cam, r.params_host


def oracle_image(ref, cam, params, w, h, spp, eval_mats=None):
    img = np.zeros((w * h, 3), np.float32)
    O.lib().orc_render_detached(O.ptr(img), w, h, O.ptr(cam), C.byref(ref.scene),
                                None if eval_mats is None else O.ptr(eval_mats), C.byref(O.Env()), O.ptr(params),
                                spp, 8)
    return img


def test_fixed_spp_forward_matches_oracle(D):
    w = h = 24
    r, ref, cam, params = setup(D, w, h, 8)
    img = r.forward(r.theta0().cuda()).cpu().numpy()
    want = oracle_image(ref, cam, params, w, h, 8)
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))


def test_material_update_equals_rebuilt_scene(D):
    import torch
    w = h = 24
    r, ref, cam, params = setup(D, w, h, 4)
    theta = r.theta0()
    theta[:, 3] = torch.clamp(theta[:, 3] * 1.1, 0.05, 0.95)   # roughness
    theta[-1, 6] = theta[-1, 6] * 1.5                          # light emission
    img = r.forward(theta.cuda()).cpu().numpy()
    mats = r.base.copy()
    mats[:, list(D.PARAM_WORDS)] = theta.numpy()
    ref2 = oracle_scene(*dense_plant(w, h, spp=4, max_depth=6, materials="diff")[0:1], mats)
    want = oracle_image(ref2, cam, params, w, h, 4)
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("bvh", ["median", "sah"])
def test_adjoint_matches_oracle_fp64(D, bvh):
    import torch
    w = h = 32
    spp = 8
    r, ref, cam, params = setup(D, w, h, spp, bvh=bvh)
    img = r.forward(r.theta0().cuda())
    target = img * 0.8 + 0.02
    grad_img = (2.0 * (img - target)).contiguous()
    g_gpu = r.adjoint(grad_img).cpu().numpy()
    g_cpu = np.zeros(7 * r.material_count)
    gi = np.ascontiguousarray(grad_img.cpu().numpy().astype(np.float64))
    O.lib().orc_render_adjoint(O.ptr(gi), w, h, O.ptr(cam), C.byref(ref.scene), C.byref(O.Env()), O.ptr(params), spp,
                               O.ptr(g_cpu), 8)
    g_cpu = g_cpu.reshape(-1, 7)
    scale = np.abs(g_cpu).max()
    assert scale > 0
    err = np.abs(g_gpu - g_cpu)
    assert np.all(err <= 1e-4 * np.abs(g_cpu) + 1e-6 * scale), f"max rel err {(err / (np.abs(g_cpu) + 1e-30)).max()}"
    # every live parameter of every material received a gradient
    live = (g_cpu != 0)
    assert live[:, :6].sum() >= 6 * (r.material_count - 1)


def test_seeded_adjoint_matches_oracle_fp64(D):
    """pathTracerAdjointRenderSeeded over seeds 37..44 (render of those samples inside the call)
    against the oracle's adjoint of the same sample set."""
    w = h = 32
    spp = 8
    r, ref, cam, params = setup(D, w, h, spp)
    img = r.forward(r.theta0().cuda(), seed_base=36)
    grad_img = (2.0 * (img - (img * 0.8 + 0.02))).contiguous()
    g_gpu = r.adjoint(grad_img, reuse_forward=False, seed_base=36).cpu().numpy()
    g_cpu = np.zeros(7 * r.material_count)
    gi = np.ascontiguousarray(grad_img.cpu().numpy().astype(np.float64))
    O.lib().orc_render_adjoint_seeded(O.ptr(gi), w, h, O.ptr(cam), C.byref(ref.scene), C.byref(O.Env()),
                                      O.ptr(params), spp, 37, O.ptr(g_cpu), 8)
    g_cpu = g_cpu.reshape(-1, 7)
    scale = np.abs(g_cpu).max()
    assert scale > 0
    err = np.abs(g_gpu - g_cpu)
    assert np.all(err <= 1e-4 * np.abs(g_cpu) + 1e-6 * scale), f"max rel err {(err / (np.abs(g_cpu) + 1e-30)).max()}"
    # the forward of seed base 36 is the oracle's detached image of seeds 37..44
    want = np.zeros((w * h, 3), np.float32)
    O.lib().orc_render_detached_seeded(O.ptr(want), w, h, O.ptr(cam), C.byref(ref.scene), None, C.byref(O.Env()),
                                       O.ptr(params), spp, 37, 8)
    got = img.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_autograd_backward_equals_adjoint(D):
    import torch
    w = h = 16
    r, ref, cam, params = setup(D, w, h, 4)
    theta = r.theta0().cuda().requires_grad_(True)
    img = D.render_image(theta, r)
    target = (img.detach() * 0.9).contiguous()
    loss = ((img - target) ** 2).sum()
    loss.backward()
    direct = r.adjoint((2.0 * (img.detach() - target)).contiguous()).to(torch.float32)
    assert torch.allclose(theta.grad, direct, rtol=1e-6, atol=0.0)


def test_detached_render_matches_oracle(D):
    import torch
    w = h = 20
    r, ref, cam, params = setup(D, w, h, 4)
    theta = r.theta0()
    base = r.forward(theta.cuda()).cpu().numpy()
    same = r.detached(theta).cpu().numpy()
    assert np.array_equal(same.view(np.uint32), base.view(np.uint32))
    pert = theta.clone()
    pert[0, 3] += 0.05
    pert[1, 0] -= 0.05
    pert[-1, 6] *= 1.2
    got = r.detached(pert).cpu().numpy()
    em = r.base.copy()
    em[:, list(D.PARAM_WORDS)] = pert.numpy()
    want = oracle_image(ref, cam, params, w, h, 4, np.ascontiguousarray(em))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_gradient_check_against_finite_differences(D):
    import torch
    from lsystemeditor_amd import scenes
    w = h = 32
    r, ref, cam, params = setup(D, w, h, 8)
    theta0 = r.theta0()
    target_theta = torch.from_numpy(scenes.perturbed_target(theta0.numpy()))
    target = r.forward(target_theta.cuda())
    res = D.gradient_check(r, theta0.cuda(), target, h=1e-3)
    assert res["max_rel_err"] <= 1e-2, res["rel_err"]
    assert res["median_rel_err"] <= 1e-3, res["rel_err"]


def test_adjoint_reusing_forward_radiance_equals_recomputed(D):
    w = h = 24
    r, ref, cam, params = setup(D, w, h, 8)
    img = r.forward(r.theta0().cuda())
    g_img = (img * 0.3 - 0.01).contiguous()
    a = r.adjoint(g_img).cpu().numpy()
    b = r.adjoint(g_img, reuse_forward=False).cpu().numpy()
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max())
