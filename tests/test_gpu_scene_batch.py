"""Batch renders (ptBatchCreate + pathTracerRenderFixedBatch / pathTracerAdjointRenderBatch):
several scenes of different sizes, materials (textures included), depths, cameras and
environments in one wavefront give, scene by scene, exactly their own renders -- the fixed-spp
render (itself equal to the oracle, tests/test_gpu_batch.py) and, for the adjoint, the seeded
single-scene adjoint (up to the order of the FP64 sums). Also: another seed range, a graph replay
and hero-wavelength mode."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi, rgb2spec
from oracle import oracle as O
from tests.helpers import dense_plant
from tests.test_gpu_parity import compare_accumulators, oracle_scene

pytestmark = pytest.mark.gpu


def _scenes(kind_a="zoo", env=True):
    """Three scenes: a textured material zoo (24x20, depth 8), a thin-lens diffuse plant with an
    HDR environment (17x23, depth 5) and a small C4 bush with two area lights (40x40, depth 8)."""
    from lsystemeditor_amd import inverse, render
    from tests.test_gpu_launchers import hdr_image
    out = []
    tris, mats, tex, cam, params = dense_plant(24, 20, spp=4, materials=kind_a)
    out.append((abi.Scene(tris, mats, tex), render.DeviceInputs(cam, params), (24, 20), (tris, mats, tex, cam, params)))
    tris, mats, tex, cam, params = dense_plant(17, 23, spp=4, materials="diff", max_depth=5, thin_lens=True)
    e = abi.Environment(hdr_image()) if env else None
    out.append((abi.Scene(tris, mats, tex, bvh="sah"), render.DeviceInputs(cam, params, env=e), (17, 23),
                (tris, mats, tex, cam, params)))
    spec = inverse.c4_specs(count=1, res=40, spp=4, iterations=(3, 3))[0]
    out.append((abi.Scene(spec.triangles, spec.materials, bvh="sah"), render.DeviceInputs(spec.camera, spec.params),
                (40, 40), (spec.triangles, spec.materials, None, spec.camera, spec.params)))
    return out


def _batch(items):
    from lsystemeditor_amd import render
    return render.SceneBatch([i[0] for i in items], [i[1] for i in items], [i[2] for i in items])


def _single(item, spp, first_iter=0):
    from lsystemeditor_amd import render
    sc, inp, (w, h), _ = item
    acc = render.Accumulator.allocate(w, h)
    if first_iter == 0:
        render.render_fixed(sc, acc, inp, spp)
    else:
        render.render_wavefront(sc, acc, inp, first_iter=first_iter, passes=spp)
    return acc.host()


def _part(host, sl):
    return {k: v[sl] for k, v in host.items()}


def _same(got, want, label):
    """Bit-identical accumulators (the batch runs the single-scene arithmetic)."""
    for k in ("acc", "counts", "mean", "m2", "conv"):
        g, w = np.ascontiguousarray(got[k]), np.ascontiguousarray(want[k])
        assert g.shape == w.shape, (label, k)
        bad = np.nonzero(g.view(np.uint8).reshape(g.shape[0], -1) != w.view(np.uint8).reshape(w.shape[0], -1))[0]
        assert bad.size == 0, f"{label}/{k}: {np.unique(bad).size} pixels differ, first {np.unique(bad)[:6]}"


def test_batch_info_offsets(cuda):
    items = _scenes()
    b = _batch(items)
    assert b.pixel_offsets == [0, 24 * 20, 24 * 20 + 17 * 23]
    assert b.total_pixels == 24 * 20 + 17 * 23 + 40 * 40
    rows = [it[3][1].reshape(-1, abi.MATERIAL_WORDS).shape[0] for it in items]
    assert b.grad_row_offsets == [0, rows[0], rows[0] + rows[1]]
    assert b.grad_rows == sum(rows)
    assert b.max_depth == 8
    b.close()


def test_batch_render_equals_single_renders(cuda):
    items = _scenes()
    b = _batch(items)
    b.render_fixed(4)
    got = b.acc.host()
    for i, it in enumerate(items):
        _same(_part(got, b.pixels(i)), _single(it, 4), f"batch scene {i}")
    # scene 0 against the oracle directly as well
    tris, mats, tex, cam, params = items[0][3]
    acc_o, _ = O.render_adaptive(oracle_scene(tris, mats, tex), cam, params, 24, 20, threads=8)
    compare_accumulators(_part(got, b.pixels(0)), acc_o, "batch scene 0 / oracle")
    b.close()


def test_batch_render_other_seed_range(cuda):
    """Seeds 9..12 equal the single-scene schedule starting at pass 8."""
    items = _scenes()
    b = _batch(items)
    b.render_fixed(4, first_seed=9)
    got = b.acc.host()
    for i, it in enumerate(items):
        _same(_part(got, b.pixels(i)), _single(it, 4, first_iter=8), f"seeded batch scene {i}")
    b.close()


def test_batch_graph_replay_equals_eager(cuda):
    import torch
    items = _scenes()
    b = _batch(items)
    b.reserve(1 << 20)
    b.render_fixed(4, max_paths=1 << 20)
    eager = b.acc.host()
    stream = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        b.render_fixed(4, max_paths=1 << 20, stream=torch.cuda.current_stream())
    for rep in range(3):
        b.acc.acc.zero_()      # other GPU work between replays
        b.acc.counts.zero_()
        g.replay()
        torch.cuda.synchronize()
        got = b.acc.host()
        for k in ("acc", "counts", "mean", "m2", "conv"):
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(eager[k]).view(np.uint8)), (rep, k)
    b.close()


def test_batch_render_spectral_equals_single_renders(cuda):
    table = rgb2spec.table(16)
    s = abi.Spectral(*table)
    s.enable(380.0, 720.0)
    try:
        items = _scenes()
        b = _batch(items)
        b.render_fixed(4)
        got = b.acc.host()
        for i, it in enumerate(items):
            _same(_part(got, b.pixels(i)), _single(it, 4), f"spectral batch scene {i}")
        b.close()
    finally:
        abi.Spectral.disable()
        s.close()


def test_batch_adjoint_equals_single_adjoints(cuda):
    import torch
    from lsystemeditor_amd import render
    items = _scenes(kind_a="diff")
    b = _batch(items)
    spp, first_seed = 4, 5
    rng = np.random.default_rng(3)
    grad_img = torch.from_numpy(rng.standard_normal((b.total_pixels, 3)).astype(np.float32)).cuda()
    out = torch.zeros((b.grad_rows, 7), dtype=torch.float64, device="cuda")
    abi.check(abi.lib().pathTracerAdjointRenderBatch(grad_img.data_ptr(), b.handle, spp, first_seed, out.data_ptr(),
                                                     None), "pathTracerAdjointRenderBatch")
    got = out.cpu().numpy()
    for i, (sc, inp, (w, h), host) in enumerate(items):
        nm = host[1].reshape(-1, abi.MATERIAL_WORDS).shape[0]
        gi = grad_img[b.pixels(i)].contiguous()
        want = torch.zeros((nm, 7), dtype=torch.float64, device="cuda")
        abi.check(abi.lib().pathTracerAdjointRenderSeeded(
            gi.data_ptr(), w, h, spp, inp.camera.data_ptr(), sc.device, inp.env_ptr, inp.params.data_ptr(), nm, None,
            want.data_ptr(), first_seed, None), "pathTracerAdjointRenderSeeded")
        want = want.cpu().numpy()
        part = got[b.grad_row_offsets[i]:b.grad_row_offsets[i] + nm]
        scale = np.abs(want).max()
        assert scale > 0, i
        assert np.all(np.abs(part - want) <= 1e-9 * scale), (i, np.abs(part - want).max() / scale)
    b.close()


def test_batch_rejects_bad_arguments(cuda):
    from lsystemeditor_amd import render
    L = abi.lib()
    assert not L.ptBatchCreate(None, 1)
    entries = (abi.PtBatchSceneGpu * 1)()
    assert not L.ptBatchCreate(entries, 1)  # null scene
    assert "invalid" in abi.last_error()
    items = _scenes()
    b = _batch(items)
    assert not L.pathTracerAdjointRenderBatch(None, b.handle, 4, 1, None, None)
    b.close()


def test_batch_autograd_equals_per_scene_autograd(cuda):
    """render_batch_decorrelated: images and parameter gradients equal the per-scene
    DecorrelatedRenderFunction's."""
    import torch
    from lsystemeditor_amd import diff, inverse
    specs = inverse.c4_specs(count=3, res=24, spp=4, iterations=(3, 4))
    rs = [diff.DiffRenderer(abi.Scene(s.triangles, s.materials, bvh="sah"), s.materials, s.camera, s.params, s.width,
                            s.height, s.spp) for s in specs]
    thetas = [r.theta0().requires_grad_(True) for r in rs]
    targets = [torch.full((s.width * s.height, 3), 0.3, device="cuda") for s in specs]
    br = diff.BatchDiffRenderer(rs)
    img = diff.render_batch_decorrelated(thetas, br, 8, 12)
    ((img - torch.cat(targets)) ** 2).sum().backward()
    got = [t.grad.clone() for t in thetas]
    for i, (r, tgt) in enumerate(zip(rs, targets)):
        th = r.theta0().requires_grad_(True)
        im = diff.render_image_decorrelated(th, r, 8, 12)
        assert torch.equal(im, img[br.pixels(i)]), i
        ((im - tgt) ** 2).sum().backward()
        assert torch.allclose(got[i], th.grad, rtol=1e-6, atol=1e-7 * float(th.grad.abs().max())), i
