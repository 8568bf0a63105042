"""Fixed-spp wavefront render (pathTracerRenderFixedWavefront) against the oracle, and the
batched HIP-graph replay of several scenes against their eager renders (several replays, with
other GPU work between them)."""
import numpy as np
import pytest

from lsystemeditor_amd import abi
from oracle import oracle as O
from tests.helpers import dense_plant
from tests.test_gpu_parity import compare_accumulators, oracle_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_paths", [0, 3000])
def test_fixed_wavefront_matches_oracle(cuda, max_paths):
    from lsystemeditor_amd import render
    w = h = 40
    tris, mats, tex, cam, params = dense_plant(w, h, spp=12, materials="zoo")
    sc = abi.Scene(tris, mats, tex)
    ref = oracle_scene(tris, mats, tex)
    acc_o, _ = O.render_adaptive(ref, cam, params, w, h, threads=8)
    dev = render.Accumulator.allocate(w, h)
    render.render_fixed(sc, dev, render.DeviceInputs(cam, params), 12, max_paths=max_paths)
    compare_accumulators(dev.host(), acc_o, f"fixed/{max_paths}")


def test_batch_graph_replay_equals_eager(cuda):
    import torch
    from lsystemeditor_amd import batch
    specs = batch.batch_specs(count=4, spp=4, iterations=(3, 4))
    for s in specs:   # keep the test small: at most 256^2
        if s.width > 256:
            s.width = s.height = 256
            s.camera[9] = np.float32(1.0)
    br = batch.BatchRenderer(specs, bvh="sah", max_paths=1 << 20)
    br.render_all()
    torch.cuda.synchronize()
    eager = [a.host() for a in br.accs]
    br.capture()
    for rep in range(3):
        for a in br.accs:        # other GPU work between replays
            a.acc.zero_()
            a.counts.zero_()
        br.replay()
        torch.cuda.synchronize()
        for i, a in enumerate(br.accs):
            got = a.host()
            for k in ("acc", "counts", "mean", "m2"):
                assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(eager[i][k]).view(np.uint8)), \
                    (rep, i, k)


def test_build_and_render_graph_equals_eager(cuda):
    """Every scene's tree rebuild and the batch wavefront in one graph: each replay gives the eager
    batch render, and the last rebuild reproduces its scene's tree."""
    import torch
    from lsystemeditor_amd import batch
    specs = batch.batch_specs(count=4, spp=4, iterations=(3, 4))
    for s in specs:   # keep the test small: at most 256^2
        if s.width > 256:
            s.width = s.height = 256
            s.camera[9] = np.float32(1.0)
    br = batch.BatchRenderer(specs, bvh="sah", max_paths=1 << 20)
    br.render_batched()
    torch.cuda.synchronize()
    eager = br._batch().acc.host()
    br.capture_build_and_render()
    for rep in range(3):
        br._batch().acc.acc.zero_()   # other GPU work between replays
        br.replay_build_and_render()
        torch.cuda.synchronize()
        br.check_rebuild()
        got = br._batch().acc.host()
        for k in ("acc", "counts", "mean", "m2", "conv"):
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(eager[k]).view(np.uint8)), (rep, k)
