"""Graph-capturable tree rebuilds (ptBuildWorkspaceCreate, ptSceneRebuild, ptSceneRebuildCheck):
after the scene's tree buffers are overwritten, a rebuild restores the binary tree, its leaf slots
and the 4-wide tree bit for bit; rebuilds of several scenes and their batch render captured in one
graph give, on every replay (with other GPU work and the buffers overwritten between replays),
the eager batch render; and the entry points reject what they cannot rebuild."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi
from tests.test_gpu_sah import random_soup

pytestmark = pytest.mark.gpu

DESC_NODES, DESC_TRIPOS, DESC_TRISHADE, DESC_WIDE = 0, 8, 16, 112  # MeshAccelSceneGpu pointer offsets


def _fill(ptr: int, nbytes: int, byte: int = 0xFF) -> None:
    """Overwrites device memory with `byte` (hipMemset of the HIP runtime the library uses)."""
    if ptr == 0 or nbytes == 0:
        return
    memset = abi.lib().hipMemset
    memset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    memset.restype = C.c_int
    assert memset(C.c_void_p(ptr), byte, C.c_size_t(nbytes)) == 0


def _clobber(sc: abi.Scene) -> None:
    import torch
    torch.cuda.synchronize()
    desc = np.frombuffer(sc.description(), np.uint8)
    ptr = lambda off: int(np.frombuffer(desc[off:off + 8].tobytes(), np.uint64)[0])
    n = sc.triangle_count
    wide_count = int(np.frombuffer(desc[120:124].tobytes(), np.uint32)[0])
    _fill(ptr(DESC_NODES), max(n - 1, 0) * 64)
    _fill(ptr(DESC_TRIPOS), n * 48)
    _fill(ptr(DESC_TRISHADE), n * 64)
    _fill(ptr(DESC_WIDE), wide_count * 128)
    torch.cuda.synchronize()


def _layout(sc: abi.Scene, wide: bool):
    """The tree's defined words: node boxes and child references (words 14-15 of a node are unused
    and never written), leaf slots, levels, and the 4-wide tree with its root."""
    nodes, slots, levels = sc.read_layout()
    out = [nodes[:, :14].view(np.uint32).copy(), slots.view(np.uint32).copy(), levels]
    if wide:
        w, root = sc.read_wide()
        out += [w.view(np.uint32).copy(), root]
    return out


@pytest.mark.parametrize("bvh", ["sah", "sah-binary"])
@pytest.mark.parametrize("n", [2, 3, 7, 64, 1000, 4097])
def test_rebuild_restores_the_tree(cuda, bvh, n):
    tris = random_soup(n, 300 + n)
    sc = abi.Scene(tris, abi.pack_material()[None], bvh=bvh)
    want = _layout(sc, bvh == "sah")
    ws = abi.BuildWorkspace(n + 5)
    _clobber(sc)
    assert not np.array_equal(_layout(sc, bvh == "sah")[1], want[1])
    sc.rebuild(ws)
    sc.rebuild_check(ws)
    got = _layout(sc, bvh == "sah")
    for k, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(np.asarray(a), np.asarray(b)), (bvh, n, k)
    ws.close()


def test_rebuild_rejects_what_it_cannot_rebuild(cuda):
    tris = random_soup(50, 7)
    med = abi.Scene(tris, abi.pack_material()[None], bvh="median")
    sah = abi.Scene(tris, abi.pack_material()[None], bvh="sah")
    ws = abi.BuildWorkspace(49)
    with pytest.raises(RuntimeError, match="binned-SAH"):
        med.rebuild(ws)
    with pytest.raises(RuntimeError, match="50 triangles for a workspace of 49"):
        sah.rebuild(ws)
    with pytest.raises(RuntimeError, match="capacity 0"):
        abi.BuildWorkspace(0)
    ws.close()


def test_rebuilds_and_batch_render_in_one_graph(cuda):
    import torch
    from lsystemeditor_amd import inverse, render
    from tests.helpers import dense_plant
    items = []
    tris, mats, tex, cam, params = dense_plant(24, 20, spp=4, materials="zoo")
    items.append((abi.Scene(tris, mats, tex, bvh="sah"), render.DeviceInputs(cam, params), (24, 20)))
    tris, mats, tex, cam, params = dense_plant(17, 23, spp=4, materials="diff", max_depth=5, thin_lens=True)
    items.append((abi.Scene(tris, mats, tex, bvh="sah-binary"), render.DeviceInputs(cam, params), (17, 23)))
    spec = inverse.c4_specs(count=1, res=40, spp=4, iterations=(3, 3))[0]
    items.append((abi.Scene(spec.triangles, spec.materials, bvh="sah"), render.DeviceInputs(spec.camera, spec.params),
                  (40, 40)))
    b = render.SceneBatch([i[0] for i in items], [i[1] for i in items], [i[2] for i in items])
    b.reserve(1 << 20)
    b.render_fixed(4, max_paths=1 << 20)
    eager = b.acc.host()
    trees = [_layout(sc, sc.bvh == "sah") for sc, _, _ in items]
    ws = abi.BuildWorkspace(max(sc.triangle_count for sc, _, _ in items))
    stream = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        s = torch.cuda.current_stream()
        for sc, _, _ in items:
            sc.rebuild(ws, s.cuda_stream)
        b.render_fixed(4, max_paths=1 << 20, stream=s)
    for rep in range(3):
        for sc, _, _ in items:
            _clobber(sc)
        b.acc.acc.zero_()      # other GPU work between replays
        b.acc.counts.zero_()
        g.replay()
        torch.cuda.synchronize()
        for sc, _, _ in items:          # every scene's own status record, not only the last rebuild's
            sc.rebuild_check(ws)
        got = b.acc.host()
        for k in ("acc", "counts", "mean", "m2", "conv"):
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(eager[k]).view(np.uint8)), (rep, k)
        for (sc, _, _), want in zip(items, trees):
            for k, (x, y) in enumerate(zip(_layout(sc, sc.bvh == "sah"), want)):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (rep, k)
    b.close()
    ws.close()


def test_each_scene_checks_its_own_last_rebuild(cuda):
    """Several scenes rebuilt in a row on one workspace: every scene's check reads its own status
    record (a failure of an earlier rebuild is not hidden by a later one), a check with no rebuild
    since the last one is refused, and a rebuild uses the mode the scene was built with."""
    scs = [abi.Scene(random_soup(n, 40 + n), abi.pack_material()[None], bvh="sah") for n in (5, 300, 2049)]
    wants = [_layout(sc, True) for sc in scs]
    ws = abi.BuildWorkspace(2049)
    for sc in scs:
        _clobber(sc)
        sc.rebuild(ws)
    for sc, want in zip(scs, wants):
        sc.rebuild_check(ws)
        assert all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(_layout(sc, True), want))
    with pytest.raises(RuntimeError, match="no rebuild of this scene has run"):
        scs[0].rebuild_check(ws)
    # switching the scene's build mode after its build does not change what a rebuild makes
    L = abi.lib()
    abi.check(L.ptSceneSetBuildMode(scs[1].handle, 2), "ptSceneSetBuildMode")
    _clobber(scs[1])
    scs[1].rebuild(ws)
    scs[1].rebuild_check(ws)
    assert all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(_layout(scs[1], True), wants[1]))
    ws.close()
