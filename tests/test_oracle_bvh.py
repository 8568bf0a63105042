"""Oracle BVH build and traversal: structural invariants, brute force, the reference's UV test."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi
from oracle import oracle as O


def soup(n, seed, dup=True):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    v = (c + rng.normal(0, 0.6, (n, 3, 3))).astype(np.float32)
    if dup and n >= 16:
        v[: n // 8] = v[n // 8: 2 * (n // 8)]  # identical centroids: exercises the index tie-break
    return abi.host_to_gpu_triangles(abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2]))


def node_fields(nodes):
    u = nodes.view(np.uint32)
    return nodes[:, 0:3], nodes[:, 4:7], u[:, 8], u[:, 9], u[:, 10], u[:, 11]


def tri_bounds(t):
    v = t[:, 0:9].reshape(-1, 3, 3)
    return v.min(1), v.max(1)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 777])
def test_structure_invariants(n):
    tris = soup(n, n)
    sc = O.OracleScene(tris, abi.pack_material()[None])
    nodes = sc.nodes[: sc.node_count]
    assert sc.node_count == 2 * n - 1 and sc.root == 0
    bmin, bmax, left, right, start, count = node_fields(nodes)
    tmin, tmax = tri_bounds(tris)
    leaves = np.nonzero(count == 1)[0]
    assert len(leaves) == n and np.all(count[count != 1] == 0)
    assert sorted(start[leaves].tolist()) == list(range(n)), "each triangle in exactly one leaf"
    assert np.array_equal(bmin[leaves], tmin[start[leaves]]) and np.array_equal(bmax[leaves], tmax[start[leaves]])

    def walk(i, lo, cnt):
        # pre-order: left child is i + 1, right child is i + 2 * (cnt // 2)
        if cnt == 1:
            assert count[i] == 1
            return
        mid = cnt // 2
        assert left[i] == i + 1 and right[i] == i + 2 * mid
        L_, R_ = left[i], right[i]
        assert np.array_equal(bmin[i], np.minimum(bmin[L_], bmin[R_]))
        assert np.array_equal(bmax[i], np.maximum(bmax[L_], bmax[R_]))
        walk(L_, lo, mid)
        walk(R_, lo + mid, cnt - mid)

    walk(0, 0, n)


def leaf_sets(nodes, i, cnt, out):
    u = nodes.view(np.uint32)
    if cnt == 1:
        out.append([int(u[i, 10])])
        return out[-1]
    mid = cnt // 2
    a = leaf_sets(nodes, int(u[i, 8]), mid, out)
    b = leaf_sets(nodes, int(u[i, 9]), cnt - mid, out)
    return a + b


def test_median_split_rule():
    """Every split puts the count/2 smallest (centroid on the longest extent axis, index) on the left."""
    n = 300
    tris = soup(n, 5)
    sc = O.OracleScene(tris, abi.pack_material()[None])
    nodes = sc.nodes[: sc.node_count]
    v = tris[:, 0:9].reshape(-1, 3, 3)
    cent = ((v[:, 0] + v[:, 1]) + v[:, 2]) * np.float32(1.0 / 3.0)
    u = nodes.view(np.uint32)

    def check(i, cnt):
        if cnt == 1:
            return [int(u[i, 10])]
        mid = cnt // 2
        L_ = check(int(u[i, 8]), mid)
        R_ = check(int(u[i, 9]), cnt - mid)
        ext = nodes[i, 4:7] - nodes[i, 0:3]
        axis = 0 if (ext[0] >= ext[1] and ext[0] >= ext[2]) else (1 if ext[1] >= ext[2] else 2)
        keys = sorted((float(cent[t, axis]), t) for t in L_ + R_)
        assert sorted(L_) == sorted(t for _, t in keys[:mid])
        return L_ + R_

    check(0, n)


def brute(tris, ro, rd, tmin=0.0, tmax=1e6):
    L = O.lib()
    best = (tmax, None)
    rdn = L.orc_vec_normalize(O.v3(rd))
    for i in range(tris.shape[0]):
        t = C.c_float()
        n = O.Vec3()
        uv = O.Vec2()
        # the first (smallest) index wins an exact tie of t
        if L.orc_ray_triangle(O.v3(ro), rdn, tris[i].ctypes.data, tmin, best[0], C.byref(t), C.byref(n), C.byref(uv)) \
                and (best[1] is None or t.value < best[0]):
            best = (t.value, i)
    return best


def test_traversal_equals_brute_force():
    tris = soup(400, 9, dup=False)
    sc = O.OracleScene(tris, abi.pack_material()[None])
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(600):
        ro = rng.uniform(-9, 9, 3).astype(np.float32)
        rd = (rng.uniform(-5, 5, 3) - ro).astype(np.float32)
        h = sc.trace(ro, rd)
        t, i = brute(tris, ro, rd)
        if i is None:
            assert not h.hit
        else:
            hits += 1
            assert h.hit and h.t == np.float32(t) and h.triangleIndex == i
    assert hits > 200


def test_uv_interpolation_reference_case():
    """PathTracerMeshUvTests.cpp:264-287: hit at (0.25, 0.25) of the unit right triangle."""
    tri = abi.host_to_gpu_triangles(abi.pack_host_triangles(
        [[0, 0, 0]], [[1, 0, 0]], [[0, 1, 0]], [[0, 0, 1]], [[0, 0, 1]], [[0, 0, 1]],
        [[0, 0]], [[1, 0]], [[0, 1]]))
    L = O.lib()
    t = C.c_float()
    n = O.Vec3()
    uv = O.Vec2()
    assert L.orc_ray_triangle(O.Vec3(0.25, 0.25, 1.0), O.Vec3(0, 0, -1), tri[0].ctypes.data, 0.0, 10.0, C.byref(t),
                              C.byref(n), C.byref(uv))
    assert abs(uv.x - 0.25) < 1e-5 and abs(uv.y - 0.25) < 1e-5 and t.value == 1.0


def test_two_sided_and_normal_faces_ray():
    quad = abi.host_to_gpu_triangles(abi.pack_host_triangles(
        [[-1, -1, 0], [-1, -1, 0]], [[1, -1, 0], [1, 1, 0]], [[1, 1, 0], [-1, 1, 0]]))
    sc = O.OracleScene(quad, abi.pack_material()[None])
    down = sc.trace((0.2, 0.1, 3.0), (0.01, 0.02, -1))
    up = sc.trace((0.2, 0.1, -2.0), (0.01, 0.02, 1))
    assert down.hit and up.hit and abs(down.t - 3.0) < 1e-3 and abs(up.t - 2.0) < 1e-3
    assert down.normal.z == 1.0 and up.normal.z == -1.0


def test_axis_parallel_ray_quirk():
    """MeshAccelIntersectCore.h:125-128 maps |d| <= 1e-8 to invD = 0, so that slab yields
    t0 = t1 = 0 and clamps tExit to 0: an exactly axis-parallel ray misses every box it does
    not start inside. Restated as is (the camera and bounce rays are jittered)."""
    quad = abi.host_to_gpu_triangles(abi.pack_host_triangles(
        [[-1, -1, 0], [-1, -1, 0]], [[1, -1, 0], [1, 1, 0]], [[1, 1, 0], [-1, 1, 0]]))
    sc = O.OracleScene(quad, abi.pack_material()[None])
    assert not sc.trace((0.2, 0.1, 3.0), (0, 0, -1)).hit
    assert sc.trace((0.2, 0.1, 3.0), (1e-6, 1e-6, -1)).hit


def test_tmax_is_exclusive_and_tmin_inclusive():
    quad = abi.host_to_gpu_triangles(abi.pack_host_triangles([[-1, -1, 0]], [[1, -1, 0]], [[0, 1, 0]]))
    sc = O.OracleScene(quad, abi.pack_material()[None])
    d = (2.0 ** -20, 2.0 ** -20, -1.0)  # no zero component; t is exactly 2 for this origin
    h = sc.trace((0, 0, 2.0), d, 0.0, 3.0)
    assert h.hit and h.t == 2.0
    assert not sc.trace((0, 0, 2.0), d, 0.0, 2.0).hit   # t == tMax: closestT < tMax fails
    assert sc.trace((0, 0, 2.0), d, 2.0, 3.0).hit       # t == tMin accepted
    assert not sc.trace((0, 0, 2.0), d, 2.5, 3.0).hit


def test_emissive_cdf_closed_form():
    # two emissive triangles of area 2 and 0.5, one non-emissive; luminance weights 0.2126r+...
    tris = abi.host_to_gpu_triangles(abi.pack_host_triangles(
        [[0, 0, 0], [0, 0, 0], [0, 0, 5]], [[2, 0, 0], [1, 0, 0], [1, 0, 5]], [[0, 2, 0], [0, 1, 0], [0, 1, 5]],
        material=np.array([1, 0, 2], np.uint32)))
    mats = np.stack([abi.pack_material(), abi.pack_material(1, 1, 1, emission=3.0),
                     abi.pack_material(1, 0, 0, emission=2.0)])
    sc = O.OracleScene(tris, mats)
    assert sc.em_count == 2
    assert list(sc.em_idx[:2]) == [0, 2]
    w0 = np.float32(2.0) * np.float32(3.0 * np.float32(np.float32(0.2126) + np.float32(0.7152) + np.float32(0.0722)))
    w1 = np.float32(0.5) * np.float32(2.0 * 0.2126)
    assert sc.em_cdf[1] == w0 and abs(sc.em_cdf[2] - (w0 + w1)) <= 1e-6 * w0


def test_exact_ties_go_to_the_smaller_index():
    """Two copies of a triangle (indices 0 and 5, five others in between): every ray through it
    reports index 0, whichever order the tree visits the copies in; the brute force agrees."""
    rng = np.random.default_rng(2)
    v = rng.uniform(-3, 3, (6, 3, 3)).astype(np.float32)
    v[0] = [[-1, -1, 0], [1, -1, 0], [0, 1, 0]]
    v[5] = v[0]
    v[1:5, :, 2] += 6.0   # the others lie behind the viewer
    tris = abi.host_to_gpu_triangles(abi.pack_host_triangles(v[:, 0], v[:, 1], v[:, 2]))
    sc = O.OracleScene(tris, abi.pack_material()[None])
    for x, y in ((0.0, 0.0), (0.3, -0.5), (-0.2, 0.1)):
        ro = np.array([x, y, 2.0], np.float32)
        rd = np.array([0.01, 0.02, -1.0], np.float32)   # (axis-parallel rays miss: invRd = 0 quirk)
        h = sc.trace(ro, rd)
        assert h.hit and h.triangleIndex == 0
        assert brute(tris, ro, rd) == (h.t, 0)


def test_traversal_equals_brute_force_with_copies():
    """A soup in which an eighth of the triangles are exact copies: ties everywhere."""
    tris = soup(400, 13, dup=True)
    sc = O.OracleScene(tris, abi.pack_material()[None])
    rng = np.random.default_rng(6)
    v = tris[:, 0:9].reshape(-1, 3, 3)
    hits = 0
    for j in range(300):
        tgt = v[j % 50].mean(0)   # the copied triangles
        ro = (tgt + rng.normal(0, 5.0, 3)).astype(np.float32)
        rd = (tgt - ro).astype(np.float32)
        h = sc.trace(ro, rd)
        t, i = brute(tris, ro, rd)
        if i is None:
            assert not h.hit
        else:
            hits += 1
            assert h.hit and h.t == np.float32(t) and h.triangleIndex == i
    assert hits > 100
