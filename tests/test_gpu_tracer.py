"""The C++ renderer facade (class PathTracer, include/pt_tracer.h) end to end on the GPU: its
worker thread runs the reference render loop (PathTracer.cpp:2220-2543) through the launchers,
and the result must equal the oracle running the same schedule bit for bit — plain adaptive
rendering, preview levels (depth 4, seeds 1..N, then full-resolution passes with seed = iteration
+ 1 and compaction every 4 full-resolution passes), region rendering, display publish and
pickSurface."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi
from lsystemeditor_amd.tracer import Tracer
from oracle import oracle as O
from tests.helpers import dense_plant
from tests.test_gpu_parity import compare_accumulators, oracle_scene

pytestmark = pytest.mark.gpu


class _Acc:
    def __init__(self, d):
        self.acc, self.counts, self.mean, self.m2, self.conv = d["acc"], d["counts"], d["mean"], d["m2"], d["conv"]


def setup_tracer(w, h, spp, min_spp, thr, depth=8, preview=0, bvh_mode=0):
    tris, mats, tex, cam, params = dense_plant(w, h, spp=spp, min_spp=min_spp, threshold=thr, max_depth=depth)
    t = Tracer()
    t.set_preview_steps_per_level(preview)
    t.set_bvh_build_mode(bvh_mode)
    assert t.configure(w, h, tris, mats, tex), abi.last_error()
    t.set_camera(cam)
    t.set_max_samples_per_pixel(spp)
    t.set_min_samples(min_spp)
    t.set_relative_error_threshold(thr)
    t.set_max_path_depth(depth)
    bg = params[0:3]
    t.set_clear_color(float(bg[0]), float(bg[1]), float(bg[2]))
    return t, (tris, mats, tex, cam, params)


def oracle_params(params, thr):
    p = params.copy()
    p[10] = np.float32(thr)     # the facade stores the clamped float threshold
    return p


def test_facade_render_matches_oracle(cuda):
    w = h = 40
    t, (tris, mats, tex, cam, params) = setup_tracer(w, h, spp=16, min_spp=4, thr=0.08)
    frames = []
    t.set_frame_ready_callback(lambda: frames.append(1))
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    got = t.read_accumulator()
    ref = oracle_scene(tris, mats, tex)
    acc_o, passes = O.render_adaptive(ref, cam, oracle_params(params, 0.08), w, h, threads=8)
    compare_accumulators(got, acc_o, "facade")
    assert t.current_sample_count() == passes == len(frames)
    assert t.current_active_pixel_count() == 0
    assert t.is_sample_budget_exhausted()
    assert t.sample_budget_total_iterations() == 16
    t.close()


@pytest.mark.parametrize("per_launch,bvh_mode", [(8, 0), (64, 1)])
def test_facade_passes_per_launch_matches_oracle(cuda, per_launch, bvh_mode):
    """Several full-resolution passes per worker iteration (one wavefront call each, then a
    compaction), on the median tree and on the SAH build with its 4-wide tree: the accumulator
    equals the oracle's per-pass schedule bit for bit."""
    w = h = 40
    t, (tris, mats, tex, cam, params) = setup_tracer(w, h, spp=24, min_spp=4, thr=0.08, bvh_mode=bvh_mode)
    t.set_passes_per_launch(per_launch)
    frames = []
    t.set_frame_ready_callback(lambda: frames.append(1))
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    got = t.read_accumulator()
    ref = oracle_scene(tris, mats, tex)
    acc_o, passes = O.render_adaptive(ref, cam, oracle_params(params, 0.08), w, h, threads=8)
    compare_accumulators(got, acc_o, f"facade/{per_launch} per launch")
    assert t.current_active_pixel_count() == 0
    # the worker launches until no pixel is active: as often as the longest-running pixel needs
    most = int(np.asarray(acc_o.counts).max())
    assert len(frames) == (most + per_launch - 1) // per_launch
    assert t.current_sample_count() == len(frames) * per_launch
    t.close()


def test_facade_preview_levels_then_adaptive(cuda):
    w, h = 48, 40
    t, (tris, mats, tex, cam, params) = setup_tracer(w, h, spp=8, min_spp=4, thr=0.05, preview=2)
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    ref = oracle_scene(tris, mats, tex)
    env = O.Env()
    pp = oracle_params(params, 0.05)
    prev = pp.copy()
    prev[4:5] = np.array([4], np.int32).view(np.float32)   # kPreviewMaxPathDepth
    for level in range(2):
        buf, lw, lh, ds = t.read_preview_level(level)
        assert (lw, lh, ds) == (w // (1 << (2 - level)), h // (1 << (2 - level)), 1 << (2 - level))
        want = np.zeros((lw * lh, 4), np.float32)
        O.lib().orc_preview_sample(O.ptr(want), lw, lh, O.ptr(cam), C.byref(ref.scene), C.byref(env), O.ptr(prev),
                                   level + 1, 8)
        assert np.array_equal(buf.view(np.uint32), want.view(np.uint32)), f"preview level {level}"
    acc_o, passes = O.render_adaptive(ref, cam, pp, w, h, first_iter=2, threads=8)
    compare_accumulators(t.read_accumulator(), acc_o, "facade after preview")
    assert t.current_sample_count() == 2 + passes
    t.close()


def test_facade_region_render(cuda):
    w, h = 40, 36
    t, (tris, mats, tex, cam, params) = setup_tracer(w, h, spp=12, min_spp=4, thr=0.08)
    t.set_region_render_enabled(True)
    t.set_render_region(30, 25, 6, 3)           # corners in any order, clamped to the image
    assert t.render_region() == (6, 3, 30, 25)
    t.reset_accumulation()                       # region reset: only the region is cleared and listed
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    got = t.read_accumulator()
    # oracle: cleared frame, region reset, region passes with region-filtered compaction
    ref = oracle_scene(tris, mats, tex)
    pp = oracle_params(params, 0.08)
    a = O.Accumulator(w, h)
    a.clear()
    L = O.lib()
    active = a.active
    scratch = np.zeros(w * h, np.int32)
    n = L.orc_clear_region(O.ptr(a.acc), O.ptr(a.counts), O.ptr(a.mean), O.ptr(a.m2), O.ptr(a.conv), O.ptr(active), w,
                           6, 3, 30, 25)
    it = 0
    env = O.Env()
    while n > 0:
        L.orc_adaptive_sample(O.ptr(a.acc), O.ptr(a.counts), O.ptr(a.mean), O.ptr(a.m2), O.ptr(a.conv), O.ptr(active),
                              n, w, h, O.ptr(cam), C.byref(ref.scene), C.byref(env), O.ptr(pp), it + 1, 8)
        it += 1
        if it % 4 == 0:
            n = L.orc_compact_active(O.ptr(active), O.ptr(scratch), O.ptr(a.conv), O.ptr(a.counts), 12, n, w, 1, 6, 3,
                                     30, 25)
            active, scratch = scratch, active
    compare_accumulators(got, a, "facade region")
    inside = np.zeros((h, w), bool)
    inside[3:26, 6:31] = True
    assert np.all(got["counts"].reshape(h, w)[~inside] == 0)
    assert t.current_sample_count() == it
    t.close()


def test_facade_display_and_pick(cuda):
    w = h = 32
    t, (tris, mats, tex, cam, params) = setup_tracer(w, h, spp=8, min_spp=8, thr=0.001)
    t.set_physical_camera(1.2, 1.0, 100.0)
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    assert t.publish_display_frame(1)
    got = t.read_display(1)
    acc = t.read_accumulator()
    want = np.zeros((w * h, 4), np.uint8)
    exposure = t.exposure_multiplier()
    O.lib().orc_copy_to_pbo(O.ptr(acc["acc"]), O.ptr(acc["counts"]), O.ptr(acc["conv"]), O.ptr(want), w, h, w, h, 1, 1,
                            O.ptr(oracle_params(params, 0.001)), exposure)
    assert np.array_equal(got, want)
    # pickSurface: closest hit along the camera's forward axis, from the camera position
    ref = oracle_scene(tris, mats, tex)
    pos = cam[0:3]
    target = tris[:, 0:3].mean(0)
    d = (target - pos).astype(np.float32)
    hit = t.pick_surface(tuple(pos), tuple(d))
    dn = (d / np.linalg.norm(d)).astype(np.float32)
    h_o = ref.trace(pos, dn, 0.001, 1.0e30)   # PathTracer.cpp:2005
    assert (hit is not None) == bool(h_o.hit)
    if hit is not None:
        assert np.allclose(hit, pos + dn * np.float32(h_o.t), rtol=1e-6, atol=1e-5)
    t.close()


def test_facade_configure_from_procedural_instances(cuda):
    """configure(w, h, proceduralInstances, meshParams) builds the scene through meshSceneBuild: the
    render equals the oracle's on the mesh ptProceduralBuildScene gives for the same instances, and
    rebuildMeshScene(instances) swaps the scene and resets the accumulation."""
    from lsystemeditor_amd import procedural as P
    w = h = 32
    plant = ("Mat(0) = {0.5, 0.35, 0.2, 0.6}\nMat(Lamp) = {1, 0.95, 0.9, 0.5, 0, 0, 0, 1.5, 0, 30}\n"
             "A(1, 0.12)\nA(l, w) : w > 0.03 -> F(l, w) [ Pitch(30) A(l*0.7, w*0.7) ] Roll(137) [ Pitch(-25) A(l*0.7, w*0.7) ]\n"
             "A(l, w) : w <= 0.03 -> Mat(Lamp) F(l, 0.04)")
    inst = [P.ProceduralInstance(plant, 5, (0.0, -1.0, 0.0), (0.0, 90.0, 0.0)),
            P.ProceduralInstance("F(3, 0.8)", 0, (0.0, -1.8, 1.5), (0.0, 0.0, 0.0))]
    host = P.build_scene(inst, P.default_params(circularSegments=16))
    _, _, _, cam, params = dense_plant(w, h, spp=8, min_spp=4, threshold=0.08, max_depth=6)
    cam = abi.pack_camera((0.0, 0.5, 4.0), abi.quat_look((0.0, -0.1, -1.0)), fov_y=0.9, aspect=1.0, near=0.01)
    t = Tracer()
    assert t.configure_procedural(w, h, inst, circular_segments=16), abi.last_error()
    t.set_camera(cam)
    t.set_max_samples_per_pixel(8)
    t.set_min_samples(4)
    t.set_relative_error_threshold(0.08)
    t.set_max_path_depth(6)
    bg = params[0:3]
    t.set_clear_color(float(bg[0]), float(bg[1]), float(bg[2]))
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    acc_o, passes = O.render_adaptive(oracle_scene(host.triangles, host.materials, None), cam,
                                      oracle_params(params, 0.08), w, h, threads=8)
    compare_accumulators(t.read_accumulator(), acc_o, "facade procedural")
    assert t.current_sample_count() == passes
    # a failing instance is refused with the mesher's message and leaves the scene in place
    assert not t.rebuild_mesh_scene_procedural([P.ProceduralInstance("A -> B\nnot a rule")])
    assert "Procedural instance 0" in abi.last_error()
    # rebuilding with the second instance alone changes the picture
    assert t.rebuild_mesh_scene_procedural(inst[1:], circular_segments=16)
    t.start()
    assert t.wait_until_idle(60000)
    t.stop()
    host2 = P.build_scene(inst[1:], P.default_params(circularSegments=16))
    acc2, _ = O.render_adaptive(oracle_scene(host2.triangles, host2.materials, None), cam,
                                oracle_params(params, 0.08), w, h, threads=8)
    compare_accumulators(t.read_accumulator(), acc2, "facade procedural rebuild")
    t.close()
