"""The C-ABI library: loads without a GPU, exports every declared symbol, and rejects invalid
arguments before launching anything (the validation order of PathTracer.cu:589-936)."""
import ctypes as C

import numpy as np
import pytest

from lsystemeditor_amd import abi


@pytest.fixture(scope="module")
def L():
    return abi.lib()


def test_exports_every_declared_symbol(L):
    names = abi.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ten launcher entry points of PathTracerCuda.h are all present under their reference names
    for n in ("pathTracerClearAccumulator", "pathTracerClearRegionAccumulator", "pathTracerRebuildActiveList",
              "pathTracerAdaptiveSample", "pathTracerCompactActiveList", "pathTracerPreviewSample",
              "pathTracerCopyToPbo", "pathTracerUpsamplePreviewToPbo", "pathTracerWriteUvDebugToPbo",
              "spectralUploadStateToSymbol"):
        assert n in names


FAKE = 0x1000  # never dereferenced: every call below fails validation first


@pytest.mark.parametrize("call,args,msg", [
    ("pathTracerClearAccumulator", [None, FAKE, FAKE, FAKE, FAKE, FAKE, FAKE, FAKE, 4, 4, None], "invalid"),
    ("pathTracerClearAccumulator", [FAKE] * 8 + [0, 4, None], "invalid"),
    ("pathTracerClearRegionAccumulator", [FAKE] * 7 + [4, 4, 3, 0, 1, 2, None], "invalid"),   # minX > maxX
    ("pathTracerRebuildActiveList", [FAKE, FAKE, FAKE, None, 4, 4, 16, None], "invalid"),
    ("pathTracerAdaptiveSample", [FAKE] * 6 + [0, 4, 4, FAKE, FAKE, None, FAKE, 1, None], "invalid"),
    ("pathTracerAdaptiveSample", [FAKE] * 6 + [16, 4, 4, FAKE, None, None, FAKE, 1, None], "no scene"),
    ("pathTracerAdaptiveSample", [FAKE] * 6 + [16, 4, 4, FAKE, FAKE, None, None, 1, None], "invalid"),
    ("pathTracerCompactActiveList", [FAKE] * 5 + [16, 0, 4, 4, False, 0, 0, 0, 0, None], "invalid"),
    ("pathTracerCompactActiveListNoAlloc", [FAKE] * 5 + [16, 0, 4, 4, False, 0, 0, 0, 0, None], "invalid"),
    ("pathTracerCompactActiveListNoAlloc", [FAKE] * 5 + [16, 65536 * 1024 + 1, 8192, 8193, False, 0, 0, 0, 0, None],
     "exceed the maximum"),
    ("pathTracerPreviewSample", [FAKE, 4, 4, None, FAKE, None, FAKE, 1, None], "invalid"),
    ("pathTracerCopyToPbo", [FAKE, FAKE, FAKE, FAKE, 4, 4, None, 1.0, None], "invalid"),
    ("pathTracerUpsamplePreviewToPbo", [FAKE, 0, 4, 2, FAKE, 8, 8, FAKE, 1.0, None], "invalid"),
    ("pathTracerWriteUvDebugToPbo", [FAKE, None, FAKE, 4, 4, FAKE, None], "invalid"),
    ("pathTracerAdaptiveRenderPasses", [FAKE] * 6 + [16, 4, 4, FAKE, FAKE, None, FAKE, 1, 0, None], "invalid"),
    ("pathTracerTraceRays", [None, FAKE, FAKE, 4, 0.0, 1.0, FAKE, None], "invalid"),
])
def test_argument_validation(L, call, args, msg):
    assert getattr(L, call)(*args) is False
    assert msg in abi.last_error()


def test_scene_handle_lifecycle_without_gpu(L):
    h = L.ptSceneCreate()
    assert h
    assert L.ptSceneDevice(h) is None
    cnt, root = C.c_uint32(7), C.c_uint32(7)
    assert L.ptSceneReadBvh(h, None, 0, C.byref(cnt), C.byref(root)) and cnt.value == 0
    assert not L.ptSceneDescription(h, None)
    L.ptSceneDestroy(h)
    assert not L.ptSceneBuild(None, None, 0, None, 0, None, 0, None)


def test_record_layouts():
    assert abi.pack_material().nbytes == 96
    assert abi.pack_camera((0, 0, 0)).nbytes == 80
    assert abi.pack_params().nbytes == 44
    t = abi.pack_host_triangles(np.zeros((3, 3)), np.ones((3, 3)), np.ones((3, 3)))
    assert t.shape == (3, 25) and t.dtype == np.float32
    p = abi.pack_params(max_path_depth=8, max_samples_per_pixel=256)
    assert p[4:5].view(np.int32)[0] == 8 and p[7:8].view(np.int32)[0] == 256
    # reference defaults (RenderTypes.h:15-29)
    d = abi.pack_params()
    assert np.allclose(d[:4], [10 / 255, 10 / 255, 10 / 255, 1.0])
    assert list(d[4:10].view(np.int32)) == [32, 3, 16, 1024, 0, 0] and np.isclose(d[10], 0.02)


def test_rebuild_entry_points_validate_without_gpu(L):
    """The rebuild entry points reject what they cannot rebuild before any device call: a workspace
    of no triangles, null handles, and a scene that has no binned-SAH build (an empty scene)."""
    assert not L.ptBuildWorkspaceCreate(0)
    assert "capacity 0" in abi.last_error()
    assert L.ptSceneRebuild(None, None, None) is False
    assert "null scene or workspace" in abi.last_error()
    assert L.ptSceneRebuildCheck(None, None, None) is False
    h = L.ptSceneCreate()
    assert L.ptSceneRebuild(h, FAKE, None) is False
    assert "binned-SAH" in abi.last_error()
    assert L.ptSceneRebuildCheck(h, FAKE, None) is False
    L.ptSceneDestroy(h)
    L.ptBuildWorkspaceDestroy(None)   # no-op


def test_light_sampling_mode_validation_and_adjoint_refusal(L):
    """pathTracerSetLightSampling takes 0 or 1; the adjoint and detached renders refuse mode 1 before
    any device call."""
    assert L.pathTracerLightSampling() == 0
    assert L.pathTracerSetLightSampling(2) is False and "mode must be 0" in abi.last_error()
    with abi.LightSampling(1):
        assert L.pathTracerLightSampling() == 1
        assert L.pathTracerAdjointRender(FAKE, 4, 4, 2, FAKE, FAKE, None, FAKE, 2, None, FAKE, None) is False
        assert "reference light sampling only" in abi.last_error()
        assert L.pathTracerDetachedRender(FAKE, 4, 4, 2, FAKE, FAKE, FAKE, None, FAKE, None) is False
        assert "reference light sampling only" in abi.last_error()
    assert L.pathTracerLightSampling() == 0


def test_workspace_pins_refuse_growth_without_gpu(L):
    """While graph captures pin the shared wavefront workspace, a call that would need a larger one
    fails with a clear error before any device allocation; pins are released explicitly."""
    assert L.pathTracerWavefrontPins() == 0
    assert L.pathTracerWavefrontPin()
    assert L.pathTracerWavefrontPins() == 1
    assert L.pathTracerWavefrontReserve(1 << 20, 4096, 8) is False
    msg = abi.last_error()
    assert "graph capture(s) hold it" in msg and "pathTracerWavefrontUnpin" in msg
    assert L.pathTracerWavefrontUnpin(2) is False and "2 pins requested, 1 held" in abi.last_error()
    assert L.pathTracerWavefrontUnpin(1)
    assert L.pathTracerWavefrontPins() == 0
    assert L.pathTracerWavefrontUnpin(-1) is False
