"""Python handle on the C++ renderer facade (include/pt_tracer.h, class PathTracer).

Method names follow PathTracerLib/src/PathTracer.h:30-152 in snake_case; the worker thread,
render loop, clamps and reset semantics live in the C++ library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class Tracer:
    def __init__(self):
        self._lib = abi.lib()
        self.handle = self._lib.ptTracerCreate()
        if not self.handle:
            raise RuntimeError("ptTracerCreate failed")
        self.width = self.height = 0
        self._callback = None

    def close(self):
        if self.handle:
            self._lib.ptTracerDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- configuration / lifecycle ----------------------------------------------------------
    def configure(self, width: int, height: int, triangles: np.ndarray, materials: np.ndarray,
                  textures: np.ndarray | None = None) -> bool:
        tris = np.ascontiguousarray(triangles, dtype=np.float32)
        mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, abi.MATERIAL_WORDS)
        tex = None if textures is None else np.ascontiguousarray(textures, dtype=np.float32).reshape(
            -1, abi.TEXTURE_WORDS)
        ok = self._lib.ptTracerConfigure(self.handle, width, height, tris.ctypes.data if tris.size else None,
                                         tris.shape[0], mats.ctypes.data if mats.size else None, mats.shape[0],
                                         None if tex is None else tex.ctypes.data, 0 if tex is None else tex.shape[0])
        if ok:
            self.width, self.height = width, height
        return bool(ok)

    def configure_procedural(self, width: int, height: int, instances, circular_segments: int = 64,
                             crease_angle_deg: float = 50.0) -> bool:
        """configure(w, h, proceduralInstances, meshParams) (PathTracer.h:44-50): the scene is built
        from procedural.ProceduralInstance records by the library's meshSceneBuild."""
        from . import procedural
        arr, _keep = procedural._instances(instances)
        mp = procedural.MeshSceneParams(circular_segments, crease_angle_deg)
        ok = self._lib.ptTracerConfigureProcedural(self.handle, width, height, arr, len(instances), C.byref(mp))
        if ok:
            self.width, self.height = width, height
        return bool(ok)

    def rebuild_mesh_scene_procedural(self, instances, circular_segments: int = 64,
                                      crease_angle_deg: float = 50.0) -> bool:
        from . import procedural
        arr, _keep = procedural._instances(instances)
        mp = procedural.MeshSceneParams(circular_segments, crease_angle_deg)
        return bool(self._lib.ptTracerRebuildMeshSceneProcedural(self.handle, arr, len(instances), C.byref(mp)))

    def set_frame_ready_callback(self, fn):
        if fn is None:
            self._callback = None
            self._lib.ptTracerSetFrameReadyCallback(self.handle, None, None)
            return
        proto = C.CFUNCTYPE(None, C.c_void_p)
        self._callback = proto(lambda _user: fn())
        self._lib.ptTracerSetFrameReadyCallback(self.handle, C.cast(self._callback, C.c_void_p), None)

    def start(self):
        self._lib.ptTracerStart(self.handle)

    def stop(self):
        self._lib.ptTracerStop(self.handle)

    def reset_accumulation(self):
        self._lib.ptTracerResetAccumulation(self.handle)

    def reset_region_accumulation(self):
        self._lib.ptTracerResetRegionAccumulation(self.handle)

    def is_running(self) -> bool:
        return bool(self._lib.ptTracerIsRunning(self.handle))

    def wait_until_idle(self, timeout_ms: int = 60000) -> bool:
        return bool(self._lib.ptTracerWaitUntilIdle(self.handle, timeout_ms))

    # ---- display ----------------------------------------------------------------------------
    def publish_display_frame(self, slot: int = 0) -> bool:
        return bool(self._lib.ptTracerPublishDisplayFrame(self.handle, slot))

    def read_display(self, slot: int = 0) -> np.ndarray:
        out = np.zeros((self.width * self.height, 4), np.uint8)
        abi.check(self._lib.ptTracerReadDisplay(self.handle, slot, out.ctypes.data), "ptTracerReadDisplay")
        return out

    def read_accumulator(self) -> dict[str, np.ndarray]:
        n = self.width * self.height
        acc = np.zeros((n, 4), np.float32)
        counts = np.zeros(n, np.uint32)
        mean = np.zeros(n, np.float32)
        m2 = np.zeros(n, np.float32)
        conv = np.zeros(n, np.uint8)
        abi.check(self._lib.ptTracerReadAccumulator(self.handle, acc.ctypes.data, counts.ctypes.data,
                                                    mean.ctypes.data, m2.ctypes.data, conv.ctypes.data),
                  "ptTracerReadAccumulator")
        return {"acc": acc, "counts": counts, "mean": mean, "m2": m2, "conv": conv}

    def read_preview_level(self, level: int):
        w, h, ds = C.c_int(0), C.c_int(0), C.c_int(0)
        abi.check(self._lib.ptTracerReadPreviewLevel(self.handle, level, None, C.byref(w), C.byref(h), C.byref(ds)),
                  "ptTracerReadPreviewLevel")
        buf = np.zeros((w.value * h.value, 4), np.float32)
        abi.check(self._lib.ptTracerReadPreviewLevel(self.handle, level, buf.ctypes.data, None, None, None),
                  "ptTracerReadPreviewLevel")
        return buf, w.value, h.value, ds.value

    # ---- settings (same clamps as PathTracer.cpp) ------------------------------------------
    def set_camera(self, camera_words: np.ndarray):
        cam = np.ascontiguousarray(camera_words, dtype=np.float32)
        self._lib.ptTracerSetCamera(self.handle, cam.ctypes.data)

    def last_sample_camera(self) -> np.ndarray:
        cam = np.zeros(abi.CAMERA_WORDS, np.float32)
        self._lib.ptTracerLastSampleCamera(self.handle, cam.ctypes.data)
        return cam

    def pick_surface(self, ro, rd):
        hit = abi.Vec3C()
        ok = self._lib.ptTracerPickSurface(self.handle, abi.Vec3C(*ro), abi.Vec3C(*rd), C.byref(hit))
        return (hit.x, hit.y, hit.z) if ok else None

    def set_environment_image(self, rgb: np.ndarray | None) -> bool:
        if rgb is None:
            return bool(self._lib.ptTracerSetEnvironmentImage(self.handle, None, 0, 0))
        img = np.ascontiguousarray(rgb, dtype=np.float32)
        return bool(self._lib.ptTracerSetEnvironmentImage(self.handle, img.ctypes.data, img.shape[1], img.shape[0]))

    def set_focus_point(self, p):
        self._lib.ptTracerSetFocusPoint(self.handle, abi.Vec3C(*p))

    def render_region(self):
        vals = [C.c_int(0) for _ in range(4)]
        self._lib.ptTracerRenderRegion(self.handle, *[C.byref(v) for v in vals])
        return tuple(v.value for v in vals)

    def __getattr__(self, name):
        """Plain setters/getters: set_max_samples_per_pixel -> ptTracerSetMaxSamplesPerPixel, etc."""
        if name.startswith("_"):
            raise AttributeError(name)
        camel = "".join(p.capitalize() for p in name.split("_"))
        fn = getattr(abi.lib(), "ptTracer" + camel, None)
        if fn is None:
            raise AttributeError(name)
        handle = self.handle
        return lambda *args: fn(handle, *args)
