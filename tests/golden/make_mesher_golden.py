"""Writes mesher_golden.json: SHA-256 digests of the reference turtle's states and of the
reference spline's frames for every case of mesher_cases.py, computed by the reference's
Turtle.cpp / Spline.cpp / LSystemLib compiled unmodified from its sources
(oracle/_ref/libref_procedural.so, `make -C oracle ref`).

Usage: python tests/golden/make_mesher_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle as O  # noqa: E402
from tests.golden.mesher_cases import CASES  # noqa: E402

FRAME_MODES = [(0, 2), (0, 4), (0, 7), (1, 4)]   # (0 = computeLoftFrames, 1 = sampleUniform, samplesPerSpan)


def reference_turtle(R, definition: str, iterations: int, step=500.0, radius=100.0):
    cnt = (C.c_uint32 * 4)()
    d = definition.encode()
    if not R.ref_turtle_execute(d, iterations, step, radius, cnt, None, 0, None, 0, None, 0, None, 0):
        raise RuntimeError(R.ref_last_error().decode())
    st = np.zeros((cnt[1], 7), np.float32)
    sc = np.zeros(cnt[0], np.uint32)
    gr = np.zeros(cnt[2], np.uint32)
    ids = C.create_string_buffer(cnt[3] + 1)
    R.ref_turtle_execute(d, iterations, step, radius, cnt, st.ctypes.data, cnt[1], sc.ctypes.data, cnt[0], ids,
                         cnt[3] + 1, gr.ctypes.data, cnt[2])
    return st, sc, (ids.value.decode().split("\n") if cnt[0] else []), gr


def reference_frames(R, states, tension=1.0, samples=4, mode=0):
    s = np.ascontiguousarray(states, np.float32)
    n = C.c_uint32()
    arc = C.c_float()
    R.ref_spline_frames(s.ctypes.data, s.shape[0], tension, samples, mode, None, 0, C.byref(n), C.byref(arc))
    out = np.zeros((n.value, 13 if mode == 0 else 7), np.float32)
    R.ref_spline_frames(s.ctypes.data, s.shape[0], tension, samples, mode, out.ctypes.data, n.value, C.byref(n),
                        C.byref(arc))
    return out, np.float32(arc.value)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def record(R, definition: str, iterations: int) -> dict:
    st, sc, ids, gr = reference_turtle(R, definition, iterations)
    segs = np.split(st, np.cumsum(sc)[:-1]) if len(sc) else []
    out = {"segments": int(sc.size), "states": int(st.shape[0]), "turtle": digest(st, sc, gr),
           "material_ids": ids, "groups": gr.tolist()}
    for mode, samples in FRAME_MODES:
        parts = [reference_frames(R, s, 1.0, samples, mode) for s in segs]
        out[f"frames_{mode}_{samples}"] = digest(*[p for f, a in parts for p in (f, np.array([a]))])
        out[f"count_{mode}_{samples}"] = int(sum(f.shape[0] for f, _ in parts))
    return out


def main() -> int:
    R = O.ref_procedural()
    if R is None:
        print("oracle/_ref/libref_procedural.so is not built (make -C oracle ref)", file=sys.stderr)
        return 1
    out = {name: record(R, d, it) for name, d, it in CASES}
    (HERE / "mesher_golden.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"{len(out)} cases")
    return 0


if __name__ == "__main__":
    sys.exit(main())
