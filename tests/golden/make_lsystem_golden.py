"""Writes lsystem_golden.json: for every case of lsystem_cases.py, the output of the reference's
L-system library compiled unmodified from its sources (oracle/_ref/libref_procedural.so, built by
`make -C oracle ref` where the reference sources are present), or its error message.

Usage: python tests/golden/make_lsystem_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from tests.golden.lsystem_cases import CASES  # noqa: E402

KINDS = {"modules": 0, "rule": 1, "rules": 2, "system": 3}


def run_reference(R, kind: str, text: str, iterations: int, seed: int) -> dict:
    t = text.encode()
    if kind in KINDS:
        call = lambda b, n: R.ref_lsystem_parse_dump(KINDS[kind], t, b, n)  # noqa: E731
    elif kind == "eval":
        call = lambda b, n: R.ref_lsystem_evaluate_text(t, iterations, seed, 1, b, n)  # noqa: E731
    else:
        call = lambda b, n: R.ref_lsystem_materials_text(t, b, n)  # noqa: E731
    n = call(None, 0)
    if n < 0:
        return {"error": R.ref_last_error().decode()}
    buf = C.create_string_buffer(int(n) + 1)
    call(buf, n + 1)
    return {"output": buf.value.decode()}


def main() -> int:
    R = O.ref_procedural()
    if R is None:
        print("oracle/_ref/libref_procedural.so is not built (make -C oracle ref)", file=sys.stderr)
        return 1
    out = {name: dict(kind=kind, **run_reference(R, kind, text, it, seed)) for name, kind, text, it, seed in CASES}
    (HERE / "lsystem_golden.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"{len(out)} cases")
    return 0


if __name__ == "__main__":
    sys.exit(main())
