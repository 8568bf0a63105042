"""Builds the gfx950 path-tracing library (lsystemeditor_amd/libpt_hip.so) in-tree.

Usage: python -m lsystemeditor_amd.build [--jobs N]

Every translation unit is compiled with hipcc for gfx950 with FP contraction disabled, so
that device arithmetic follows the source's operation order exactly, and packed-FP32
instruction selection disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
OUT = PKG / "libpt_hip.so"
BUILD_DIR = PKG / "_build"

SOURCES = ["pt_render.hip", "pt_wavefront.hip", "pt_scene.hip", "pt_sah.hip", "pt_adjoint.hip", "pt_spectral.hip", "pt_tracer.cpp", "pt_errors.cpp", "pt_lsystem.cpp", "pt_mesher.cpp", "pt_hdr.cpp"]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build the library")
    return path


def _compile(src: str) -> Path:
    obj = BUILD_DIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".hip"):
        cmd[1:1] = DEVICE
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(jobs: int | None = None) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    jobs = jobs or min(8, len(SOURCES))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    out = build(args.jobs)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
