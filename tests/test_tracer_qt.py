"""The Qt / glm adapter (include/pt_tracer_qt.h) compiles against real Qt headers and keeps the
reference's signatures (PathTracer.h:30-152): configure(w, h, pbo0, pbo1, instances, params),
rebuildMeshScene(instances, params), setClearColor(QColor), renderRegion() -> QRect,
setEnvironmentHdrPath / environmentHdrPath. Built with g++ against the Qt 5 found in the image
(skipped when it is absent); the program runs without a GPU (nothing it calls launches work)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
QT = Path("/opt/conda/include/qt")
QT_LIB = Path("/opt/conda/lib")

PROGRAM = r"""
#include "pt_tracer_qt.h"
#include <cstdio>
int main(int argc, char** argv) {
    ptqt::PathTracer t;
    t.setClearColor(QColor());
    t.setClearColor(QColor(255, 128, 0));
    t.setRenderRegion(3, 4, 1, 2);
    const QRect r = t.renderRegion();
    const PtRegion p = t.::PathTracer::renderRegion();
    std::printf("region %d %d %d %d plain %d %d %d %d\n", r.left(), r.top(), r.right(), r.bottom(),
                p.minX, p.minY, p.maxX, p.maxY);
    t.setEnvironmentHdrPath(QString::fromUtf8(argv[1]));
    std::printf("path [%s]\n", t.environmentHdrPath().toUtf8().constData());
    std::vector<ptqt::ProceduralInstance> inst(1);
    inst[0].commandString = "F";
    ptqt::MeshSceneBuildParams mp;
    std::printf("defaults %d %.1f %zu\n", mp.circularSegments, mp.creaseAngleDeg, inst[0].iterations);
    const bool configured = t.configure(0, 0, 7, 8, inst, mp);
    std::printf("configure %d pbo %u %u\n", configured ? 1 : 0, t.glPbo(0), t.glPbo(1));
    std::printf("rebuild %d\n", t.rebuildMeshScene(inst) ? 1 : 0);
    return 0;
}
"""


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not (QT / "QtGui" / "QColor").exists() or not (QT_LIB / "libQt5Core.so.5").exists():
        pytest.skip("Qt 5 headers or libraries not found")
    d = tmp_path_factory.mktemp("qt")
    src = d / "main.cpp"
    src.write_text(PROGRAM)
    exe = d / "main"
    pkg = ROOT / "lsystemeditor_amd"
    cmd = ["g++", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT / 'include'}",
           f"-I{QT}", f"-I{QT / 'QtCore'}", f"-I{QT / 'QtGui'}", str(src), "-o", str(exe), f"-L{pkg}", "-lpt_hip",
           str(QT_LIB / "libQt5Gui.so.5"), str(QT_LIB / "libQt5Core.so.5"),
           f"-Wl,-rpath,{pkg}:/usr/lib/x86_64-linux-gnu:{QT_LIB}"]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert proc.returncode == 0, proc.stderr
    return exe


def test_adapter_builds_and_maps_types(program, tmp_path):
    from tests.test_hdr import rgbe_pixels, write_hdr
    hdr = write_hdr(tmp_path / "sky.hdr", rgbe_pixels(np.random.default_rng(1), 4, 16))
    out = subprocess.run([str(program), f"  {hdr}  "], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    region = lines[0].split()
    assert region[1:5] == region[6:10]               # QRect corners are the plain region's
    assert lines[1] == f"path [{hdr}]"              # trimmed, as the reference stores it
    assert lines[2] == "defaults 64 50.0 0"          # MeshSceneBuildParams / ProceduralInstance defaults
    assert lines[3] == "configure 0 pbo 7 8"         # invalid size refused; the PBO names are kept
    assert lines[4] == "rebuild 0"                   # nothing configured
