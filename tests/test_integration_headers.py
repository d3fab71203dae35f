"""The reference-side engine adapters (integration/HipEngine2D.hpp, integration/HipEngine3D.hpp) compile against
the reference's own headers: a translation unit that uses every method the façade calls, both 2D constructors
(the FP64 mesh and the reference's FP32 MeshT<float>) and the 3D one, the way CurvilinearSolver2D.cpp and
CurvilinearSolver.cpp call CudaEngine / CudaEngine3D, is compiled with g++ -std=c++20 (object code, not linked;
the reference's sources are read where they lie and nothing of them is kept)."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
REF = pathlib.Path("/root/reference")

USE_2D = r'''
#include "HipEngine2D.hpp"
namespace S = Seis2D::Curvilinear;
void drive(const S::Point::Mesh &m64, const S::Point::MeshT<float> &m32, std::vector<S::Receiver> &probes) {
    S::HipEngine a(m64, probes, 0), b(m64, probes, 0, S::Precision::Float32), c(m32, probes, 0);
    for (S::HipEngine *e : {&a, &b, &c}) {
        e->step(0.0, 1e-4, 30.0, 1.0, 0);
        S::Diagnostics d;
        e->diagnostics(d);
        e->rollback();
        e->sample(probes);
        std::vector<float> frame;
        e->frame(0, frame, 10, 10, 0);
        const S::Snapshot s = e->snapshot(true);
        std::vector<std::array<double, 5>> values;
        e->probes({std::size_t(42)}, values);
        (void)s, (void)e->residentBytes();
        e->reset();
    }
}
'''

USE_3D = r'''
#include "HipEngine3D.hpp"
namespace S = Seis3D::Curvilinear;
void drive(const S::Point::Mesh &m, std::vector<S::ReceiverTrace3D> &receivers) {
    S::HipEngine3D a(m, receivers, 0), b(m, receivers, 0, true);
    for (S::HipEngine3D *e : {&a, &b}) {
        e->step(0.0, 1e-3, 8.0, 0);
        S::Diagnostics3D d;
        e->diagnostics(d);
        e->rollback();
        e->sample(receivers);
        S::DisplayFrame3D frame;
        e->slices(S::DisplayRequest3D{1, 2, 3, 0}, frame);
        std::array<std::vector<double>, 9> fields;
        e->download(fields);
        (void)e->residentMemory();
        e->resetFields();
    }
}
'''


@pytest.mark.skipif(not (REF / "include").exists() or shutil.which("g++") is None,
                    reason="reference headers (/root/reference) or g++ not present")
@pytest.mark.parametrize("name,source,sub", [("2d", USE_2D, "two_d"), ("3d", USE_3D, "three_d")])
def test_adapter_compiles_against_reference_headers(name, source, sub, tmp_path):
    src = tmp_path / f"use_{name}.cpp"
    src.write_text(source)
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-c", str(src), "-o", str(tmp_path / "use.o"),
           f"-I{ROOT / 'integration'}", f"-I{ROOT / 'include'}", f"-I{REF / 'include'}", f"-I{REF / 'src' / sub}",
           f"-I{REF / 'src'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
