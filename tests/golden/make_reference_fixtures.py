"""Writes tests/golden/reference_*.npz, segy_*.sgy and vts_*.vts.gz: outputs of the reference's own code.

The reference's 2D and 3D CPU paths are compiled from its sources into oracle/_ref/libswlref.so
(oracle/Makefile, target `ref`).  This script runs that build on the small cases of
tests/test_oracle_2d.py (REF_CASES) and tests/test_oracle_3d.py (CASES) and keeps, per case:
dt, total steps, the energy and peak energy after every step, the receiver traces, and per final
field its L2 norm, sum and max |value|.  The inputs are the case definitions in those test files.
It also runs the reference's SEG-Y writer on a synthetic record set, and its model-document and
wavefield-snapshot writers on the small VTS cases of tests/reference_fixtures.py (gzipped text), and its
3D VTS and receiver-CSV writers on VTS3D_CASES.

    python tests/golden/make_reference_fixtures.py      (needs oracle/_ref/libswlref.so)
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import conftest  # noqa: E402,F401  registers the package
from cpu_sessions import load  # noqa: E402
from reference_fixtures import (FIXTURES_2D, FULL_2D, RECEIVERS2D_CASES, RECEIVERS3D_EXPORTS,  # noqa: E402
                                SEGY_VARIANTS, VTS3D_CASES, VTS_ALL_COMPONENTS, VTS_MODEL_VARIANTS,
                                VTS_SNAPSHOT_CASES, fixture_path, receivers2d_fixture_path, receivers3d_fixture_path,
                                reduce_full_2d, run_oracle_2d, run_oracle_3d, segy_fixture_path, segy_records,
                                vts3d_fixture_path, vts_case, vts_fixture_path)
from test_oracle_2d import ref_case  # noqa: E402
from test_oracle_3d import CASES, case  # noqa: E402


def write_segy_fixtures():
    """The reference's writeSegy on the synthetic record set of tests/reference_fixtures.py."""
    import ctypes as C
    lib = load("ref")
    lib.ref_segy_write.restype = C.c_int
    lib.ref_segy_write.argtypes = [C.c_char_p] + [C.c_void_p] * 4 + [C.c_int, C.c_int64, C.c_double, C.c_char_p,
                                                                     C.c_char_p, C.c_int, C.c_double, C.c_double,
                                                                     C.c_char_p, C.c_int]
    pos, src, start, s = (np.ascontiguousarray(a, dtype=np.float64) for a in segy_records())
    for name, (profile, gain, interval, dt, comp, units) in SEGY_VARIANTS.items():
        err = C.create_string_buffer(256)
        ok = lib.ref_segy_write(str(segy_fixture_path(name)).encode(), pos.ctypes.data, src.ctypes.data,
                                start.ctypes.data, s.ctypes.data, s.shape[0], s.shape[1], dt, comp.encode(),
                                units.encode(), profile, gain, interval, err, 256)
        assert ok, err.value


def write_vts_fixtures():
    """The reference's writeModelVts (ComputeSettings3D{1, -1}, CPU session) for each model variant, and its
    writeSnapshot (all seven components and the materials) at step 0 and after VTS_SNAPSHOT_CASES steps."""
    import ctypes as C
    import gzip
    import shutil
    import tempfile
    from cpu_sessions import CpuSession2D
    lib = load("ref")
    lib.ref2d_write_model_vts.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
    lib.ref2d_write_snapshot.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_char_p, C.c_int]
    err = C.create_string_buffer(512)
    comps = (C.c_int * 7)(*VTS_ALL_COMPONENTS)
    with tempfile.TemporaryDirectory() as tmp:
        def keep(src, dst):
            with open(src, "rb") as a, gzip.GzipFile(dst, "wb", mtime=0) as b:
                shutil.copyfileobj(a, b)
        for name in VTS_MODEL_VARIANTS:
            s = CpuSession2D("ref", vts_case(name))
            assert lib.ref2d_write_model_vts(C.c_void_p(s.h), f"{tmp}/m.vts".encode(), 1, -1, err, 512), err.value
            keep(f"{tmp}/m.vts", vts_fixture_path(name))
        for name, steps in VTS_SNAPSHOT_CASES.items():
            s = CpuSession2D("ref", vts_case(name))
            for step in (0, steps):
                ok, e = s.advance(step - s.step)
                assert ok, e
                assert lib.ref2d_write_snapshot(C.c_void_p(s.h), f"{tmp}/s.vts".encode(), comps, 1, err, 512), err.value
                keep(f"{tmp}/s.vts", vts_fixture_path(name, step))


def write_export3d_fixtures():
    """The reference's 3D writeVts (all fields, materials, velocity vector) at step 0 and its writeReceiversCsv
    after maxSteps, on reference CPU sessions of VTS3D_CASES."""
    import ctypes as C
    import gzip
    import shutil
    import tempfile
    from cpu_sessions import CpuSession3D
    lib = load("ref")
    lib.ref3d_write_vts.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
    lib.ref3d_write_receivers_csv.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_char_p, C.c_int]
    lib.ref3d_write_receivers_vz_csv.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    err = C.create_string_buffer(512)
    with tempfile.TemporaryDirectory() as tmp:
        def keep(src, dst):
            with open(src, "rb") as a, gzip.GzipFile(dst, "wb", mtime=0) as b:
                shutil.copyfileobj(a, b)
        for name in VTS3D_CASES:
            p = case(name)
            s = CpuSession3D("ref", p)
            assert lib.ref3d_write_vts(C.c_void_p(s.h), f"{tmp}/s.vts".encode(), (C.c_int * 11)(*[1] * 11), 1, 1,
                                       err, 512), err.value
            keep(f"{tmp}/s.vts", vts3d_fixture_path(name))
            ok, e = s.advance(p.maxSteps)
            assert ok, e
            if RECEIVERS3D_EXPORTS[name] == "vz":
                ok = lib.ref3d_write_receivers_vz_csv(C.c_void_p(s.h), f"{tmp}/r.csv".encode(), 0, err, 512)
            else:
                ok = lib.ref3d_write_receivers_csv(C.c_void_p(s.h), f"{tmp}/r.csv".encode(), (C.c_int * 3)(1, 0, 1),
                                                   err, 512)
            assert ok, err.value
            keep(f"{tmp}/r.csv", receivers3d_fixture_path(name))
            s.close()


def write_receivers2d_fixtures():
    """The reference's writeReceivers after each RECEIVERS2D_CASES case's full run (reference CPU session)."""
    import ctypes as C
    import gzip
    import shutil
    import tempfile
    from cpu_sessions import CpuSession2D
    lib = load("ref")
    lib.ref2d_write_receivers.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
    err = C.create_string_buffer(512)
    with tempfile.TemporaryDirectory() as tmp:
        for name in RECEIVERS2D_CASES:
            s = CpuSession2D("ref", ref_case(name))
            ok, e = s.advance(s.total)
            assert ok, e
            assert lib.ref2d_write_receivers(C.c_void_p(s.h), f"{tmp}/r.csv".encode(), err, 512), err.value
            with open(f"{tmp}/r.csv", "rb") as a, gzip.GzipFile(receivers2d_fixture_path(name), "wb", mtime=0) as b:
                shutil.copyfileobj(a, b)
            s.close()


def write_full_2d_fixture():
    """The 500x200 sine-terrain workload (seed 1) in FP64 for all its 2000 steps on the reference build."""
    from seiswavelab_curvilinear_amd import workloads
    p = workloads.sine_terrain_500x200(seed=1)
    p.cpuThreads = 8
    np.savez_compressed(fixture_path("2d", FULL_2D), **reduce_full_2d(run_oracle_2d("ref", p)))


def main():
    if load("ref") is None:
        sys.exit("oracle/_ref/libswlref.so is missing: build it with `make -C oracle ref` first")
    for name, prec in FIXTURES_2D:
        p = ref_case(name)
        p.precision = prec
        np.savez_compressed(fixture_path("2d", name, prec), **run_oracle_2d("ref", p))
    for name in CASES:
        np.savez_compressed(fixture_path("3d", name), **run_oracle_3d("ref", case(name)))
    write_segy_fixtures()
    write_vts_fixtures()
    write_export3d_fixtures()
    write_receivers2d_fixtures()
    write_full_2d_fixture()
    for f in sorted(list(HERE.glob("reference_*.npz")) + list(HERE.glob("segy_*.sgy")) + list(HERE.glob("*.gz"))):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
