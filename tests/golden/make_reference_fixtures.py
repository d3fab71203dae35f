"""Writes tests/golden/reference_*.npz: input/output vectors of the reference's own CPU solver.

The reference's 2D and 3D CPU paths are compiled from its sources into oracle/_ref/libswlref.so
(oracle/Makefile, target `ref`).  This script runs that build on the small cases of
tests/test_oracle_2d.py (REF_CASES) and tests/test_oracle_3d.py (CASES) and keeps, per case:
dt, total steps, the energy and peak energy after every step, the receiver traces, and per final
field its L2 norm, sum and max |value|.  The inputs are the case definitions in those test files.

    python tests/golden/make_reference_fixtures.py      (needs oracle/_ref/libswlref.so)
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import conftest  # noqa: E402,F401  registers the package
from cpu_sessions import load  # noqa: E402
from reference_fixtures import (FIXTURES_2D, SEGY_VARIANTS, fixture_path, run_oracle_2d, run_oracle_3d,  # noqa: E402
                                segy_fixture_path, segy_records)
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
    for f in sorted(list(HERE.glob("reference_*.npz")) + list(HERE.glob("segy_*.sgy"))):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
