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
from reference_fixtures import FIXTURES_2D, fixture_path, run_oracle_2d, run_oracle_3d  # noqa: E402
from test_oracle_2d import ref_case  # noqa: E402
from test_oracle_3d import CASES, case  # noqa: E402


def main():
    if load("ref") is None:
        sys.exit("oracle/_ref/libswlref.so is missing: build it with `make -C oracle ref` first")
    for name, prec in FIXTURES_2D:
        p = ref_case(name)
        p.precision = prec
        np.savez_compressed(fixture_path("2d", name, prec), **run_oracle_2d("ref", p))
    for name in CASES:
        np.savez_compressed(fixture_path("3d", name), **run_oracle_3d("ref", case(name)))
    for f in sorted(HERE.glob("reference_*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
