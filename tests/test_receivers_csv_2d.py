"""2D receiver CSV (writeReceivers, src/two_d/CurvilinearSolver2D.cpp:811-854) against files the reference build
wrote after each case's full run (tests/golden/receivers2d_*.csv.gz).

The backend line names the device that computed the traces and is the one line allowed to differ.  The CPU part
checks the provenance header and the t = 0 row of every receiver from a device-free plan; the GPU part steps an
FP64 HIP session through the case and checks every line: header lines and the receiver, x, z and time columns
byte for byte, the velocity columns to 1e-10 of the record's peak (the path tolerance)."""
import gzip

import numpy as np
import pytest

from reference_fixtures import RECEIVERS2D_CASES, receivers2d_fixture_path
from seiswavelab_curvilinear_amd import BackendKind, Precision, Simulation, writeReceivers
from test_oracle_2d import ref_case


def reference_lines(name):
    with gzip.open(receivers2d_fixture_path(name), "rt") as f:
        return f.read().split("\n")


def split(lines):
    head = [l for l in lines if l.startswith("#") or l.startswith("receiver,")]
    return head, [l for l in lines[len(head):] if l]


def compare_headers(ours, theirs):
    assert len(ours) == len(theirs)
    for a, b in zip(ours, theirs):
        if b.startswith("# backend="):
            assert a.startswith("# backend=HIP ")
        else:
            assert a == b


@pytest.mark.parametrize("name", RECEIVERS2D_CASES)
def test_receiver_csv_header_and_first_rows(name, tmp_path):
    s = Simulation()
    ok, err = s.initialize(ref_case(name), BackendKind.Cpu)
    assert ok, err
    ok, err = writeReceivers(s, tmp_path / "r.csv")
    assert ok, err
    ours_head, ours_rows = split((tmp_path / "r.csv").read_text().split("\n"))
    ref_head, ref_rows = split(reference_lines(name))
    compare_headers(ours_head, ref_head)
    per = len(ref_rows) // len(ours_rows)  # the reference's records hold the full run, ours t = 0 only
    assert per * len(ours_rows) == len(ref_rows)
    assert ours_rows == ref_rows[::per]


@pytest.mark.gpu
@pytest.mark.parametrize("name", RECEIVERS2D_CASES)
def test_receiver_csv_after_stepping(name, tmp_path):
    p = ref_case(name)
    p.precision = Precision.Float64
    print(f"[receiver csv] {name}", flush=True)
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    assert s.advance(s.totalSteps())[0]
    ok, err = writeReceivers(s, tmp_path / "r.csv")
    assert ok, err
    ours_head, ours_rows = split((tmp_path / "r.csv").read_text().split("\n"))
    ref_head, ref_rows = split(reference_lines(name))
    compare_headers(ours_head, ref_head)
    assert len(ours_rows) == len(ref_rows)
    a = [r.split(",") for r in ours_rows]
    b = [r.split(",") for r in ref_rows]
    assert [x[:4] for x in a] == [x[:4] for x in b]
    va = np.array([[float(x[4]), float(x[5])] for x in a])
    vb = np.array([[float(x[4]), float(x[5])] for x in b])
    assert np.abs(vb).max() > 0
    assert np.abs(va - vb).max() <= 1e-10 * np.abs(vb).max()
