"""Multi-shot batch (csrc/batch2d_host.inl, csrc/batch2d.cpp) on the GPU.

A shot stepped inside a batch runs the same stage arithmetic as alone, so its receivers and fields
are bit-identical to the single-session run; only the energy sum is reduced in a different order
(bar 1e-12 relative).  Covers both runner modes (per-shot kernels on several streams; one fused launch per stage over all
shots), mixed precision / GSLS mechanism counts / grid sizes, graph capture and replay, the
persistent-grid autotuner, shots that finish at different steps, and fault isolation.
"""
import copy

import numpy as np
import pytest

from seiswavelab_curvilinear_amd import Batch2D, BackendKind, Parameters, Precision, Simulation
from test_oracle_2d import precision_case, ref_case

pytestmark = pytest.mark.gpu


def shots():
    """Four different shots: elastic FP32, GSLS-3 FP32, elastic FP64 (moment), GSLS-2 FP64."""
    a = ref_case("sine_terrain_small")
    a.precision, a.maxSteps = Precision.Float32, 70
    b = precision_case("gsls3_curved_pml")
    b.precision, b.maxSteps = Precision.Float32, 90
    c = ref_case("moment_external_pml")
    c.precision, c.maxSteps = Precision.Float64, 60
    d = ref_case("gsls2_layered_pml")
    d.precision, d.maxSteps = Precision.Float64, 80
    return [a, b, c, d]


def session(p):
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    return s


def alone(p):
    s = session(p)
    ok, err = s.advance(s.totalSteps())
    assert ok, err
    return s


def same(a, b):
    ra, rb = a.receiver_arrays(), b.receiver_arrays()
    assert a.step() == b.step()
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])
    for f in range(5):
        assert np.array_equal(a.snapshot_field(f), b.snapshot_field(f)), f
    da, db = a.diagnostics(), b.diagnostics()
    assert abs(da["energy"] - db["energy"]) <= 1e-12 * abs(da["energy"])


MODES = {"lanes": {"fused": 0, "lanes": 3, "lane_steps": 5},
         "fused_graphs": {"fused": 1, "graphs": 1, "chunk": 16},
         "fused_direct": {"fused": 1, "graphs": 0, "chunk": 16}}


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_equals_alone(mode):
    ps = shots()
    print(f"[batch] mode={mode}", flush=True)
    ref = [alone(copy.deepcopy(p)) for p in ps]
    batch = Batch2D(0)
    for k, v in MODES[mode].items():
        batch.set_option(k, v)
    members = [session(copy.deepcopy(p)) for p in ps]
    for s in members:
        batch.add(s)
    assert len(batch) == 4
    ok, err = batch.advance(1000)
    assert ok, err
    for s, r in zip(members, ref):
        assert s.completed()
        same(s, r)
    st = batch.stats()
    if mode == "lanes":
        assert st["fused"] == 0 and st["lanes"] == 3 and st["lane_runs"] >= 1
    else:
        assert st["chunks"] >= 5
    if mode == "fused_graphs":
        assert st["captures"] >= 1 and st["replays"] >= 1
    elif mode == "fused_direct":
        assert st["direct_chunks"] == st["chunks"]


def test_parallel_build_equals_alone():
    """Batch2D.build creates the sessions on several host threads: every shot is still bit-identical to running
    it alone, in the order given."""
    ps = shots()
    ref = [alone(copy.deepcopy(p)) for p in ps]
    batch, members = Batch2D.build([copy.deepcopy(p) for p in ps], 0, workers=4)
    assert len(batch) == 4 and len(members) == 4
    ok, err = batch.advance(1000)
    assert ok, err
    for s, r in zip(members, ref):
        assert s.completed()
        same(s, r)
    batch.close()
    bad = copy.deepcopy(ps[0])
    bad.nx = 3  # rejected by the session's validation
    with pytest.raises(RuntimeError, match="shot 1"):
        Batch2D.build([copy.deepcopy(ps[1]), bad], 0, workers=2)


def test_lanes_uneven_calls_equal_alone():
    """Lanes with members finishing at different steps and calls that split runs unevenly."""
    ps = shots()
    ref = [alone(copy.deepcopy(p)) for p in ps]
    batch = Batch2D(0)
    batch.set_option("lanes", 2)
    batch.set_option("lane_steps", 3)
    members = [session(copy.deepcopy(p)) for p in ps]
    for s in members:
        batch.add(s)
    for n in (1, 13, 29, 1000):
        ok, err = batch.advance(n)
        assert ok, err
    for s, r in zip(members, ref):
        same(s, r)


def test_autotuner_choice_does_not_change_results():
    ps = shots()[:2]
    ref = [alone(copy.deepcopy(p)) for p in ps]
    batch = Batch2D(0)
    batch.set_option("fused", 1)
    batch.set_option("chunk", 8)
    members = [session(copy.deepcopy(p)) for p in ps]
    for s in members:
        batch.add(s)
    for n in (13, 40, 1000):  # uneven calls: members finish at different steps
        ok, err = batch.advance(n)
        assert ok, err
    st = batch.stats()
    assert st["grid"] >= 0  # a candidate was chosen after the timed chunks
    for s, r in zip(members, ref):
        same(s, r)


@pytest.mark.parametrize("fused", [0, 1])
def test_fault_stays_with_its_session(fused):
    ps = shots()[:3]
    # density 1e-300 overflows the source term (BackendValidation.cpp:55-62): the first step is rejected
    bad = Parameters(nx=40, nz=40, sourceZ=5, receiverCount=4, density=1e-300, vp=3000, vs=1700, maxSteps=20)
    ref = [alone(copy.deepcopy(p)) for p in ps[1:]]
    batch = Batch2D(0)
    batch.set_option("fused", fused)
    members = [session(bad)] + [session(copy.deepcopy(p)) for p in ps[1:]]
    for s in members:
        batch.add(s)
    ok, err = batch.advance(1000)
    assert ok, err
    assert members[0].fault().startswith("2D step rejected") and members[0].step() == 0
    for s, r in zip(members[1:], ref):
        assert s.fault() == ""
        same(s, r)
    ok, err = members[0].advance(1)
    assert not ok and err == members[0].fault()
