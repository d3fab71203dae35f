"""The fault rule's late-energy-growth branch, driven from the host (GPU).

The reference rejects a step when the fields are nonfinite or, after the source window, when the energy
exceeds 1e4 times the peak it compares against (2D: the accepted peak energy,
src/two_d/CurvilinearSolver2D.cpp:576-590; 3D: the peak over (s+1) dt <= 3/f,
src/three_d/CurvilinearSolver.cpp:477-488).  A stable model never grows 1e4-fold, so the tests replace
that peak with a tiny positive value through the sessions' override entry points (and the oracle's
test hook) and check, on both sides: no rejection for the last step inside the window, rejection of the
first step after it, the state kept at the last accepted step, the latched message, and that a peak of
zero disables the branch.  Single sessions, the multi-shot batch and 3D are covered.
"""
import math

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, CpuSession3D, rel_l2
from seiswavelab_curvilinear_amd import Batch2D, BackendKind, Precision, Simulation, Simulation3D
from test_oracle_2d import ref_case
from test_oracle_3d import case

pytestmark = pytest.mark.gpu

TINY = 1e-300
MSG2D = "2D step rejected: nonfinite fields or excessive late energy growth. Rebuild the session."
MSG3D = "Energy grew above 10,000 times the source-window peak."


def first_late_step(dt, frequency):
    """Smallest step index s whose step ends after the source window: (s+1)*dt > 3/f."""
    s = max(0, int(3 / frequency / dt) - 2)
    while (s + 1) * dt <= 3 / frequency:
        s += 1
    return s


def case2d(prec):
    p = ref_case("moment_external_pml")
    p.precision = prec
    p.maxSteps = 400
    return p


@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_energy_growth_rule_2d_matches_oracle(prec):
    p = case2d(prec)
    print(f"[fault 2d] {prec.name}", flush=True)
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    o = CpuSession2D("orc", p)
    late = first_late_step(s.dt(), p.frequency)
    assert 0 < late < s.totalSteps() - 5
    assert s.advance(late - 1)[0] and o.advance(late - 1)[0]
    # last step inside the window: a tiny peak does not reject it
    s.override_peak_energy(TINY), o.set_peak(TINY)
    assert s.advance(1)[0] and o.advance(1)[0]
    assert s.step() == o.step == late
    # first step after the window: rejected on both sides, state kept, fault latched
    before = [s.snapshot_field(f) for f in range(5)]
    s.override_peak_energy(TINY), o.set_peak(TINY)
    ok, err = s.advance(3)
    ok_o, err_o = o.advance(3)
    assert not ok and not ok_o and err == err_o == MSG2D
    assert s.step() == o.step == late
    assert s.fault() == MSG2D
    for f in range(5):
        assert np.array_equal(s.snapshot_field(f), before[f]), f
        bar = 1e-10 if prec == Precision.Float64 else 1e-4
        assert rel_l2(o.field(f), before[f]) <= bar, f
    assert not s.advance(1)[0]
    # a peak of zero disables the growth branch (peak > 0 is part of the rule)
    assert s.reset()[0]
    assert s.advance(late)[0]
    s.override_peak_energy(0.0)
    assert s.advance(3)[0] and s.step() == late + 3


def test_energy_growth_rule_in_a_batch_faults_one_member_only():
    """Batch2D passes each member's own peak: the overridden member is rejected at its first late step and
    latches; the other members run on, bit-identical to running alone."""
    params = [case2d(Precision.Float32), case2d(Precision.Float32), case2d(Precision.Float64)]
    params[1].sourceX = 30
    alone = []
    for p in params:
        s = Simulation()
        assert s.initialize(p, BackendKind.Hip, 0)[0]
        alone.append(s)
    late = first_late_step(alone[0].dt(), params[0].frequency)
    for s in alone[1:]:
        assert s.advance(late + 5)[0]
    sims = []
    for p in params:
        s = Simulation()
        assert s.initialize(p, BackendKind.Hip, 0)[0]
        sims.append(s)
    batch = Batch2D(0)
    for s in sims:
        batch.add(s)
    batch.advance(late)
    sims[0].override_peak_energy(TINY)
    batch.advance(5)
    assert sims[0].step() == late and sims[0].fault() == MSG2D
    for s, a in zip(sims[1:], alone[1:]):
        assert s.step() == late + 5 and s.fault() == ""
        r, q = s.receiver_arrays(), a.receiver_arrays()
        assert np.array_equal(r["vx"], q["vx"]) and np.array_equal(r["vz"], q["vz"])


@pytest.mark.parametrize("precision", [0, 1])
def test_energy_growth_rule_3d_matches_oracle(precision):
    p = case("analytic_sponge_force")
    p.maxSteps = 400
    print(f"[fault 3d] precision {precision}", flush=True)
    s = Simulation3D(BackendKind.Hip)
    ok, err = s.initialize(p, 0, precision)
    assert ok, err
    o = CpuSession3D("orc", p)
    late = first_late_step(s.dt(), p.dominantFrequency)
    assert 0 < late < s.totalSteps() - 5
    assert s.advance(late - 1)[0] and o.advance(late - 1)[0]
    # inside the window the source-window peak is still being raised by the step's own energy
    s.override_source_peak(TINY), o.set_source_peak(TINY)
    assert s.advance(1)[0] and o.advance(1)[0]
    assert s.currentStep() == o.step == late
    before = [s.field(f) for f in range(9)]
    s.override_source_peak(TINY), o.set_source_peak(TINY)
    ok, err = s.advance(3)
    ok_o, err_o = o.advance(3)
    assert not ok and not ok_o
    assert err.startswith(MSG3D) and err_o.startswith(MSG3D)
    assert s.currentStep() == o.step == late
    for f in range(9):
        assert np.array_equal(s.field(f), before[f]), f
        assert rel_l2(o.field(f), before[f]) <= (1e-10 if precision == 0 else 1e-4), f
    assert not s.advance(1)[0]
