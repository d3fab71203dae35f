"""FP32 elastic CFS-PML strips updated as vector lanes (engine option pml_vector) (GPU).

Below the surface band (depth >= 4) the absorbing strips of an FP32 elastic CFS-PML model carry no closure,
traction image or reduced stencil, so they are updated two x-adjacent nodes per lane, with vector stencil and
memory loads, instead of one node per lane in the general kernel.  The equations are the general kernel's
(PointKernels2D.hpp:57-64 stretched derivatives, :121-172 RHS; CurvilinearSolver2D.cpp:80-97 RK combine), the
FP32 rounding is not (fused multiply-adds, the interior's accumulator form).  Bars: every receiver gather and
field within 1e-4 relative L2 of the FP64 oracle (the FP32 bar of the solver), and within 2e-5 of the same run
with the strips left general (pml_vector 0), PML memories within 1e-4 of it; the lanes exist (or, for GSLS,
FP64 and sponge-damped models, do not); the automatic choice by grid size.  Cases: layered imported terrain, a moment source, an odd computational
width (the last lane holds one node) and a source inside a strip (the lanes add its body force).
"""
import copy

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, rel_l2
from seiswavelab_curvilinear_amd import Batch2D, BackendKind, Precision, Simulation
from test_oracle_2d import precision_case, ref_case

pytestmark = pytest.mark.gpu

ORACLE_TOL = 1e-4
SAME_TOL = 2e-5
PML_TOL = 1e-4


def case(name):
    if name == "odd_width":
        p = ref_case("sine_terrain_small")
        p.nx, p.receiverCount = 117, 117
        p.surfaceSamples = [(x, z) for x, z in p.surfaceSamples if x <= 2 * 117]
    elif name == "source_in_strip":
        p = ref_case("sine_terrain_small")
        p.sourceX, p.sourceZ = 6, 12
    elif name == "curved_pml":
        p = precision_case("elastic_curved_pml")
        p.maxSteps = 150
    else:
        p = ref_case(name)
    p.precision = Precision.Float32
    return p


def run(p, vector):
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    s.set_option("pml_vector", int(vector))
    ok, err = s.advance(s.totalSteps())
    assert ok, err
    out = {"rec": s.receiver_arrays(), "fields": [s.snapshot_field(f) for f in range(5)], "pml": s.pml(),
           "diag": s.diagnostics(), "timing": s.kernel_timing()}
    s.close()
    return out


def gather(rec):
    return np.concatenate([rec["vx"].ravel(), rec["vz"].ravel()])


@pytest.mark.parametrize("name", ["sine_terrain_small", "moment_external_pml", "curved_pml", "odd_width",
                                  "source_in_strip"])
def test_pml_lanes_match_general_strips_and_oracle(name):
    p = case(name)
    print(f"[pml lanes] {name}", flush=True)
    on, off = run(p, True), run(p, False)
    assert on["timing"]["pml_lanes"] > 0 and off["timing"]["pml_lanes"] == 0
    assert on["timing"]["border_tiles"] < off["timing"]["border_tiles"]
    assert on["diag"]["finite"] and off["diag"]["finite"]
    assert rel_l2(gather(off["rec"]), gather(on["rec"])) <= SAME_TOL
    for f in range(5):
        assert rel_l2(off["fields"][f], on["fields"][f]) <= SAME_TOL, f
    assert rel_l2(off["pml"], on["pml"]) <= PML_TOL
    assert abs(on["diag"]["energy"] / off["diag"]["energy"] - 1) <= SAME_TOL
    q = copy.deepcopy(p)
    q.precision = Precision.Float64
    o = CpuSession2D("orc", q)
    assert o.advance(o.total)[0]
    assert rel_l2(gather(o.receivers()), gather(on["rec"])) <= ORACLE_TOL
    for f in range(5):
        assert rel_l2(o.field(f), on["fields"][f]) <= ORACLE_TOL, f


def test_pml_lanes_automatic_by_grid_size():
    """Automatic choice (pml_vector -1, the default): lanes from 1024 interior tiles on (a 2400 x 900 grid), none
    on small grids, where one node per lane measured faster; switching it on and back rebuilds the lists."""
    from test_step_close_2d import grf_case
    small, large = case("sine_terrain_small"), grf_case(Precision.Float32, False)
    for p, want in ((small, False), (large, True)):
        s = Simulation()
        ok, err = s.initialize(p, BackendKind.Hip, 0)
        assert ok, err
        t = s.kernel_timing()
        assert (t["interior_tiles"] >= 1024) == want and (t["pml_lanes"] > 0) == want
        s.set_option("pml_vector", 1 - int(want))
        assert (s.kernel_timing()["pml_lanes"] > 0) != want
        s.set_option("pml_vector", -1)
        assert s.kernel_timing()["pml_lanes"] == t["pml_lanes"]
        assert s.advance(3)[0] and s.diagnostics()["finite"]
        s.close()


@pytest.mark.parametrize("variant", ["fp64", "gsls", "sponge"])
def test_no_pml_lanes_outside_fp32_elastic_cfs_pml(variant):
    if variant == "fp64":
        p = ref_case("sine_terrain_small")
        p.precision = Precision.Float64
    elif variant == "gsls":
        p = precision_case("gsls1_curved_pml")
        p.precision = Precision.Float32
    else:
        p = ref_case("sponge_tilted_force")
        p.precision = Precision.Float32
    p.maxSteps = 5
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    assert s.advance(5)[0]
    assert s.kernel_timing()["pml_lanes"] == 0
    s.close()


def test_fused_batch_with_pml_lanes_equals_alone():
    """Two FP32 elastic CFS-PML shots of different widths in one fused launch per stage: the PML work items
    reproduce each shot's lone run to the bit (the energy to 1e-12, summation order)."""
    ps = [case("sine_terrain_small"), case("odd_width")]
    ref = [run(copy.deepcopy(p), True) for p in ps]
    batch = Batch2D(0)
    for k, v in {"fused": 1, "graphs": 1, "chunk": 16}.items():
        batch.set_option(k, v)
    members = []
    for p in ps:
        s = Simulation()
        ok, err = s.initialize(copy.deepcopy(p), BackendKind.Hip, 0)
        assert ok, err
        s.set_option("pml_vector", 1)
        batch.add(s)
        members.append(s)
    ok, err = batch.advance(1000)
    assert ok, err
    for s, r in zip(members, ref):
        assert s.completed()
        assert s.kernel_timing()["pml_lanes"] > 0
        rec = s.receiver_arrays()
        assert np.array_equal(rec["vx"], r["rec"]["vx"]) and np.array_equal(rec["vz"], r["rec"]["vz"])
        for f in range(5):
            assert np.array_equal(s.snapshot_field(f), r["fields"][f]), f
        assert np.array_equal(s.pml(), r["pml"])
        assert abs(s.diagnostics()["energy"] - r["diag"]["energy"]) <= 1e-12 * abs(r["diag"]["energy"])
