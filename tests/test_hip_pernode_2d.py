"""Per-node materials on the HIP path (GPU).

Layered models take per-row material tables in the stage kernel; heterogeneous models (the
5001x2001 GRF-terrain shot, every served request) read lambda, mu, rho and the GSLS dP, dS per node.
These tests run the per-node branch against the oracle (whose per-node branch is pinned in
test_oracle_2d.py) on small GRF-terrain models with 5 % velocity perturbations, elastic and GSLS-3,
and check that forcing the per-node arrays on a layered model changes nothing, bit for bit.
Bars: FP64 <= 1e-10 relative L2, FP32 <= 1e-4 relative L2 against the FP64 oracle.
"""
import numpy as np
import pytest

from cpu_sessions import CpuSession2D, rel_l2
from seiswavelab_curvilinear_amd import BackendKind, GslsSettings, Precision, Simulation, workloads
from test_oracle_2d import precision_case, ref_case

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-4


def grf_case(kind, prec, seed=11, steps=300):
    """160x90 logical nodes of the GRF-terrain workload (30-node external CFS-PML: 220x120 computational
    nodes), explosive source 5 nodes deep, receivers every 5 nodes."""
    p = workloads.grf_terrain_5001x2001(seed=seed, precision=prec, nx=160, nz=90, steps=steps, topo_rms=6.0).params
    if kind == "gsls3":
        p.attenuation = GslsSettings(enabled=True, mechanisms=3, minFrequency=6.0, maxFrequency=90.0,
                                     referenceFrequency=30.0, qp=40.0, qs=20.0)
    p.cpuThreads = 8
    return p


def gather(r):
    return np.concatenate([r["vx"].ravel(), r["vz"].ravel()])


def hip(p, options=None):
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    for k, v in (options or {}).items():
        s.set_option(k, v)
    return s


@pytest.mark.parametrize("kind", ["elastic", "gsls3"])
def test_pernode_fp64_matches_oracle(kind):
    p = grf_case(kind, Precision.Float64)
    print(f"[pernode] {kind} FP64", flush=True)
    s = hip(p)
    assert s.kernel_timing()["material_by_row"] is False
    o = CpuSession2D("orc", p)
    assert s.dt() == o.dt and s.totalSteps() == o.total
    assert np.array_equal(s.probe_indices(), o.probe_indices())
    assert s.advance(s.totalSteps())[0] and o.advance(o.total)[0]
    a, b = s.receiver_arrays(), o.receivers()
    assert a["vx"].shape == b["vx"].shape
    assert np.abs(gather(b)).max() > 0
    assert rel_l2(gather(b), gather(a)) <= FP64_TOL
    for f in range(s.layout()["fields"]):
        assert rel_l2(o.field(f), s.snapshot_field(f)) <= FP64_TOL, f
    assert rel_l2(o.pml(), s.pml()) <= FP64_TOL
    assert abs(s.diagnostics()["energy"] / o.diagnostics()["energy"] - 1) < 1e-12


@pytest.mark.parametrize("kind", ["elastic", "gsls3"])
def test_pernode_fp32_within_tolerance_of_fp64_oracle(kind):
    p = grf_case(kind, Precision.Float32)
    print(f"[pernode] {kind} FP32", flush=True)
    s = hip(p)
    assert s.kernel_timing()["material_by_row"] is False
    assert s.advance(s.totalSteps())[0]
    p.precision = Precision.Float64
    o = CpuSession2D("orc", p)
    assert o.advance(o.total)[0]
    a, b = s.receiver_arrays(), o.receivers()
    assert a["vx"].shape == b["vx"].shape
    assert rel_l2(gather(b), gather(a)) <= FP32_TOL
    for f in range(s.layout()["fields"]):
        assert rel_l2(o.field(f), s.snapshot_field(f)) <= FP32_TOL, f
    assert s.diagnostics()["tractionRelative"] < 5e-6


def layered(name):
    if name == "dry_loess_gsls3_pml":
        p = precision_case(name)
        p.maxSteps = 300
        return p
    return ref_case(name)


@pytest.mark.parametrize("name", ["sine_terrain_small", "gsls2_layered_pml", "dry_loess_gsls3_pml"])
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_per_node_arrays_equal_row_tables(name, prec):
    """A layered model read through the per-node arrays (engine option material_rows = 0) gives the same
    traces, fields and PML memories, bit for bit, as through the per-row tables."""
    p = layered(name)
    p.precision = prec
    print(f"[pernode] rows vs nodes {name} {prec.name}", flush=True)
    runs = []
    for rows in (1, 0):
        s = hip(p, {"material_rows": rows})
        assert s.kernel_timing()["material_by_row"] is bool(rows)
        assert s.advance(s.totalSteps())[0]
        runs.append((s.receiver_arrays(), [s.snapshot_field(f) for f in range(s.layout()["fields"])], s.pml()))
        s.close()
    (ra, fa, pa), (rb, fb, pb) = runs
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])
    for f in range(len(fa)):
        assert np.array_equal(fa[f], fb[f]), f
    assert np.array_equal(pa, pb)
