"""3D oracle (oracle/oracle3d.c) pinned against the reference build, closed forms and invariants.

The reference's 3D CPU path (src/three_d/CurvilinearSolver.cpp, PointKernels3D.hpp) is compiled
from its own sources into oracle/_ref/libswlref.so; every comparison with it is bit for bit, live where
the build is present and against its committed outputs (tests/golden/reference_3d_*.npz) everywhere.
"""
import numpy as np
import pytest

from cpu_sessions import CpuSession3D, have_ref, load
from reference_fixtures import load_fixture, run_oracle_3d
from seiswavelab_curvilinear_amd.params import (LayerDepthMode3D, MaterialLayer3D, ReceiverArrayKind3D,
                                                SimulationParameters, SourceKind, SourceMechanism, SurfaceGrid3D)


def smooth_dem(nx, ny, spacing, seed, rms=15.0):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((ny, nx))
    k = np.exp(-0.5 * (np.arange(-4, 5) / 2.0) ** 2)
    k /= k.sum()
    for axis in (0, 1):
        z = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), axis, z)
    z *= rms / z.std()
    return SurfaceGrid3D(nx=nx, ny=ny, spacing=spacing, originX=0, originY=0, datum=float(z.mean()) + 100.0,
                         elevations=100.0 + z)


def case(name):
    p = SimulationParameters(nx=24, ny=20, nz=22, spacing=10.0, maxSteps=24, duration=1.0, dominantFrequency=12.0,
                             cpuThreads=2)
    if name == "analytic_sponge_force":
        p.topographyAmplitude, p.topographyWavelength = 20.0, 120.0
        p.source = SourceMechanism(SourceKind.DirectionalForce, (0.3, -0.2, 1.0))
        p.sourceZ = 3
        p.absorbingBoundary.cells = (6, 5, 6)
        p.receiverArray.z = 1
    elif name == "dem_layers_moment":
        p.surfaceGrid = smooth_dem(30, 26, 10.0, seed=7)
        p.topographyAmplitude = 0.0
        p.layeredMaterial, p.layerDepthMode = True, LayerDepthMode3D.BelowSurface
        p.layers = [MaterialLayer3D("a", 40, 1500, 800, 2000), MaterialLayer3D("b", 60, 2500, 1400, 2300),
                    MaterialLayer3D("c", 0, 3500, 2000, 2500)]
        p.vp, p.vs, p.density = 3500.0, 2000.0, 2500.0
        p.source = SourceMechanism(SourceKind.IsotropicMoment)
        p.sourceX, p.sourceY, p.sourceZ = 12, 9, 7
        p.receiverArray.kind, p.receiverArray.startX, p.receiverArray.startY = ReceiverArrayKind3D.GridXY, 1, 2
        p.receiverArray.countX, p.receiverArray.countY, p.receiverArray.strideX, p.receiverArray.strideY = 7, 5, 3, 3
        p.receiverArray.z = 0
        p.absorbingBoundary.cells = (5, 5, 5)
    elif name == "dem_gridindex_layers":
        p.surfaceGrid = smooth_dem(28, 24, 10.0, seed=11, rms=8.0)
        p.topographyAmplitude = 0.0
        p.layeredMaterial, p.layerDepthMode = True, LayerDepthMode3D.GridIndex
        p.layers = [MaterialLayer3D("a", 35, 1500, 800, 2000), MaterialLayer3D("c", 0, 3500, 2000, 2500)]
        p.source = SourceMechanism(SourceKind.DirectionalForce, (0.0, 0.0, 1.0))
        p.sourceZ = 2
        p.receiverArray.kind, p.receiverArray.startX, p.receiverArray.countX, p.receiverArray.strideX = \
            ReceiverArrayKind3D.XLine, 0, 12, 2
        p.receiverArray.z = 0
    elif name == "flat_datum_layers":
        p.topographyAmplitude = 0.0
        p.layeredMaterial, p.layerDepthMode = True, LayerDepthMode3D.Datum
        p.layers = [MaterialLayer3D("a", 55, 1800, 900, 2100), MaterialLayer3D("c", 0, 3000, 1700, 2400)]
        p.source = SourceMechanism(SourceKind.IsotropicMoment)
        p.sourceX, p.sourceY, p.sourceZ = 11, 10, 8
        p.receiverArray.kind, p.receiverArray.startY, p.receiverArray.countY = ReceiverArrayKind3D.YLine, 0, 20
        p.receiverArray.z = 2
    else:
        raise KeyError(name)
    return p


CASES = ["analytic_sponge_force", "dem_layers_moment", "dem_gridindex_layers", "flat_datum_layers"]
HAVE_REF = have_ref()


@pytest.mark.skipif(not HAVE_REF, reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("name", CASES)
def test_oracle3d_bitwise_equals_reference_build(name):
    p = case(name)
    o, r = CpuSession3D("orc", p), CpuSession3D("ref", p)
    assert o.layout() == r.layout()
    assert o.dt == r.dt and o.total == r.total
    for which in (0, 1, 2, 3, 7):
        assert np.array_equal(o.model(which), r.model(which)), which
    if p.layeredMaterial:
        lam, mu, rho = o.model(4), o.model(5), o.model(6)
        assert np.array_equal(np.sqrt((lam + 2 * mu) / rho), r.model(12))
        assert np.array_equal(np.sqrt(mu / rho), r.model(13))
        assert np.array_equal(rho, r.model(14))
    do, dr = o.diagnostics(), r.diagnostics()
    for k in do:
        if k != "sourcePeakEnergy":
            assert do[k] == dr[k], (k, do[k], dr[k])
    assert o.advance(p.maxSteps)[0] and r.advance(p.maxSteps)[0]
    for f in range(9):
        assert np.array_equal(o.field(f), r.field(f)), f
    ro, rr = o.receivers(), r.receivers()
    for k in ro:
        assert np.array_equal(ro[k], rr[k]), k
    do, dr = o.diagnostics(), r.diagnostics()
    for k in do:
        if k != "sourcePeakEnergy":
            assert do[k] == dr[k], (k, do[k], dr[k])
    assert np.abs(ro["vz"]).max() > 0


def test_flat_grid_closed_forms():
    p = case("flat_datum_layers")
    s = CpuSession3D("orc", p)
    L = s.layout()
    n = L["nx"] * L["ny"] * L["nz"]
    m = s.model(0).reshape(n, 9)
    jac = s.model(1)
    phys = np.zeros((L["nz"], L["ny"], L["nx"]), bool)
    phys[3:-3, 3:-3, 3:-3] = True
    phys = phys.ravel()
    assert np.array_equal(m[phys], np.tile(np.eye(3).ravel(), (phys.sum(), 1)))
    assert np.all(jac[phys] == 1.0)
    d = s.diagnostics()
    assert d["minCellLength"] == p.spacing / np.sqrt(3.0) and d["maxCellEdge"] == p.spacing
    assert d["maxMetricNorm"] == np.sqrt(3.0)
    vp_max = max(l.vp for l in p.layers)
    assert d["waveDtLimit"] == p.cflSafety * min(p.spacing / np.sqrt(3.0), p.spacing / np.sqrt(3.0)) / vp_max
    assert d["energy"] == 0.0 and d["allFinite"] == 1.0


def test_moment_pattern_zero_force_and_unit_tensor():
    """-div(w I) on a flat grid: zero net force, first moment = identity * sum(w) = identity."""
    p = case("flat_datum_layers")
    s = CpuSession3D("orc", p)
    L = s.layout()
    g = [s.model(9 + a) for a in range(3)]
    w = s.model(8)
    coords = s.model(3).reshape(-1, 3)
    assert abs(w.sum() - 1.0) < 1e-14
    for a in range(3):
        assert abs(g[a].sum()) < 1e-12 * np.abs(g[a]).sum()
        for b in range(3):
            moment = np.sum(coords[:, b] * g[a])
            assert abs(moment - (1.0 if a == b else 0.0)) < 1e-12, (a, b, moment)


def test_traction_rate_residual_and_energy_invariants():
    """Invariants the reference's own CFL validation checks (src/three_d/CflValidation.cpp:16-31)."""
    p = case("analytic_sponge_force")
    p.maxSteps = 40
    s = CpuSession3D("orc", p)
    ok, err = s.advance(40)
    assert ok, err
    d = s.diagnostics()
    assert d["allFinite"] == 1.0 and d["maxTractionRateRelative"] < 1e-9
    assert s.dt <= d["waveDtLimit"] and s.dt <= d["dampingDtLimit"] and s.dt <= d["sourceDtLimit"]
    assert d["energy"] > 0 and d["peakEnergy"] >= d["energy"]


def test_invalid_parameters_rejected_with_reference_messages():
    p = case("analytic_sponge_force")
    p.cflSafety = 0.9
    with pytest.raises(ValueError, match="CFL coefficient must be finite and in 0.02..0.8."):
        CpuSession3D("orc", p)
    p = case("dem_layers_moment")
    p.sourceX = 3
    with pytest.raises(ValueError, match="Isotropic moment needs 11\\+ nodes per axis"):
        CpuSession3D("orc", p)
    p = case("analytic_sponge_force")
    p.topographyAmplitude = 200.0
    with pytest.raises(ValueError, match="Terrain transition must fit the model depth"):
        CpuSession3D("orc", p)
    p = case("analytic_sponge_force")
    p.receiverArray.kind, p.receiverArray.countX = ReceiverArrayKind3D.XLine, 500
    with pytest.raises(ValueError, match="Receiver counts exceed"):
        CpuSession3D("orc", p)


@pytest.mark.parametrize("name", CASES)
def test_oracle3d_matches_reference_fixtures(name):
    """Bitwise against the reference build's committed outputs (dt, step count, energy per step,
    receiver traces, final field statistics)."""
    got, want = run_oracle_3d("orc", case(name)), load_fixture("3d", name)
    assert sorted(got) == sorted(want)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
