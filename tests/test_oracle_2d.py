"""Pins the 2D oracle (oracle/oracle2d.c) before it is trusted as the HIP checker.

* golden data from the reference (docs/data/precision-2d-2026-10-28/storage.csv and
  dry_loess_traces.csv, produced by the reference on its own build), copied under
  tests/golden/;
* the reference's known answers (src/two_d/Validation2D.cpp:44, :69-70);
* the reference build itself (oracle/_ref, same inputs, bitwise) where available, and its
  committed outputs on four cases (tests/golden/reference_2d_*.npz) everywhere;
* closed forms: scipy PCHIP, flat-metric identity, conservative moment pattern.
"""
import ctypes as C
import math
import pathlib

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, have_ref, load, rel_l2
from reference_fixtures import FIXTURES_2D, load_fixture, run_oracle_2d
from seiswavelab_curvilinear_amd.params import (Absorber, CParams2D, GslsSettings, Layer, Parameters, Precision,
                                                 SourceKind, SourceMechanism)

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def precision_fixture():
    """PrecisionValidation2D.cpp:26-31 fixture()."""
    p = Parameters(nx=80, nz=72, spacing=1, vp=550, vs=250, density=1550, relief=.5, wavelength=80, frequency=20,
                   duration=.35, maxSteps=0, cfl=.5, sourceX=40, sourceZ=8, forcePerLength=100, receiverStart=0,
                   receiverCount=16, receiverStride=5, receiverDepth=1, spongeX=12, spongeZ=12,
                   externalAbsorption=True, sampleIntervalMs=.25)
    p.attenuation = GslsSettings(referenceFrequency=20, minFrequency=5, maxFrequency=60, qp=40, qs=25)
    p.cpuThreads = 4
    return p


def precision_case(name):
    """The six cases of PrecisionValidation2D.cpp:94-99."""
    p = precision_fixture()
    if name == "elastic_flat":
        p.relief = 0
    elif name == "elastic_curved_moment":
        p.source = SourceMechanism(SourceKind.IsotropicMoment)
        p.momentPerLength = 100
    elif name == "elastic_curved_pml":
        p.absorber = Absorber.CfsPml
    elif name == "gsls1_curved_pml":
        p.absorber = Absorber.CfsPml
        p.attenuation.enabled = True
        p.attenuation.mechanisms = 1
    elif name == "gsls3_curved_pml":
        p.absorber = Absorber.CfsPml
        p.attenuation.enabled = True
    elif name == "dry_loess_gsls3_pml":
        p.absorber = Absorber.CfsPml
        p.attenuation.enabled = True
        p.layered = True
        p.layers = [Layer(3, 500, 150, 1450, 40, 25), Layer(0, 1200, 600, 1900, 80, 50)]
        p.spacing, p.frequency, p.relief, p.wavelength, p.sourceZ, p.duration = .25, 15, .15, 20, 5, .4
    return p


def storage_rows():
    rows = {}
    for line in (GOLDEN / "precision2d_storage.csv").read_text().splitlines()[1:]:
        f = line.split(",")
        rows[(f[0], f[1], f[2])] = {"energy_ratio": float(f[9]), "dt": float(f[10])}
    return rows


def test_params_layout_matches_product_header():
    lib = load("orc")
    lib.orc2d_params_size.restype = C.c_size_t
    assert lib.orc2d_params_size() == C.sizeof(CParams2D)


@pytest.mark.parametrize("name", ["elastic_flat", "elastic_curved_moment", "elastic_curved_pml", "gsls1_curved_pml",
                                  "gsls3_curved_pml"])
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_storage_golden_dt_and_energy(name, prec):
    """Exact FP64 dt and final/peak energy ratio of the reference's CPU rows."""
    row = storage_rows()[(name, "CPU", "FP64" if prec == Precision.Float64 else "FP32")]
    p = precision_case(name)
    p.precision = prec
    s = CpuSession2D("orc", p)
    assert s.dt == row["dt"]
    ok, err = s.advance(s.total)
    assert ok, err
    d = s.diagnostics()
    assert d["finite"] == 1.0
    assert abs(d["energy"] / d["peakEnergy"] / row["energy_ratio"] - 1) < 1e-13
    assert d["tractionRelative"] < (5e-6 if prec == Precision.Float32 else 1e-9)


def test_dry_loess_golden_trace():
    """dry_loess_traces.csv: middle receiver Vz of the GSLS3+CFS-PML+layered+curved case."""
    g = np.genfromtxt(GOLDEN / "precision2d_dry_loess_traces.csv", delimiter=",", names=True)
    rows = storage_rows()
    p = precision_case("dry_loess_gsls3_pml")
    p.cpuThreads = 8
    for prec, col in ((Precision.Float64, "cpu64_vz"), (Precision.Float32, "cpu32_vz")):
        p.precision = prec
        s = CpuSession2D("orc", p)
        assert s.dt == rows[("dry_loess_gsls3_pml", "CPU", "FP64")]["dt"]
        ok, err = s.advance(s.total)
        assert ok, err
        vz = s.receivers()["vz"][8]
        assert vz.size == len(g)
        np.testing.assert_allclose(np.arange(vz.size) * s.record_dt, g["time_s"], rtol=1e-15, atol=0)
        assert rel_l2(g[col], vz) < (1e-13 if prec == Precision.Float64 else 1e-12)


def test_known_answer_43370_steps_and_damping_dt():
    """Validation2D.cpp:34-44 and :62-71."""
    p = Parameters(nx=48, nz=48, spacing=.25, relief=0, vp=2300, vs=1250, density=2200, sourceZ=4, receiverCount=4,
                   duration=2, frequency=20, cfl=.6, maxSteps=0, sampleIntervalMs=.5)
    s = CpuSession2D("orc", p)
    assert s.total == 43370 and s.total * s.dt == p.duration
    p.maxSteps = 25000
    assert CpuSession2D("orc", p).total == 25000
    p.nx, p.nz, p.duration, p.maxSteps, p.spongeX, p.spongeZ = 601, 401, .001, 2, 200, 100
    s = CpuSession2D("orc", p)
    wx, wz = min(p.spongeX, p.nx // 3), min(p.spongeZ, p.nz // 3)
    expected = .5 / (p.spongeStrength * p.vp / p.spacing * (1. / wx + 1. / wz))
    assert abs(s.diagnostics()["dampingDt"] / expected - 1) < 1e-12


def test_invalid_parameters_rejected():
    p = Parameters(nx=7)
    with pytest.raises(ValueError, match="8..16384"):
        CpuSession2D("orc", p)
    p = Parameters(source=SourceMechanism(SourceKind.IsotropicMoment), sourceZ=3)
    with pytest.raises(ValueError, match="five nodes inside"):
        CpuSession2D("orc", p)


def test_flat_metric_identity():
    p = Parameters(nx=40, nz=30, relief=0, receiverCount=4, sourceZ=10)
    s = CpuSession2D("orc", p)
    assert np.all(s.model(0) == 0) and np.all(s.model(1) == 1) and np.all(s.model(2) == 1)
    n = s.model(10).reshape(-1, 2)
    assert np.all(n[:, 0] == 0) and np.all(n[:, 1] == 1)


def test_imported_surface_is_pchip():
    """SurfaceProfile interior/endpoint slopes are SciPy's PCHIP (SurfaceProfile.cpp:28-46)."""
    scipy_interp = pytest.importorskip("scipy.interpolate")
    rng = np.random.default_rng(3)
    xs = np.cumsum(rng.uniform(5, 25, 60)) - 10.0
    zs = 6 * np.sin(xs / 37.0) + rng.normal(0, .8, xs.size)
    p = Parameters(nx=200, nz=60, spacing=4, importedSurface=True, surfaceSamples=list(zip(xs, zs)),
                   receiverCount=4, sourceZ=10, flattenDepth=200)
    s = CpuSession2D("orc", p)
    L = s.layout()
    coord = s.model(11).reshape(L["nz"], L["nx"], 2)
    x = coord[3, 3:-3, 0]
    z_surface = -coord[3, 3:-3, 1]  # datum 0: depth = -elevation on the surface row
    ref = scipy_interp.PchipInterpolator(xs, zs)(x)
    np.testing.assert_allclose(z_surface, ref, rtol=0, atol=1e-12 * np.max(np.abs(zs)))


def test_moment_pattern_conservative():
    """SourceMechanism.hpp:43-56: the mapped FD4 divergence of a compact kernel has zero net force on a
    curved grid; on a flat grid its first moment is the identity tensor times the kernel mass."""
    for relief in (3.0, 0.0):
        p = Parameters(nx=61, nz=51, spacing=2, relief=relief, wavelength=120,
                       source=SourceMechanism(SourceKind.IsotropicMoment), sourceX=30, sourceZ=20, receiverCount=4)
        s = CpuSession2D("orc", p)
        gx, gz, jac = s.model(8), s.model(9), s.model(2)
        coord = s.model(11).reshape(-1, 2)
        w = jac * p.spacing
        scale = np.sum(np.abs(gx * w))
        assert abs(np.sum(gx * w)) < 1e-13 * scale and abs(np.sum(gz * w)) < 1e-13 * scale
        assert abs(np.sum(s.model(6) * jac) - 1) < 1e-14
        if relief == 0:
            m = np.array([[np.sum(coord[:, a] * g * w) for g in (gx, gz)] for a in (0, 1)])
            np.testing.assert_allclose(m, np.eye(2) * p.spacing, rtol=0, atol=1e-12)


REF_CASES = {
    "sine_terrain_small": dict(nx=120, nz=60, spacing=2, frequency=30, duration=1, cfl=.5, maxSteps=120, sourceX=-1,
                          sourceZ=0, absorber=Absorber.CfsPml, spongeX=20, spongeZ=20, receiverStart=0,
                          receiverCount=120, receiverStride=1, receiverDepth=0, layered=True),
    "moment_external_pml": dict(nx=90, nz=50, spacing=2, relief=4, wavelength=100, frequency=30, maxSteps=150,
                                source=SourceMechanism(SourceKind.IsotropicMoment), sourceZ=6,
                                absorber=Absorber.CfsPml, externalAbsorption=True, spongeX=15, spongeZ=15,
                                receiverStart=0, receiverCount=45, receiverStride=2, receiverDepth=0),
    "sponge_tilted_force": dict(nx=70, nz=60, spacing=1.5, relief=2, wavelength=60, maxSteps=140,
                                source=SourceMechanism(direction=(1.0, 0.0, 2.0)), sourceZ=3, spongeX=10,
                                spongeZ=12, receiverStart=1, receiverCount=30, receiverStride=2, receiverDepth=2,
                                sampleIntervalMs=.13),
    "gsls2_layered_pml": dict(nx=64, nz=48, spacing=1, relief=1, wavelength=50, frequency=25, maxSteps=160,
                              absorber=Absorber.CfsPml, spongeX=10, spongeZ=10, layered=True, sourceZ=2,
                              receiverStart=0, receiverCount=32, receiverStride=2, receiverDepth=0),
}


def ref_case(name):
    kw = dict(REF_CASES[name])
    p = Parameters(**kw)
    if name == "sine_terrain_small":
        phase = 1.234
        x = np.arange(0, 240, 2.0)
        p.importedSurface = True
        p.surfaceSamples = list(zip(x, 10 * np.sin(2 * np.pi * x / 400) + 4 * np.sin(2 * np.pi * x / 137 + phase)))
        p.layers = [Layer(60, 800, 400, 1800), Layer(140, 1800, 900, 2100), Layer(0, 3000, 1700, 2400)]
    if name == "gsls2_layered_pml":
        p.layers = [Layer(6, 900, 400, 1800, 30, 15), Layer(0, 1800, 900, 2100, 70, 35)]
        p.attenuation = GslsSettings(enabled=True, mechanisms=2, minFrequency=5, maxFrequency=75,
                                     referenceFrequency=25)
    p.cpuThreads = 2
    return p


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("name", sorted(REF_CASES))
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_oracle_bitwise_equals_reference_build(name, prec):
    p = ref_case(name)
    p.precision = prec
    a, b = CpuSession2D("ref", p), CpuSession2D("orc", p)
    assert a.dt == b.dt and a.total == b.total and a.layout() == b.layout()
    assert np.array_equal(a.probe_indices(), b.probe_indices())
    cast = (lambda x: x.astype(np.float32).astype(np.float64)) if prec == Precision.Float32 else (lambda x: x)
    for which in (0, 1, 2, 5, 10, 11, 12, 13):  # the reference stores model arrays in the run precision
        assert np.array_equal(a.model(which), cast(b.model(which))), which
    assert a.advance(a.total)[0] and b.advance(b.total)[0]
    ra, rb = a.receivers(), b.receivers()
    for k in ("vx", "vz", "position", "xz"):
        assert np.array_equal(ra[k], rb[k]), k
    for f in range(a.layout()["fields"]):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.pml(), b.pml())
    da, db = a.diagnostics(), b.diagnostics()
    for k in da:
        assert da[k] == db[k] or (math.isinf(da[k]) and math.isinf(db[k])), k


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
def test_pml_slot_map_matches_reference():
    ref, orc = load("ref"), load("orc")
    for nx, nz, wx, wz in ((26, 20, 4, 3), (57, 41, 9, 12), (14, 14, 2, 2)):
        for q in range(nx * nz):
            for axis in (0, 1):
                i, k = q % nx, q // nx
                if not (3 <= i < nx - 3 and 3 <= k < nz - 3):
                    continue
                assert ref.ref2d_pml_slot(nx, nz, wx, wz, q, axis) == orc.orc2d_pml_slot(nx, nz, wx, wz, q, axis)


@pytest.mark.parametrize("name,prec", FIXTURES_2D)
def test_oracle_matches_reference_fixtures(name, prec):
    """Bitwise against the reference build's committed outputs (dt, step count, energy per step,
    receiver traces, final field statistics)."""
    p = ref_case(name)
    p.precision = prec
    got, want = run_oracle_2d("orc", p), load_fixture("2d", name, prec)
    assert sorted(got) == sorted(want)
    for k in want:
        assert np.array_equal(got[k], want[k]), k


# ---- per-node materials (the extension for heterogeneous models, oracle2d.c per-node branch) ----
# The reference's 2D API has layered models only, so the per-node branch is pinned two ways:
#  * a layered model restated as per-node arrays (every logical node carries its layer's values) must
#    reproduce the layered run bit for bit: the reference build's fixtures, its live runs and the
#    reference's own golden dry-loess trace (GSLS-3, external CFS-PML, curved, layered);
#  * point anomalies must land on the node whose coordinates (pinned to the reference build) are those
#    of the logical node, and the ghost and external-strip nodes carry the nearest logical node's values.


def pernode_twin(p):
    """The same model given as per-node Vp, Vs, density on the logical nz x nx grid."""
    import copy
    from seiswavelab_curvilinear_amd.workloads import layered_base
    q = copy.deepcopy(p)
    if p.layered:
        rows = layered_base(p.nz, p.spacing, p.layers)
    else:
        rows = tuple(np.full(p.nz, v) for v in (p.vp, p.vs, p.density))
    q.nodeVp, q.nodeVs, q.nodeDensity = (np.repeat(a[:, None], p.nx, axis=1) for a in rows)
    return q


@pytest.mark.parametrize("name,prec", FIXTURES_2D)
def test_pernode_twin_matches_reference_fixtures(name, prec):
    p = ref_case(name)
    p.precision = prec
    got, want = run_oracle_2d("orc", pernode_twin(p)), load_fixture("2d", name, prec)
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_pernode_twin_of_dry_loess_is_bitwise_the_layered_run():
    """GSLS-3 + external CFS-PML + curved + two layers, given per node (each node refits its own GSLS
    material): models, traces, fields and PML memories bitwise equal to the layered run, which
    test_dry_loess_golden_trace holds to the reference's own trace (600 of its 5664 steps here)."""
    p = precision_case("dry_loess_gsls3_pml")
    p.cpuThreads = 8
    a, b = CpuSession2D("orc", p), CpuSession2D("orc", pernode_twin(p))
    for which in (3, 4, 5, 12, 13, 16, 17, 18, 19, 20, 21):
        assert np.array_equal(a.model(which), b.model(which)), which
    assert a.dt == b.dt and a.total == b.total
    assert a.advance(600)[0] and b.advance(600)[0]
    ra, rb = a.receivers(), b.receivers()
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])
    for f in range(a.layout()["fields"]):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.pml(), b.pml())


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_pernode_twin_equals_reference_build_layered_external(prec):
    """Layers under an external CFS-PML: the strips below the model take the bottom row's layer in the
    reference (its depth is clamped to the model's bottom); the per-node branch clamps to the last
    logical row.  Live against the reference build, every array and trace bit for bit."""
    p = ref_case("moment_external_pml")
    p.precision = prec
    p.layered = True
    p.layers = [Layer(20, 900, 450, 1800), Layer(30, 1800, 900, 2100), Layer(0, 3000, 1700, 2400)]
    a, b = CpuSession2D("ref", p), CpuSession2D("orc", pernode_twin(p))
    assert a.dt == b.dt and a.total == b.total
    cast = (lambda x: x.astype(np.float32).astype(np.float64)) if prec == Precision.Float32 else (lambda x: x)
    for which in (5, 12, 13):  # the reference build exposes density, Vp and Vs
        assert np.array_equal(a.model(which), cast(b.model(which))), which
    assert a.advance(a.total)[0] and b.advance(b.total)[0]
    ra, rb = a.receivers(), b.receivers()
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])
    for f in range(a.layout()["fields"]):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.pml(), b.pml())


@pytest.mark.parametrize("external", [True, False])
def test_pernode_index_map_follows_coordinates(external):
    """Point anomalies at logical nodes (corners, edges, interior) of a flat model.  The padded node that
    carries each is found from the coordinate arrays (x = i h, depth = k h on a flat grid, pinned to the
    reference build), not from the index formula; ghost nodes and external-strip nodes must carry the
    values of the nearest logical node."""
    nx, nz, h = 41, 33, 2.0
    p = Parameters(nx=nx, nz=nz, spacing=h, relief=0, vp=3000, vs=1700, density=2400, sourceZ=10,
                   receiverCount=4, absorber=Absorber.CfsPml, spongeX=7, spongeZ=6, externalAbsorption=external)
    vp, vs, rho = np.full((nz, nx), 3000.0), np.full((nz, nx), 1700.0), np.full((nz, nx), 2400.0)
    spots = [(0, 0), (nx - 1, 0), (0, nz - 1), (nx - 1, nz - 1), (17, 11), (0, 20), (25, nz - 1), (nx - 1, 5), (9, 0)]
    for n, (x, z) in enumerate(spots):
        vp[z, x], vs[z, x], rho[z, x] = 3100.0 + 10 * n, 1750.0 + 5 * n, 2450.0 + 3 * n
    p.nodeVp, p.nodeVs, p.nodeDensity = vp, vs, rho
    s = CpuSession2D("orc", p)
    L = s.layout()
    coord = s.model(11).reshape(L["nz"], L["nx"], 2)
    li = np.clip(np.rint(coord[..., 0] / h).astype(int), 0, nx - 1)
    lk = np.clip(np.rint(coord[..., 1] / h).astype(int), 0, nz - 1)
    assert L["offsetX"] == (7 if external else 0)
    for which, want in ((12, vp), (13, vs), (5, rho)):
        got = s.model(which).reshape(L["nz"], L["nx"])
        assert np.array_equal(got, want[lk, li]), which
    # each interior anomaly sits at exactly one padded node, the one whose coordinates are its own
    got = s.model(12).reshape(L["nz"], L["nx"])
    k, i = np.nonzero(got == vp[11, 17])
    assert len(k) == 1 and coord[k[0], i[0], 0] == 17 * h and coord[k[0], i[0], 1] == 11 * h


def test_oracle_full_size_sine_terrain_matches_reference_build():
    """The 500x200 sine-terrain workload (seed 1: three layers, imported PCHIP surface at every node, 30 Hz
    surface force, 20-node CFS-PML, CFL 0.5) in FP64 for all its 2000 steps, bit for bit against the
    reference build's run (tests/golden/reference_2d_sine_terrain_500x200_f64.npz)."""
    from reference_fixtures import FULL_2D, reduce_full_2d
    from seiswavelab_curvilinear_amd import workloads
    p = workloads.sine_terrain_500x200(seed=1)
    p.cpuThreads = 8
    got, want = reduce_full_2d(run_oracle_2d("orc", p)), load_fixture("2d", FULL_2D)
    assert int(want["total"]) == 2000
    assert sorted(got) == sorted(want)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
