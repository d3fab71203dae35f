"""3D HIP engine (csrc/engine3d.hip) against the CPU oracle, through the C-ABI.

FP64: the engine follows the reference's operation order, so fields and traces match the oracle to
rounding (bar: 1e-10 relative L2; the energy sum is a parallel reduction, bar 1e-12).  FP32: state
in FP32, surface closure / projection / energy in FP64; bar 1e-4 relative L2 over the whole gather.
The same bars hold against the reference build's committed outputs (tests/golden/reference_3d_*.npz).
"""
import numpy as np
import pytest

from cpu_sessions import CpuSession3D, rel_l2
from reference_fixtures import load_fixture, record_2d
from seiswavelab_curvilinear_amd import BackendKind, Simulation3D
from test_oracle_3d import CASES, case

pytestmark = pytest.mark.gpu


def _gpu(p, precision):
    print(f"[3d] case nx={p.nx} ny={p.ny} nz={p.nz} precision={precision}", flush=True)
    s = Simulation3D(BackendKind.Hip)
    ok, err = s.initialize(p, 0, precision)
    assert ok, err
    return s


@pytest.mark.parametrize("name", CASES)
def test_fp64_matches_oracle_3d(name):
    p = case(name)
    o = CpuSession3D("orc", p)
    s = _gpu(p, 0)
    assert s.dt() == o.dt and s.totalSteps() == o.total
    ok, err = s.advance(p.maxSteps)
    assert ok, err
    assert o.advance(p.maxSteps)[0]
    ro, rs = o.receivers(), s.receiver_arrays()
    gather_o = np.concatenate([ro["vx"], ro["vy"], ro["vz"]])
    gather_s = np.concatenate([rs["vx"], rs["vy"], rs["vz"]])
    assert np.abs(gather_o).max() > 0
    assert rel_l2(gather_o, gather_s) <= 1e-10
    for f in range(9):
        assert rel_l2(o.field(f), s.field(f)) <= 1e-10, f
    do, ds = o.diagnostics(), s.diagnostics()
    assert abs(ds["elasticEnergy"] - do["energy"]) <= 1e-12 * do["energy"]
    assert abs(ds["peakEnergy"] - do["peakEnergy"]) <= 1e-12 * do["peakEnergy"]
    for k, kk in (("maxAbsVelocity", "maxAbsVelocity"), ("peakVz", "peakVz"), ("peakPressure", "peakPressure")):
        assert abs(ds[k] - do[kk]) <= 1e-10 * max(abs(do[kk]), 1e-300), k
    assert ds["allFinite"] and ds["maxTractionRateRelative"] < 1e-9


@pytest.mark.parametrize("name", ["analytic_sponge_force", "dem_layers_moment"])
def test_fp32_within_tolerance_of_fp64_oracle_3d(name):
    p = case(name)
    o = CpuSession3D("orc", p)
    s = _gpu(p, 1)
    ok, err = s.advance(p.maxSteps)
    assert ok, err
    assert o.advance(p.maxSteps)[0]
    ro, rs = o.receivers(), s.receiver_arrays()
    gather_o = np.concatenate([ro["vx"], ro["vy"], ro["vz"]])
    gather_s = np.concatenate([rs["vx"], rs["vy"], rs["vz"]])
    assert rel_l2(gather_o, gather_s) <= 1e-4
    d = s.diagnostics()
    assert d["allFinite"] and d["maxTractionRateRelative"] < 5e-6


def test_reset_is_bitwise_reproducible_and_chunking_invariant_3d():
    p = case("dem_layers_moment")
    s = _gpu(p, 0)
    assert s.advance(p.maxSteps)[0]
    first = s.receiver_arrays()
    f0 = s.field(2)
    assert s.resetFields()[0]
    assert s.currentStep() == 0
    for n in (5, 1, 7, p.maxSteps - 13):
        assert s.advance(n)[0]
    again = s.receiver_arrays()
    assert np.array_equal(first["vz"], again["vz"]) and np.array_equal(first["vx"], again["vx"])
    assert np.array_equal(f0, s.field(2))


def test_first_download_is_complete_3d():
    """The first field download allocates its staging buffer; the allocation's clear must be finished
    before the conversion kernel on the engine's stream writes into it (it once raced and zeroed part of
    the first field read)."""
    p = case("dem_layers_moment")
    for _ in range(3):
        s = _gpu(p, 0)
        assert s.advance(p.maxSteps)[0]
        first = s.field(0)
        assert np.array_equal(first, s.field(0))
        s.close()


def test_display_frame_planes_3d():
    p = case("analytic_sponge_force")
    s = _gpu(p, 0)
    assert s.advance(10)[0]
    ok, err, planes = s.readDisplayFrame(p.nx // 2, p.ny // 2, 1, 0)
    assert ok, err
    vz = s.field(2).reshape(p.nz + 6, p.ny + 6, p.nx + 6)
    assert np.array_equal(planes[0], vz[3:-3, p.ny // 2 + 3, 3:-3].astype(np.float32))
    assert np.array_equal(planes[1], vz[3:-3, 3:-3, p.nx // 2 + 3].astype(np.float32))
    assert np.array_equal(planes[2], vz[4, 3:-3, 3:-3].astype(np.float32))
    ok, err, _ = s.readDisplayFrame(p.nx, 0, 0, 0)
    assert not ok and "outside the active physical grid" in err


def test_nonfinite_model_rejects_step_and_latches_fault_3d():
    p = case("analytic_sponge_force")
    p.sourceForce = 1e308
    s = _gpu(p, 0)
    ok, err = s.advance(p.maxSteps)
    assert not ok
    assert err in ("Non-finite RK4 step rejected. Last accepted fields retained; rebuild with smaller CFL.",
                   "Energy grew above 10,000 times the source-window peak. Last accepted step retained; "
                   "rebuild with smaller CFL or gentler terrain.")
    step = s.currentStep()
    assert np.all(np.isfinite(s.field(2)))
    ok, err2 = s.advance(1)
    assert not ok and err2 == err and s.currentStep() == step


@pytest.mark.parametrize("name", CASES)
def test_fp64_matches_reference_fixtures_3d(name):
    """Against the reference build's own outputs: dt and step count exactly; traces and final field
    statistics <= 1e-10; energy after every step <= 1e-12 (a parallel sum)."""
    want = load_fixture("3d", name)
    p = case(name)
    s = _gpu(p, 0)
    assert s.dt() == float(want["dt"]) and s.totalSteps() == int(want["total"])
    energy, peak, stats = record_2d(s, p.maxSteps,
                                    lambda x: (x.diagnostics()["elasticEnergy"], x.diagnostics()["peakEnergy"]),
                                    lambda x: [x.field(f) for f in range(9)])
    r = s.receiver_arrays()
    got = np.concatenate([r["vx"].ravel(), r["vy"].ravel(), r["vz"].ravel()])
    ref = np.concatenate([want["vx"].ravel(), want["vy"].ravel(), want["vz"].ravel()])
    assert r["vx"].shape == want["vx"].shape
    assert rel_l2(ref, got) <= 1e-10
    np.testing.assert_allclose(energy, want["energy"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(peak, want["peak_energy"], rtol=1e-12, atol=0)
    for f in range(9):
        np.testing.assert_allclose(stats[f], want["field_stats"][f], rtol=1e-10, atol=1e-10 * want["field_stats"][f][0])


@pytest.mark.parametrize("name", CASES)
def test_device_geometry_equals_host_geometry_3d(name):
    """A HIP session builds its coordinates, inverse metrics, Jacobian, normals and CFL figures on the GPU in FP64
    (csrc/geometry3d.hip); they must equal the host builder's (= the oracle's, = the reference's) bit for bit, and
    so must everything derived from them (materials, sponge, source, dt, step count)."""
    from seiswavelab_curvilinear_amd import BackendKind as BK
    p = case(name)
    host = Simulation3D(BK.Cpu)
    ok, err = host.initialize(p)
    assert ok, err
    dev = _gpu(p, 0)
    assert dev.dt() == host.dt() and dev.totalSteps() == host.totalSteps()
    for which in range(12):
        assert np.array_equal(dev.model_array(which), host.model_array(which)), which
    dh, dd = host.diagnostics(), dev.diagnostics()
    for k in ("minJacobian", "maxMetricNorm", "minCellLength", "maxCellEdge", "waveDtLimit", "dampingDtLimit",
              "sourceDtLimit"):
        assert dh[k] == dd[k], k


def test_device_geometry_dem_workload_3d():
    """The 400x288x200 DEM workload: GPU-built geometry equals the host-built one (dt, CFL figures, metrics)."""
    from seiswavelab_curvilinear_amd import BackendKind as BK
    from seiswavelab_curvilinear_amd import workloads
    import time
    p, _ = workloads.dem_terrain_400x288x200(steps=4)
    host = Simulation3D(BK.Cpu)
    t0 = time.perf_counter()
    ok, err = host.initialize(p)
    assert ok, err
    t_host = time.perf_counter() - t0
    dev = _gpu(p, 0)
    assert dev.dt() == host.dt() and dev.totalSteps() == host.totalSteps()
    dh, dd = host.diagnostics(), dev.diagnostics()
    for k in ("minJacobian", "maxMetricNorm", "minCellLength", "maxCellEdge", "waveDtLimit"):
        assert dh[k] == dd[k], k
    for which in (0, 1, 2, 3):
        assert np.array_equal(dev.model_array(which), host.model_array(which)), which
    print(f"[3d geometry] host plan {1e3 * t_host:.0f} ms, HIP session (geometry on the GPU, engine set-up, "
          f"upload) {dd['initializationMs']:.0f} ms", flush=True)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["dem_layers_moment", "flat_datum_layers"])
def test_lds_tiles_equal_direct_loads_3d(name, precision):
    """The deep bands' LDS-staged stencil input (option "lds_tiles") reads the same values as direct loads, so
    fields and traces are bit-identical either way."""
    p = case(name)
    runs = []
    for tiles in (0, 1):
        s = _gpu(p, precision)
        s.set_option("lds_tiles", tiles)
        ok, err = s.advance(p.maxSteps)
        assert ok, err
        runs.append((s.receiver_arrays(), [s.field(f) for f in range(9)]))
        s.close()
    (ra, fa), (rb, fb) = runs
    for k in ("vx", "vy", "vz"):
        assert np.array_equal(ra[k], rb[k]), k
    for f in range(9):
        assert np.array_equal(fa[f], fb[f]), f


def _odd_case(name):
    """A case whose flat band has ragged edges: 27 physical columns (not a multiple of 4 or 2 nodes per lane) and
    37 rows (two 32-row vector tiles), the source inside the flat band."""
    p = case(name)
    p.nx, p.ny, p.nz, p.maxSteps = 27, 37, 26, 16
    p.sourceX, p.sourceY, p.sourceZ = 13, 17, 12
    return p


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["flat_datum_layers", "analytic_sponge_force"])
def test_vector_bands_equal_node_lanes_3d(name, precision):
    """The deep bands with 2 or 4 x-adjacent nodes per lane (options "vector_terrain", "vector_flat"), with the terrain
    band reading all nine metrics or only their z row ("z_row_metrics", where the other six are exactly 1, 0, 0, 0,
    1, 0), compute every node with the one-node-per-lane expressions: fields and traces bit-identical in every
    combination; the energy
    (a sum over tiles of another shape) to 1e-12.  analytic_sponge_force has a terrain band (planes 7..14) and a
    flat band."""
    p = _odd_case(name)
    runs = []
    # (terrain, flat, z_row_metrics): z_row_metrics reads only the z row of the terrain band's metrics
    widths = (((0, 0, 1), (2, 2, 1), (2, 0, 1), (0, 2, 1), (2, 2, 0)) if precision == 0 else
              ((0, 0, 1), (2, 4, 1), (4, 2, 1), (4, 0, 1), (0, 4, 1), (2, 4, 0), (4, 2, 0)))
    for terrain, flat, zrow in widths:
        s = _gpu(p, precision)
        s.set_option("vector_terrain", terrain)
        s.set_option("vector_flat", flat)
        s.set_option("z_row_metrics", zrow)
        s.set_option("z_row_surface", zrow)
        assert s.kernel_timing()["z_row_metrics"] == bool(terrain and zrow), (terrain, zrow)
        t = s.kernel_timing()
        assert max(7, t["metric_planes"]) <= p.nz + 3 - 4, "the case has no flat band of 4 planes"
        ok, err = s.advance(p.maxSteps)
        assert ok, err
        runs.append((s.receiver_arrays(), [s.field(f) for f in range(9)], s.diagnostics()))
        s.close()
    ra, fa, da = runs[0]
    for rb, fb, db in runs[1:]:
        for k in ("vx", "vy", "vz"):
            assert np.array_equal(ra[k], rb[k]), k
        for f in range(9):
            assert np.array_equal(fa[f], fb[f]), f
        assert abs(db["elasticEnergy"] - da["elasticEnergy"]) <= 1e-12 * da["elasticEnergy"]
        assert db["maxAbsVelocity"] == da["maxAbsVelocity"] and db["allFinite"] == da["allFinite"]
