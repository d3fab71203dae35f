"""A step ended in the next step's first stage launch equals a step ended by separate kernels (GPU).

Device-resident 2D runs end step j inside the first RK stage launch of step j + 1: one block reduces the
inspection partials that step j's last stage wrote, applies the fault rule
(src/two_d/CurvilinearSolver2D.cpp:576-590), writes the diagnostics row and reads the receivers from the
stage input, which is step j's state.  The run's last step is ended by the separate kernels.  The other path
launches a reduction kernel and a finalisation kernel after every step (engine option end_in_stage = 0).

Both must give the same run: fields, PML memories and every receiver sample bit for bit, and the energy within
1e-12 (the partials are summed in another order).  Runs are split into uneven advance() calls so that they
start and end at odd steps; fault handling on this path is covered by test_hip_faults.py.
"""
import numpy as np
import pytest

from seiswavelab_curvilinear_amd import BackendKind, GslsSettings, Precision, Simulation, workloads
from test_oracle_2d import ref_case

pytestmark = pytest.mark.gpu


def grf_case(prec, gsls):
    p = workloads.grf_terrain_5001x2001(seed=5, precision=prec, nx=2400, nz=900, steps=90, topo_rms=8.0).params
    p.receiverCount = 200
    if gsls:
        p.attenuation = GslsSettings(enabled=True, mechanisms=3, minFrequency=6.0, maxFrequency=90.0,
                                     referenceFrequency=30.0, qp=40.0, qs=20.0)
    return p


def run(p, options, chunks):
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Hip, 0)
    assert ok, err
    for name, value in options.items():
        s.set_option(name, value)
    s.profile(True)
    for n in chunks:
        ok, err = s.advance(n)
        assert ok, err
    fields = [s.snapshot_field(f) for f in range(s.layout()["fields"])]
    out = {"rec": s.receiver_arrays(), "diag": s.diagnostics(), "fields": fields, "pml": s.pml(), "step": s.step(),
           "timing": s.kernel_timing()}
    s.close()
    return out


def assert_same(a, b):
    assert a["step"] == b["step"]
    for key in ("vx", "vz"):
        assert np.array_equal(a["rec"][key], b["rec"][key]), key
    assert len(a["fields"]) == len(b["fields"])
    for f, (x, y) in enumerate(zip(a["fields"], b["fields"])):
        assert np.array_equal(x, y), f
    assert np.array_equal(a["pml"], b["pml"])
    da, db = a["diag"], b["diag"]
    assert da["finite"] and db["finite"]
    for key in ("energy", "peakEnergy"):
        assert abs(da[key] - db[key]) <= 1e-12 * abs(db[key]), key


@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
@pytest.mark.parametrize("gsls", [False, True])
def test_step_end_in_stage_equals_separate_kernels(prec, gsls):
    p = grf_case(prec, gsls)
    chunks = [37, 1, 52]
    print(f"[step close] {prec.name} gsls={gsls} 2400x900 GRF, chunks {chunks}", flush=True)
    sep = run(p, {"end_in_stage": 0}, chunks)
    inl = run(p, {}, chunks)
    assert sep["timing"]["finalize_launches"] > 0 and inl["timing"]["finalize_launches"] == 0
    assert inl["timing"]["interior_tiles"] + inl["timing"]["border_tiles"] > 2048  # two-level reduction at run ends
    assert_same(inl, sep)


def test_step_end_in_stage_small_grid_single_steps():
    """A small grid (direct finalisation at run ends) advanced in runs of one step as well."""
    p = ref_case("moment_external_pml")
    p.maxSteps = 60
    chunks = [13, 1, 1, 2, 43]
    a = run(p, {"end_in_stage": 0}, chunks)
    b = run(p, {}, chunks)
    assert_same(b, a)


def test_pipelined_runs_equal_short_runs():
    """advance() over more than one run of 512 steps enqueues the next run before it folds this one; the
    result equals advancing in runs that never overlap."""
    p = ref_case("moment_external_pml")
    p.maxSteps = 1300
    long_ = run(p, {}, [1300])
    short = run(p, {}, [100] * 13)
    assert_same(long_, short)


@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_fault_inside_pipelined_runs(prec):
    """A step rejected inside the first of several pipelined runs: the run after it accepts nothing, the state
    stays at the last accepted step and the fault is latched, as when the rejected step ends a short run."""
    from test_hip_faults import MSG2D, TINY, first_late_step
    p = ref_case("moment_external_pml")
    p.precision = prec
    p.maxSteps = 1400
    out = []
    for span in (1200, 3):
        s = Simulation()
        ok, err = s.initialize(p, BackendKind.Hip, 0)
        assert ok, err
        late = first_late_step(s.dt(), p.frequency)
        assert 10 < late < 400
        assert s.advance(late)[0]
        s.override_peak_energy(TINY)
        ok, err = s.advance(span)
        assert not ok and err == MSG2D and s.step() == late and s.fault() == MSG2D
        out.append(([s.snapshot_field(f) for f in range(5)], s.receiver_arrays()))
        assert not s.advance(1)[0]
        s.close()
    (fa, ra), (fb, rb) = out
    for f in range(5):
        assert np.array_equal(fa[f], fb[f]), f
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])


@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
@pytest.mark.parametrize("gsls", [False, True])
def test_launch_layouts_equal(prec, gsls):
    """The border blocks launched on the engine's second stream beside the interior tiles (split_border), spread
    evenly over the one launch (spread_border), the small-grid FP32 elastic kernel forced on this large grid
    (latency_kernel), or the elastic interior's direct loads instead of its LDS tiles (lds_tiles 0; the tiles are
    automatic at this size) change only the schedule: every field, receiver sample and the energy are the same to
    the bit."""
    p = grf_case(prec, gsls)
    chunks = [40, 50]
    print(f"[launch layouts] {prec.name} gsls={gsls}", flush=True)
    base = run(p, {"split_border": 0, "spread_border": 0}, chunks)
    for opts in ({"split_border": 1}, {"spread_border": 1}, {"latency_kernel": 1}, {"lds_tiles": 0}):
        other = run(p, opts, chunks)
        assert_same(other, base)
        assert other["diag"]["energy"] == base["diag"]["energy"], opts
        assert other["timing"]["latency_kernel"] == (opts.get("latency_kernel") == 1 and not gsls
                                                     and prec == Precision.Float32)
    assert not base["timing"]["latency_kernel"]  # 2400 x 900: more blocks than the automatic choice takes
    assert base["timing"]["lds_tiles"] == (not gsls)  # elastic: the interior's input staged through LDS


def test_latency_kernel_automatic_on_small_grids():
    """A small FP32 elastic grid takes the small-grid kernel by itself; forced off it gives the same run, bit for
    bit (the two kernels differ only in how their loads are scheduled)."""
    p = ref_case("moment_external_pml")
    p.precision = Precision.Float32
    chunks = [37, 1, 52]
    auto = run(p, {}, chunks)
    off = run(p, {"latency_kernel": 0}, chunks)
    assert auto["timing"]["latency_kernel"] and not off["timing"]["latency_kernel"]
    assert_same(auto, off)
    assert auto["diag"]["energy"] == off["diag"]["energy"]
