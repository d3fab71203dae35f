"""The benchmark configurations themselves, on the GPU (the small-case parity tests live elsewhere).

* 5001x2001 GRF-terrain elastic shot (30-node external CFS-PML, 1001 receivers): FP32 and FP64 HIP sessions
  agree with the oracle's plan of the same model on dt, step count, layout and every receiver index; after
  400 steps (past the 30 Hz Ricker's peak at 1.5/f) both are finite, record 401 samples per receiver, run the
  per-node material branch, and the FP32 gather is within 1e-4 relative L2 of FP64.
* 64-shot GSLS-3 batch: every shot stepped in the batch is bit-identical to the same shot alone, and within
  1e-4 of the shot's FP64 run.
* Serving stream (elastic and GSLS-3 requests of 200x100 to 3000x1200 nodes): every served result is
  bit-identical to the request run alone.
"""
import copy

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, rel_l2
from seiswavelab_curvilinear_amd import Batch2D, BackendKind, Precision, Simulation, workloads
from seiswavelab_curvilinear_amd.serving import Server

pytestmark = pytest.mark.gpu


def gather(r):
    return np.concatenate([r["vx"].ravel(), r["vz"].ravel()])


def session(p):
    s = Simulation()
    ok, err = s.initialize(copy.deepcopy(p), BackendKind.Hip, 0)
    assert ok, err
    return s


def test_grf_terrain_5001x2001_full_size():
    w = workloads.grf_terrain_5001x2001(seed=2, precision=Precision.Float32, steps=400)
    p32 = w.params
    p64 = copy.deepcopy(p32)
    p64.precision = Precision.Float64
    print("[config] 5001x2001 GRF terrain", flush=True)
    plan = CpuSession2D("orc", p64)  # the oracle's construction of the same model (not stepped)
    runs = {}
    for label, p in (("f32", p32), ("f64", p64)):
        s = session(p)
        assert s.dt() == plan.dt and s.totalSteps() == plan.total == 400
        assert s.layout()["nx"] == plan.layout()["nx"] and s.layout()["nz"] == plan.layout()["nz"]
        assert np.array_equal(s.probe_indices(), plan.probe_indices())
        assert s.kernel_timing()["material_by_row"] is False
        assert s.advance(400)[0] and s.step() == 400
        d = s.diagnostics()
        assert d["finite"] and d["energy"] > 0
        r = s.receiver_arrays()
        assert r["vx"].shape == (1001, 401)
        runs[label] = gather(r)
        s.close()
    plan.close()
    assert rel_l2(runs["f64"], runs["f32"]) <= 1e-4


def test_gsls_batch_64_shots_equal_alone():
    steps = 100
    shots = workloads.gsls_batch(shots=64, precision=Precision.Float32, steps=steps)
    print("[config] 64-shot GSLS-3 batch", flush=True)
    sims = [session(p) for p in shots]
    batch = Batch2D(0)
    for s in sims:
        batch.add(s)
    ok, err = batch.advance(steps)
    assert ok, err
    for i, (p, s) in enumerate(zip(shots, sims)):
        assert s.step() == steps, i
        a = session(p)
        assert a.advance(steps)[0]
        ra, rb = a.receiver_arrays(), s.receiver_arrays()
        assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"]), i
        a.close()
        q = copy.deepcopy(p)
        q.precision = Precision.Float64
        d = session(q)
        assert d.advance(steps)[0]
        assert rel_l2(gather(d.receiver_arrays()), gather(rb)) <= 1e-4, i
        d.close()
    batch.close()
    for s in sims:
        s.close()


def test_serving_stream_results_equal_alone():
    steps = 100
    reqs = list(workloads.serving_requests(count=12, steps=steps))
    print("[config] serving stream, 12 requests", flush=True)
    report = Server().run(copy.deepcopy(reqs))
    assert all(not r.error for r in report.results)
    for r, (i, _, p) in zip(report.results, reqs):
        assert r.steps == steps
        a = session(p)
        assert a.advance(steps)[0]
        ra = a.receiver_arrays()
        assert np.array_equal(r.traces["vx"], ra["vx"]) and np.array_equal(r.traces["vz"], ra["vz"]), i
        a.close()
    s = report.summary()
    assert s["completed"] == 12 and 0 < s["p50_s"] <= s["p99_s"]
