"""Continuous-batching server (seiswavelab--curvilinear_amd/serving.py) on the GPU: every request's
traces are identical alone and in the mix; admission respects the HBM budget; latency percentiles."""
import copy

import numpy as np
import pytest

from seiswavelab_curvilinear_amd import BackendKind, GslsSettings, Precision, Simulation, workloads
from seiswavelab_curvilinear_amd.serving import Server, footprint_bytes, serve_sequential

pytestmark = pytest.mark.gpu


def small_requests(n=6, steps=60):
    """Staggered arrivals, mixed sizes, elastic and GSLS-3, FP32 (models as the 5001x2001 workload)."""
    out = []
    for i in range(n):
        nx, nz = 120 + 37 * i, 70 + 11 * i
        w = workloads.grf_terrain_5001x2001(seed=40 + i, precision=Precision.Float32, nx=nx, nz=nz, steps=steps,
                                            topo_rms=min(25.0, 0.05 * nz * 2.0))
        p = w.params
        p.receiverCount = min(p.receiverCount, 30)
        if i % 2:
            p.attenuation = GslsSettings(enabled=True, mechanisms=3, minFrequency=6.0, maxFrequency=90.0,
                                         referenceFrequency=30.0, qp=40.0, qs=20.0)
        out.append((i, 0.004 * i, p))
    return out


def alone(p):
    s = Simulation()
    ok, err = s.initialize(copy.deepcopy(p), BackendKind.Hip, 0)
    assert ok, err
    ok, err = s.advance(s.totalSteps())
    assert ok, err
    a = s.receiver_arrays()
    s.close()
    return a


@pytest.mark.parametrize("close_in_background", [True, False])
def test_each_result_is_the_same_alone_and_in_the_mix(close_in_background):
    reqs = small_requests()
    report = Server(chunk=16, close_in_background=close_in_background).run(copy.deepcopy(reqs))
    assert all(not r.error for r in report.results)
    for r, (i, _, p) in zip(report.results, reqs):
        ref = alone(p)
        assert r.steps == p.maxSteps
        assert np.array_equal(r.traces["vx"], ref["vx"]) and np.array_equal(r.traces["vz"], ref["vz"]), i
    s = report.summary()
    assert s["completed"] == len(reqs) and 0 < s["p50_s"] <= s["p99_s"]


def test_admission_respects_the_budget():
    reqs = small_requests()
    budget = max(footprint_bytes(p) for _, _, p in reqs) * 2.2
    server = Server(budget_bytes=budget, chunk=16)
    report = server.run(copy.deepcopy(reqs))
    assert all(not r.error for r in report.results)
    assert server.max_used <= budget and server.max_running < len(reqs)
    for r, (i, _, p) in zip(report.results, reqs):
        assert np.array_equal(r.traces["vz"], alone(p)["vz"]), i


def test_sequential_fp64_path_matches_fp32_within_tolerance():
    reqs = small_requests(n=3)
    fast = Server(chunk=16).run(copy.deepcopy(reqs))
    slow = serve_sequential(copy.deepcopy(reqs), precision=Precision.Float64)
    for a, b in zip(fast.results, slow.results):
        ga = np.concatenate([a.traces["vx"].ravel(), a.traces["vz"].ravel()])
        gb = np.concatenate([b.traces["vx"].ravel(), b.traces["vz"].ravel()])
        assert np.sqrt(np.sum((ga - gb) ** 2) / np.sum(gb * gb)) <= 1e-4


def test_session_and_batch_options_reach_the_server_path():
    """Server(options=..., batch_options=...): engine switches that change only the schedule (no nontemporal
    stores, two lanes of 5 steps) leave every trace equal to the request's lone run; a switch that changes FP32
    rounding (PML lanes forced on) keeps it within the FP32 bar of the lone run."""
    reqs = small_requests(n=4)
    same = Server(chunk=16, options={"nontemporal_stores": 0},
                  batch_options={"lanes": 2, "lane_steps": 5}).run(copy.deepcopy(reqs))
    pml = Server(chunk=16, options={"pml_vector": 1}).run(copy.deepcopy(reqs))
    assert same.stats["lanes"] == 2
    for r, q, (i, _, p) in zip(same.results, pml.results, reqs):
        ref = alone(p)
        assert not r.error and not q.error
        assert np.array_equal(r.traces["vx"], ref["vx"]) and np.array_equal(r.traces["vz"], ref["vz"]), i
        ga = np.concatenate([q.traces["vx"].ravel(), q.traces["vz"].ravel()])
        gb = np.concatenate([ref["vx"].ravel(), ref["vz"].ravel()])
        assert np.sqrt(np.sum((ga - gb) ** 2) / np.sum(gb * gb)) <= 2e-5, i
