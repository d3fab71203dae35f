#!/usr/bin/env python3
"""Headline benchmark: 2D elastic P-SV, 5001x2001 nodes at h=2 m (30-node external CFS-PML,
5061x2031 computational nodes), FP32, on one MI355X.

A "step" is one RK4 time step of the whole grid: four fused stage-kernel launches; the device-side
inspection / fault rule / receiver gather of a step runs inside the next step's first launch (a run's
last step has its own small kernels), inputs resident in HBM.  ``value`` is computational node updates
per second (G pt-updates/s) over the timed steps, which carry no timing events.  ``roofline`` divides
the algorithmic bytes of one stage-kernel launch (400 B per computational node per step for FP32
elastic, i.e. 100 B per node per launch) by the average stage-kernel launch time measured with HIP
events on the engine's stream over a run of steps right after the timed ones.  ``cpu_baseline`` times the CPU oracle (a bitwise restatement of
the reference's CPU/OpenMP path) on a bounded sample of the same model on the host cores.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "seiswavelab--curvilinear_amd"
HBM_PEAK_GBS = 8000.0
ALGO_BYTES = {"f32": 400.0, "f64": 800.0}  # per computational node per RK4 step, 2D elastic (5 fields, 5 coefficients)


def _package():
    name = "seiswavelab_curvilinear_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def advance_exactly(sim, count, label):
    """advance() stops at the session's total without complaint; the bench must time exactly `count` steps."""
    step = getattr(sim, "currentStep", None) or sim.step  # the 3D session keeps the reference's name
    before = step()
    ok, err = sim.advance(count)
    if not ok:
        raise RuntimeError(f"{label}: {err}")
    taken = step() - before
    if taken != count:
        raise RuntimeError(f"{label}: asked for {count} steps, the session took {taken} (total {sim.totalSteps()})")
    return taken


def profile_steps(steps):
    """Steps of the event-timed run that follows a timed run (at least 32: the finalisation is sampled every 32nd
    step)."""
    return max(32, min(steps, 200))


def measure(pkg, params, steps, warmup, label, options=None):
    sim = pkg.Simulation()
    t = time.perf_counter()
    ok, err = sim.initialize(params, pkg.BackendKind.Hip, 0)
    if not ok:
        raise RuntimeError(err)
    log(f"[bench] {label}: initialized in {time.perf_counter() - t:.1f} s, dt={sim.dt():.6g} s, "
        f"{sim.totalSteps()} steps in the session")
    for name, value in (options or {}).items():
        sim.set_option(name, value)
    advance_exactly(sim, warmup, label)
    sim.profile(False)
    t0 = time.perf_counter()
    taken = advance_exactly(sim, steps, label)
    t1 = time.perf_counter()
    # per-launch timing with HIP events on the engine's stream, over steps run after the timed ones: events
    # recorded between launches add their own gaps, so they stay out of the timed steps
    sim.profile(True)
    advance_exactly(sim, profile_steps(steps), label)
    timing = sim.kernel_timing()
    sim.profile(False)
    L = sim.layout()
    nodes = (L["nx"] - 6) * (L["nz"] - 6)
    wall = t1 - t0
    log(f"[bench] {label}: {taken} steps in {wall:.3f} s ({1e3 * wall / taken:.3f} ms/step)")
    d = sim.diagnostics()
    sim.close()
    return {"wall": wall, "steps": taken, "timing": timing, "nodes": nodes, "finite": d["finite"],
            "energy": d["energy"]}


def accuracy(pkg, params, steps):
    """FP32 against FP64 on the same model and steps: relative L2 and peak error of the whole receiver gather
    (Vx and Vz of every receiver) and of the final wavefield snapshot (all five fields, each relative to its
    own FP64 peak)."""
    import copy
    import numpy as np
    runs = {}
    for label, prec in (("f32", pkg.Precision.Float32), ("f64", pkg.Precision.Float64)):
        p = copy.deepcopy(params)
        p.precision = prec
        s = pkg.Simulation()
        ok, err = s.initialize(p, pkg.BackendKind.Hip, 0)
        if not ok:
            raise RuntimeError(err)
        advance_exactly(s, steps, f"accuracy {label}")
        r = s.receiver_arrays()
        runs[label] = (np.concatenate([r["vx"].ravel(), r["vz"].ravel()]), [s.snapshot_field(f) for f in range(5)])
        s.close()
    (g32, f32), (g64, f64) = runs["f32"], runs["f64"]
    rel = lambda a, b: float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)))  # noqa: E731
    peak = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))  # noqa: E731
    out = {"steps": steps, "receivers": int(len(g32) // 2 // max(1, steps + 1)),
           "gather_rel_l2": rel(g32, g64), "gather_peak_error": peak(g32, g64),
           "snapshot_rel_l2": max(rel(a, b) for a, b in zip(f32, f64)),
           "snapshot_peak_error": max(peak(a, b) for a, b in zip(f32, f64))}
    log(f"[bench] FP32 vs FP64 after {steps} steps: {out}")
    return out


def roofline(res, prec):
    t = res["timing"]
    avg_ms = t["stage_ms"] / max(t["stage_launches"], 1)
    # the dominant kernel is the fused RK-stage kernel; one launch updates every computational node
    # (interior and border tiles), and a step is four such launches
    per_launch = ALGO_BYTES[prec] / 4 * t["nodes_per_launch"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(avg_ms, 5),
            "algorithmic_bytes_per_launch": per_launch, "nodes_per_launch": t["nodes_per_launch"],
            "metric_on_the_fly": t["metric_on_the_fly"], "material_by_row": t["material_by_row"], "vgprs": t["vgprs"]}


def host_cpu():
    """CPU model, logical CPUs, CPUs this process may run on, physical cores and SMT, OpenMP environment."""
    model, cores = "", set()
    phys = core = None
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            key, _, value = line.partition(":")
            key, value = key.strip(), value.strip()
            if key == "model name" and not model:
                model = value
            elif key == "physical id":
                phys = value
            elif key == "core id":
                core = value
            elif not key and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    physical = len(cores) or logical
    smt = None
    try:
        smt = pathlib.Path("/sys/devices/system/cpu/smt/active").read_text().strip() == "1"
    except OSError:
        pass
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES", "OMP_WAIT_POLICY")}
    return {"model": model, "logical_cpus": logical, "usable_cpus": usable, "physical_cores": physical,
            "smt_active": smt, "omp_env": env}


def cpu_baseline(pkg, params, steps=4, warmup=1, repeats=3):
    """The reference's FP64 CPU path, restated by the oracle (OpenMP rows-parallel RHS, fields-parallel
    combine, serial inspection, as src/two_d/CurvilinearSolver2D.cpp:68-119), on the same 5061x2031 model:
    `repeats` runs of `steps` production advance() steps after `warmup` steps (median reported), at the
    reference's default thread count min(8, nproc) (CurvilinearSolver2D.cpp:278) and at its cap of 64
    (:183-185) limited to the physical cores this process may use."""
    import copy
    import statistics
    sys.path.insert(0, str(ROOT / "tests"))
    from cpu_sessions import CpuSession2D

    info = host_cpu()
    counts = sorted({min(8, info["usable_cpus"]), min(64, info["physical_cores"], info["usable_cpus"])})
    p = copy.deepcopy(params)
    p.precision = pkg.Precision.Float64
    runs = {}
    for threads in counts:
        p.cpuThreads = threads
        t = time.perf_counter()
        sess = CpuSession2D("orc", p)
        log(f"[bench] cpu FP64 oracle: model built in {time.perf_counter() - t:.1f} s; {threads} threads")
        L = sess.layout()
        nodes = (L["nx"] - 6) * (L["nz"] - 6)
        walls = []
        for _ in range(repeats):
            ok, err = sess.advance(warmup)
            t0 = time.perf_counter()
            ok2, err2 = sess.advance(steps)
            walls.append(time.perf_counter() - t0)
            if not (ok and ok2):
                raise RuntimeError(err or err2)
        sess.close()
        wall = statistics.median(walls)
        runs[threads] = {"value": round(nodes * steps / wall / 1e9, 5), "ms_per_step": round(1e3 * wall / steps, 1),
                         "repeats_s": [round(w, 3) for w in walls]}
        log(f"[bench] cpu FP64 oracle, {threads} threads: {runs[threads]}")
    best = max(runs, key=lambda k: runs[k]["value"])
    calib = None
    cpath = ROOT / "oracle" / "calibration.json"
    if cpath.exists():
        calib = json.loads(cpath.read_text())
    ratio = calib["cases"]["elastic_cfs_pml_2048"]["oracle_over_reference"] if calib else None
    return {"value": runs[best]["value"], "unit": "G pt-updates/s", "cores": best, "kind": "port",
            "reference_equivalent_value": round(runs[best]["value"] * ratio, 5) if ratio else None,
            "sample": f"same 5061x2031-node model in FP64 (the reference's CPU precision), median of {repeats} runs "
                      f"of {steps} production advance() steps (RHS, combine, projection, energy inspection, "
                      f"receivers) each after {warmup} warm-up step(s), OpenMP as the reference; thread counts "
                      f"min(8, nproc) and min(64, physical cores)",
            "by_threads": runs, "host": info, "oracle_vs_reference_build": calib,
            "note": "reference_equivalent_value = value x the oracle/reference per-step time ratio measured on the "
                    "2048^2 elastic CFS-PML case at 8 threads where both builds exist (oracle/calibration.json)"}


ALGO_BYTES_3D = {"f32": 784.0, "f64": 1568.0}  # per node per RK4 step, 3D nine-field (9 fields, 13 coefficients)


def measure_3d(pkg, steps, warmup, precision):
    """400x288x200 DEM-terrain 3D elastic shot; FP64 (precision 0) or FP32 (1)."""
    from seiswavelab_curvilinear_amd import workloads
    p, seed = workloads.dem_terrain_400x288x200(steps=max(1500, warmup + steps + max(4, min(steps, 20))))
    sim = pkg.Simulation3D(pkg.BackendKind.Hip)
    t = time.perf_counter()
    ok, err = sim.initialize(p, 0, precision)
    if not ok:
        raise RuntimeError(err)
    label = "FP64" if precision == 0 else "FP32"
    log(f"[bench] 3D {label}: initialized in {time.perf_counter() - t:.1f} s, dt={sim.dt():.6g} s")
    advance_exactly(sim, warmup, f"3D {label}")
    sim.profile(False)
    t0 = time.perf_counter()
    steps = advance_exactly(sim, steps, f"3D {label}")
    wall = time.perf_counter() - t0
    # per-launch timing with HIP events over steps run after the timed ones (as for 2D)
    sim.profile(True)
    advance_exactly(sim, max(4, min(steps, 20)), f"3D {label}")
    timing = sim.kernel_timing()
    d = sim.diagnostics()
    sim.close()
    nodes = timing["nodes_per_launch"]
    avg_ms = timing["stage_ms"] / max(timing["stage_launches"], 1)
    prec = "f64" if precision == 0 else "f32"
    per_launch = ALGO_BYTES_3D[prec] / 4 * nodes
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    log(f"[bench] 3D {label}: {steps} steps in {wall:.3f} s ({1e3 * wall / steps:.3f} ms/step)")
    return {"value": round(nodes * steps / wall / 1e9, 4), "unit": "G pt-updates/s",
            "ms_per_step": round(1e3 * wall / steps, 4), "dtype": prec, "seed": seed, "finite": d["allFinite"],
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "avg_launch_ms": round(avg_ms, 5),
                         "algorithmic_bytes_per_launch": per_launch, "nodes_per_launch": nodes,
                         "metric_planes": timing["metric_planes"], "material_by_depth": timing["material_by_depth"],
                         "vgprs": timing["vgprs"]}}


def measure_batch(pkg, shots, steps):
    """64-shot GSLS-3 viscoelastic batch (FP32 state, FP64 metrics / closure) through one Batch2D versus the
    plain single-shot FP64 path (each shot alone, one after another).  Both paths run every shot's full
    `steps` steps and both are timed from the first session build to the last step: shots/s = shots / that."""
    import numpy as np
    from seiswavelab_curvilinear_amd import workloads
    p32 = workloads.gsls_batch(shots=shots, precision=pkg.Precision.Float32, steps=steps)
    p64 = workloads.gsls_batch(shots=shots, precision=pkg.Precision.Float64, steps=steps)
    t0 = time.perf_counter()
    batch, sims = pkg.Batch2D.build(p32, 0, workers=8)  # sessions built on 8 host threads
    t1 = time.perf_counter()
    ok, err = batch.advance(steps)
    t2 = time.perf_counter()
    if not ok:
        raise RuntimeError(err)
    if any(s.step() != steps for s in sims):
        raise RuntimeError(f"batch: not every shot ran {steps} steps: {[s.step() for s in sims]}")
    stats = batch.stats()
    nodes = sum((s.layout()["nx"] - 6) * (s.layout()["nz"] - 6) for s in sims)
    log(f"[bench] batch: {shots} FP32 GSLS-3 sessions built in {t1 - t0:.2f} s "
        f"({nodes / shots / 1e6:.3f} M computational nodes per shot on average), {steps} steps of every shot "
        f"in {t2 - t1:.2f} s; stats {stats}")
    plain_build = plain_run = 0.0
    errs = []
    for i, p in enumerate(p64):
        t3 = time.perf_counter()
        s = pkg.Simulation()
        ok, err = s.initialize(p, pkg.BackendKind.Hip, 0)
        if not ok:
            raise RuntimeError(err)
        t4 = time.perf_counter()
        advance_exactly(s, steps, f"plain FP64 shot {i}")
        plain_build += t4 - t3
        plain_run += time.perf_counter() - t4
        a, b = s.receiver_arrays(), sims[i].receiver_arrays()
        ref = np.concatenate([a["vx"].ravel(), a["vz"].ravel()])
        val = np.concatenate([b["vx"].ravel(), b["vz"].ravel()])
        errs.append(float(np.sqrt(np.sum((ref - val) ** 2) / max(np.sum(ref * ref), 1e-300))))
        s.close()
    batch.close()
    for s in sims:
        s.close()
    batch_total, plain_total = t2 - t0, plain_build + plain_run
    log(f"[bench] plain FP64: {shots} shots one after another, builds {plain_build:.2f} s + steps {plain_run:.2f} s")
    return {"shots": shots, "steps_per_shot": steps, "mean_nodes_per_shot": round(nodes / shots),
            "batched_shots_per_s": round(shots / batch_total, 4), "plain_fp64_shots_per_s": round(shots / plain_total, 4),
            "speedup": round(plain_total / batch_total, 3),
            "batched_s": {"builds": round(t1 - t0, 3), "steps": round(t2 - t1, 3)},
            "plain_fp64_s": {"builds": round(plain_build, 3), "steps": round(plain_run, 3)},
            "speedup_steps_only": round(plain_run / (t2 - t1), 3),
            "batched_ms_per_step_all_shots": round(1e3 * (t2 - t1) / steps, 4),
            "batched_G_pt_updates_per_s": round(nodes * steps / (t2 - t1) / 1e9, 4),
            "max_gather_rel_l2_fp32_batch_vs_fp64_plain": max(errs), "graph": stats,
            "note": "every shot runs its full step count on both paths; times run from the first session build "
                    "to the last step"}


def measure_serving(pkg, count, steps, rate=20.0):
    """Poisson stream (20 req/s) of 2D shots (elastic / GSLS-3, FP32) through the continuous-batching server
    versus the plain path (one shot at a time, FP64, arrival order)."""
    import copy
    import numpy as np
    from seiswavelab_curvilinear_amd import serving, workloads
    reqs = list(workloads.serving_requests(count=count, rate=rate, steps=steps))
    window = reqs[-1][1]
    log(f"[bench] serving: {count} requests over {window:.2f} s of arrivals, {steps} steps each")
    served = serving.Server().run(copy.deepcopy(reqs))
    log(f"[bench] serving: {count} requests, makespan {served.wall:.2f} s; {served.summary()}")
    plain = serving.serve_sequential(copy.deepcopy(reqs), precision=pkg.Precision.Float64)
    log(f"[bench] plain FP64 sequential: makespan {plain.wall:.2f} s; {plain.summary()}")
    errs = []
    for a, b in zip(served.results, plain.results):
        ga = np.concatenate([a.traces["vx"].ravel(), a.traces["vz"].ravel()])
        gb = np.concatenate([b.traces["vx"].ravel(), b.traces["vz"].ravel()])
        errs.append(float(np.sqrt(np.sum((ga - gb) ** 2) / max(np.sum(gb * gb), 1e-300))))
    node_steps = sum(r.nodes * r.steps for r in served.results)
    s, q = served.summary(), plain.summary()
    # capacity: the same requests all present at t = 0, so neither path waits for arrivals
    burst = [(i, 0.0, p) for i, _, p in copy.deepcopy(reqs)]
    served_sat = serving.Server().run(copy.deepcopy(burst))
    plain_sat = serving.serve_sequential(copy.deepcopy(burst), precision=pkg.Precision.Float64, keep_traces=False)
    ss, qs = served_sat.summary(), plain_sat.summary()
    log(f"[bench] serving at saturation: served {ss}; plain FP64 {qs}")
    return {"requests": count, "rate_per_s": rate, "steps_per_request": steps,
            "mean_nodes": round(np.mean([r.nodes for r in served.results])),
            "arrival_window_s": round(window, 3), "offered_shots_per_s": round(count / window, 3),
            "offered_G_pt_updates_per_s": round(node_steps / window / 1e9, 4),
            "served_G_pt_updates_per_s": round(node_steps / served.wall / 1e9, 4),
            "served": s, "plain_fp64_sequential": q,
            "shots_per_s_ratio": round(s["shots_per_s"] / q["shots_per_s"], 3),
            "p99_ratio_plain_over_served": round(q["p99_s"] / s["p99_s"], 3),
            "max_gather_rel_l2_served_fp32_vs_plain_fp64": max(errs), "graph": served.stats,
            "saturation": {"served": ss, "plain_fp64_sequential": qs,
                           "shots_per_s_ratio": round(ss["shots_per_s"] / qs["shots_per_s"], 3),
                           "served_G_pt_updates_per_s": round(node_steps / served_sat.wall / 1e9, 4),
                           "note": "the same requests all arriving at t = 0: throughput without the 20 req/s cap"},
            "note": "each request's latency = completion - arrival; both paths see the same arrival times"}


def traffic_from_profile(prec):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary, if present."""
    path = ROOT / "profiles" / f"stage_kernel_elastic2d_5061x2031_{prec}_pmc.json"
    if not path.exists():
        return None
    try:
        data = json.loads(path.read_text())
        return data.get("traffic_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nx", type=int, default=5001)
    ap.add_argument("--nz", type=int, default=2001)
    ap.add_argument("--nt", type=int, default=1, help="nontemporal stores in the interior kernel")
    ap.add_argument("--stored-metrics", type=int, default=0, help="read a, b, J from memory instead of recomputing")
    ap.add_argument("--accuracy-steps", type=int, default=4000,
                    help="steps of the FP32-vs-FP64 seismogram / snapshot comparison on the 5001x2001 shot "
                         "(default: all 4000 of the workload; 0 = skip)")
    ap.add_argument("--steps-3d", type=int, default=40, help="timed steps of the 400x288x200 3D shot (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=3000,
                    help="steps of each of the 64 GSLS shots (default: all 3000 of the workload; 0 = skip)")
    ap.add_argument("--batch-shots", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=120, help="requests of the serving stream (0 = skip)")
    ap.add_argument("--serve-steps", type=int, default=500, help="steps per served request")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30,
                    help="size of the plain copy used as the sustained-bandwidth reference (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: a shot (or a batch of shots) runs on one device, and "
              "runs across several GPUs are not part of this benchmark.", file=sys.stderr)
        sys.exit(2)
    pkg = _package()
    from seiswavelab_curvilinear_amd import workloads

    t = time.perf_counter()
    # the workload's own 4000 steps (more only if asked to time more); every leg checks the steps it ran
    w = workloads.grf_terrain_5001x2001(seed=2, precision=pkg.Precision.Float32, nx=args.nx, nz=args.nz,
                                        steps=max(4000, args.warmup + args.steps + profile_steps(args.steps),
                                                  10 + args.fp64_steps + profile_steps(args.fp64_steps),
                                                  args.accuracy_steps))
    log(f"[bench] 5001x2001 GRF-terrain model generated in {time.perf_counter() - t:.1f} s")
    copy = None
    if args.probe_bytes > 0:
        ms = pkg.native.stream_probe(args.probe_bytes, 10)
        copy = {"bytes": args.probe_bytes, "ms": round(ms, 5), "GBs": round(2 * args.probe_bytes / (ms * 1e-3) / 1e9, 1)}
        log(f"[bench] plain copy of {args.probe_bytes} B: {ms:.4f} ms = {copy['GBs']} GB/s (read+write)")
    options = {"nontemporal_stores": args.nt, "stored_metrics": args.stored_metrics}
    r32 = measure(pkg, w.params, args.steps, args.warmup, "FP32", options)
    roof32 = roofline(r32, "f32")
    roof32["traffic"] = traffic_from_profile("f32")
    if copy:
        roof32["copy_GBs"] = copy["GBs"]
        roof32["frac_of_copy"] = round(roof32["achieved"] / copy["GBs"], 4)
    extra = {}
    if args.fp64_steps > 0:
        w.params.precision = pkg.Precision.Float64
        r64 = measure(pkg, w.params, args.fp64_steps, min(args.warmup, 10), "FP64", options)
        roof64 = roofline(r64, "f64")
        roof64["traffic"] = traffic_from_profile("f64")
        extra["fp64"] = {"value": round(r64["nodes"] * r64["steps"] / r64["wall"] / 1e9, 4), "steps": r64["steps"],
                         "ms_per_step": round(1e3 * r64["wall"] / r64["steps"], 4), "roofline": roof64}
        w.params.precision = pkg.Precision.Float32
    if args.accuracy_steps > 0:
        extra["fp32_vs_fp64"] = accuracy(pkg, w.params, args.accuracy_steps)
    if args.steps_3d > 0:
        extra["3d_elastic_400x288x200"] = {"fp64": measure_3d(pkg, args.steps_3d, 5, 0),
                                           "fp32": measure_3d(pkg, args.steps_3d, 5, 1)}
    if args.batch_steps > 0:
        extra["batch_gsls3"] = measure_batch(pkg, args.batch_shots, args.batch_steps)
    if args.serve_requests > 0:
        extra["serving"] = measure_serving(pkg, args.serve_requests, args.serve_steps)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pkg, w.params)
    value = r32["nodes"] * r32["steps"] / r32["wall"] / 1e9
    info = pkg.native.device_query(0)
    line = {
        "metric": "Gpt-updates/s & %HBM peak, 2D elastic FP32 vs FP64; batched shots/s, p99",
        "value": round(value, 4),
        "unit": "G pt-updates/s",
        "n_gpus": 1,
        "steps": r32["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["wall"] / r32["steps"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded GRF topography and 5% RMS velocity perturbation, seed 2)",
        "config": {"workload": "2D elastic P-SV single shot, 5001x2001 nodes h=2 m + 30-node external CFS-PML "
                               "(5061x2031 computational nodes), 30 Hz explosive source, 1001 surface receivers",
                   "computational_nodes": r32["nodes"], "device": info.arch.decode(), "finite": r32["finite"]},
        "roofline": roof32,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
