#!/usr/bin/env python3
"""Calibrates the CPU oracle's speed against the reference's own CPU path (test infrastructure).

On the GPU box the CPU baseline is the oracle (a C restatement of the reference's OpenMP path: rows-parallel RHS,
fields-parallel combine, serial inspection), because the reference cannot travel there.  This script runs in the
container where the reference build exists (oracle/_ref/libswlref.so, `make -C oracle ref`) and times both on the
same models with the same thread count, so the oracle's figure can be read as the reference's:

* the 500x200 sine-terrain workload (three layers, imported surface, 20-node internal CFS-PML, force), FP64;
* the reference's 2048^2 elastic + CFS-PML performance case (docs/二维性能与精度.md:17-19: h = 2 m,
  Vp/Vs/rho = 1500/700/1800, relief 5 m over 500 m, 15 Hz force, CFL 0.5, 32 receivers every 0.25 ms,
  16-cell external CFS-PML), FP64.

Each is timed as the reference times it (production advance(), warm-up steps first, the median of repeats).
The ratio oracle / reference of the per-step times is written to oracle/calibration.json.

    python oracle/calibrate.py [--threads 8] [--repeats 3]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import statistics
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent / "tests"))
import conftest  # noqa: E402,F401  registers the package
from cpu_sessions import CpuSession2D, have_ref  # noqa: E402
from seiswavelab_curvilinear_amd import Absorber, Parameters, Precision, workloads  # noqa: E402


def cases():
    c1 = workloads.sine_terrain_500x200(seed=1, precision=Precision.Float64)
    big = Parameters(precision=Precision.Float64, nx=2048, nz=2048, spacing=2, vp=1500, vs=700, density=1800, relief=5,
                     wavelength=500, frequency=15, cfl=.5, duration=10, maxSteps=0, sourceZ=20,
                     absorber=Absorber.CfsPml, externalAbsorption=True, spongeX=16, spongeZ=16, receiverStart=0,
                     receiverCount=32, receiverStride=64, receiverDepth=0, sampleIntervalMs=.25)
    big.layered = False
    return {"sine_terrain_500x200": (c1, 20, 200), "elastic_cfs_pml_2048": (big, 3, 12)}


def time_kind(kind, p, warmup, steps, repeats):
    s = CpuSession2D(kind, p)
    walls = []
    for _ in range(repeats):
        assert s.advance(warmup)[0]
        t = time.perf_counter()
        ok, err = s.advance(steps)
        walls.append(time.perf_counter() - t)
        assert ok, err
    s.close()
    return statistics.median(walls) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    if not have_ref():
        sys.exit("oracle/_ref/libswlref.so is missing: build it with `make -C oracle ref` first")
    model = next((l.split(":", 1)[1].strip() for l in pathlib.Path("/proc/cpuinfo").read_text().splitlines()
                  if l.startswith("model name")), platform.machine())
    out = {"threads": args.threads, "repeats": args.repeats, "cpu": model, "nproc": os.cpu_count(), "cases": {}}
    for name, (p, warmup, steps) in cases().items():
        p.cpuThreads = args.threads
        # alternate the order so neither build always runs on a warmer machine
        ref = time_kind("ref", p, warmup, steps, args.repeats)
        orc = time_kind("orc", p, warmup, steps, args.repeats)
        ref2 = time_kind("ref", p, warmup, steps, args.repeats)
        ref_ms = 1e3 * min(ref, ref2)
        out["cases"][name] = {"steps_timed": steps, "reference_ms_per_step": round(ref_ms, 3),
                              "oracle_ms_per_step": round(1e3 * orc, 3), "oracle_over_reference": round(1e3 * orc / ref_ms, 4)}
        print(name, out["cases"][name], flush=True)
    (HERE / "calibration.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
