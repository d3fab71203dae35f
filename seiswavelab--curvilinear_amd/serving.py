"""Continuous-batching shot server on one GPU.

Requests (2D shots) arrive over time.  The server admits a request when its estimated device
footprint fits the HBM budget, builds its session (FP64 model construction + upload) on a pool of
builder threads, and adds it to the running batch between chunks of steps, so a request that arrives
while others are on the device joins them as soon as memory allows; finished shots leave the batch
and return their traces.  Every shot keeps the single-session semantics (its result is the same alone
and in the mix).  Latency is completion time minus arrival time.

No counterpart in the reference (its sessions step one shot each, src/two_d/AsyncSimulation2D.cpp
runs one session per background worker); its per-session memory model (MemoryBudget2D.hpp:12-38) is
the basis of the footprint estimate here.
"""
from __future__ import annotations

import concurrent.futures as cf
import copy
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .params import Absorber, BackendKind, Parameters, Precision
from .simulation import Batch2D, Simulation, memoryBudget


def footprint_bytes(p: Parameters) -> int:
    """Device bytes a session of p allocates at most: the HIP engine's budget, the device part of the reference's
    per-session memory model (MemoryBudget2D.hpp:12-38) computed for this library's engine."""
    return int(memoryBudget(p, BackendKind.Hip)["deviceBytes"])


@dataclass
class Result:
    index: int
    arrival: float
    admitted: float = 0.0
    started: float = 0.0
    finished: float = 0.0
    steps: int = 0
    nodes: int = 0
    footprint: int = 0
    error: str = ""
    traces: Optional[Dict[str, np.ndarray]] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeReport:
    results: List[Result]
    wall: float
    offered_node_steps: float
    stats: dict = field(default_factory=dict)

    def latencies(self):
        return np.array([r.latency for r in self.results if not r.error])

    def summary(self) -> dict:
        lat = self.latencies()
        done = [r for r in self.results if not r.error]
        return {"requests": len(self.results), "completed": len(done),
                "shots_per_s": round(len(done) / self.wall, 4) if self.wall > 0 else 0.0,
                "p50_s": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
                "p99_s": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
                "mean_s": round(float(lat.mean()), 4) if len(lat) else None,
                "makespan_s": round(self.wall, 4)}


class Server:
    """budget_bytes: HBM admission budget; chunk: steps between admission / completion checks; options: engine
    switches (Simulation.set_option) applied to every session when it is built; batch_options: runner switches
    (Batch2D.set_option) of the batch the sessions join; close_in_background: finished sessions are closed on the
    builder threads (freeing device memory waits for the device, which would otherwise stall the loop between
    chunks; two threads of their own, so that closes never queue behind builds); the run's wall time still includes
    every close."""

    def __init__(self, device: int = 0, budget_bytes: float = 200e9, chunk: int = 64, builders: int = 4,
                 keep_traces: bool = True, largest_first: bool = True, options: Dict[str, int] = None,
                 batch_options: Dict[str, int] = None, close_in_background: bool = True):
        self.device, self.budget, self.chunk, self.builders = device, budget_bytes, chunk, builders
        self.keep_traces = keep_traces
        self.options = dict(options or {})
        self.batch_options = dict(batch_options or {})
        self.close_in_background = close_in_background
        # requests admitted together (a backlog) are built largest first, so that the last sessions to join
        # the batch are the quick ones; admission itself stays in arrival order
        self.largest_first = largest_first

    @staticmethod
    def _build(p: Parameters, device: int, options: Dict[str, int]):
        s = Simulation()
        ok, err = s.initialize(p, BackendKind.Hip, device)
        if ok:
            for name, value in options.items():
                s.set_option(name, value)
        return s, ("" if ok else err)

    def run(self, requests) -> ServeReport:
        """requests: iterable of (index, arrival_s, Parameters); each shot runs its full step count."""
        reqs = sorted(list(requests), key=lambda r: r[1])
        results = {i: Result(i, a) for i, a, _ in reqs}
        batch = Batch2D(self.device)
        for name, value in self.batch_options.items():
            batch.set_option(name, value)
        pool = cf.ThreadPoolExecutor(max_workers=self.builders)
        closer = cf.ThreadPoolExecutor(max_workers=2)
        queue = list(reqs)  # not yet admitted, in arrival order
        building: Dict[int, cf.Future] = {}
        closing: List[cf.Future] = []
        running: Dict[int, Simulation] = {}
        used, self.max_running, self.max_used = 0, 0, 0
        offered = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while queue or building or running:
            # admission: arrived requests, FIFO, while the footprint fits the budget
            admitted = []
            while queue and queue[0][1] <= now():
                i, arrival, p = queue[0]
                need = footprint_bytes(p)
                if running or building:
                    if used + need > self.budget:
                        break
                queue.pop(0)
                used += need
                self.max_used = max(self.max_used, used)
                r = results[i]
                r.admitted, r.footprint = now(), need
                admitted.append((need, i, p))
            if self.largest_first:
                admitted.sort(key=lambda x: -x[0])
            for _, i, p in admitted:
                building[i] = pool.submit(self._build, p, self.device, self.options)
            # built sessions join the batch between chunks
            for i in [i for i, f in building.items() if f.done()]:
                s, err = building.pop(i).result()
                r = results[i]
                if err:
                    r.error, r.finished = err, now()
                    used -= r.footprint
                    continue
                L = s.layout()
                r.nodes = (L["nx"] - 6) * (L["nz"] - 6)
                r.steps = s.totalSteps()
                offered += r.nodes * r.steps
                r.started = now()
                batch.add(s)
                running[i] = s
                self.max_running = max(self.max_running, len(running))
            if running:
                # every member advances up to `chunk` steps and stops at its own total
                ok, err = batch.advance(self.chunk)
                if not ok:
                    raise RuntimeError(err)
                for i in [i for i, s in running.items() if s.step() >= s.totalSteps() or s.fault()]:
                    s = running.pop(i)
                    r = results[i]
                    r.finished = now()
                    r.error = s.fault()
                    if self.keep_traces:
                        r.traces = s.receiver_arrays()
                    batch.remove(s)
                    if self.close_in_background:
                        closing.append(closer.submit(s.close))
                    else:
                        s.close()
                    used -= r.footprint
            elif building:
                time.sleep(0.0005)
            elif queue:
                time.sleep(max(0.0, min(0.002, queue[0][1] - now())))
        for f in closing:
            f.result()
        wall = now()
        stats = batch.stats()
        batch.close()
        pool.shutdown(wait=True)
        closer.shutdown(wait=True)
        return ServeReport([results[i] for i, _, _ in reqs], wall, offered, stats)


def serve_sequential(requests, device: int = 0, precision=None, keep_traces: bool = True) -> ServeReport:
    """The plain path: one shot at a time in arrival order (each starts when it has arrived and the
    previous one has finished), optionally forcing the precision (e.g. FP64)."""
    reqs = sorted(list(requests), key=lambda r: r[1])
    results = []
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    for i, arrival, p in reqs:
        if precision is not None:
            p = copy.copy(p)
            p.precision = precision
        while now() < arrival:
            time.sleep(min(0.002, arrival - now()))
        r = Result(i, arrival, admitted=now())
        s = Simulation()
        ok, err = s.initialize(p, BackendKind.Hip, device)
        if not ok:
            r.error, r.finished = err, now()
            results.append(r)
            continue
        r.started = now()
        r.steps = s.totalSteps()
        ok, err = s.advance(r.steps)
        r.finished = now()
        r.error = "" if ok else err
        if keep_traces:
            r.traces = s.receiver_arrays()
        s.close()
        results.append(r)
    return ServeReport(results, now(), 0.0)
