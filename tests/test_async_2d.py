"""Background stepping (csrc/async2d.cpp, AsyncSimulation) against direct stepping.

The reference's own check (src/two_d/PrecisionValidation2D.cpp:108-116): an FP32 GSLS + CFS-PML session
stepped in background batches (queue 8 steps, synchronize, poll) has the same fields, memories and
receiver samples, bit for bit, as a session advanced directly to the same step; display frames and reset
work through the façade.  Also: a poll-only loop to completion, pause, and the uninitialized façade.
"""
import time

import numpy as np
import pytest

from seiswavelab_curvilinear_amd import Absorber, AsyncSimulation, BackendKind, Precision, Simulation
from test_oracle_2d import precision_fixture, ref_case


def same_state(live, ref):
    a, b = live.snapshot(True), ref.snapshot(True)
    for x, y in zip(a["fields"], b["fields"]):
        assert np.array_equal(x, y)
    for key in ("syy", "memory", "boundaryMemory"):
        if key in b:
            xa, xb = a[key], b[key]
            if isinstance(xb, list):
                assert all(np.array_equal(x, y) for x, y in zip(xa, xb)), key
            else:
                assert np.array_equal(xa, xb), key
    ra, rb = live.receiver_arrays(), ref.receiver_arrays()
    assert np.array_equal(ra["vx"], rb["vx"]) and np.array_equal(ra["vz"], rb["vz"])


@pytest.mark.gpu
def test_background_publication_equals_direct():
    p = precision_fixture()
    p.precision, p.absorber = Precision.Float32, Absorber.CfsPml
    p.attenuation.enabled = True
    p.cpuThreads = 1
    print("[async] background publication", flush=True)
    live, ref = AsyncSimulation(), Simulation()
    assert live.initialize(p, BackendKind.Hip, 0)[0] and ref.initialize(p, BackendKind.Hip, 0)[0]
    ok, err, image = live.readFrame(0)
    assert ok, err
    for batch in range(4):
        assert live.queue(8, 1000)
        if batch < 2:  # let these finish: poll until published (synchronize cancels a batch between steps)
            deadline = time.time() + 30
            while live.busy() and time.time() < deadline:
                assert live.poll()[0]
                time.sleep(0.001)
        live.synchronize()
        ok, err = live.poll()
        assert ok, err
    assert 16 <= live.step() <= 32
    assert ref.advance(live.step())[0]
    same_state(live, ref)
    ok, err, image = live.readFrame(0)
    assert ok and image.shape == (p.nz, p.nx)
    t = live.backgroundTiming()
    assert t["batches"] == 4 and t["steps"] == live.step()
    ok, err = live.reset()
    assert ok and live.step() == 0
    live.close()


@pytest.mark.gpu
def test_poll_loop_to_completion_equals_direct():
    p = ref_case("gsls2_layered_pml")
    p.precision = Precision.Float32
    print("[async] poll loop", flush=True)
    live, ref = AsyncSimulation(), Simulation()
    assert live.initialize(p, BackendKind.Hip, 0)[0] and ref.initialize(p, BackendKind.Hip, 0)[0]
    deadline = time.time() + 60
    while not live.completed() and time.time() < deadline:
        if not live.busy():
            live.queue(8, 50)
        ok, err = live.poll()
        assert ok, err
        time.sleep(0.001)
    live.synchronize()
    assert live.poll()[0] and live.completed()
    assert ref.advance(ref.totalSteps())[0]
    same_state(live, ref)
    d, e = live.diagnostics(), ref.diagnostics()
    assert d["energy"] == e["energy"] and d["peakEnergy"] == e["peakEnergy"]
    live.close()


@pytest.mark.gpu
def test_pause_stops_between_steps():
    p = ref_case("sine_terrain_small")
    live = AsyncSimulation()
    assert live.initialize(p, BackendKind.Hip, 0)[0]
    assert live.queue(8, 60_000)
    live.pause()
    live.synchronize()
    assert live.poll()[0]
    assert 0 <= live.step() <= 8 and not live.busy()
    before = live.step()
    assert live.queue(3, 60_000)
    live.synchronize()
    assert live.step() <= before + 3
    ok, err = live.advance(2)
    assert ok, err
    live.close()


def test_uninitialized_facade():
    live = AsyncSimulation()
    assert not live.initialized() and not live.completed() and not live.busy()
    assert not live.queue(4, 100)
    ok, err = live.advance(1)
    assert not ok and err == "The 2D session is not initialized."
    assert live.backgroundTiming()["batches"] == 0
    live.close()
