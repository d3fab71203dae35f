"""Background stepping of 3D sessions (csrc/async3d.cpp, AsyncSimulation3D) against direct stepping (GPU).

The reference's façade (src/three_d/AsyncSimulation3D.cpp) publishes, after each background batch, the step,
the diagnostics and the receiver samples recorded in the batch; a session stepped that way must hold the same
fields and traces, bit for bit, as one advanced directly to the same step.  Also: display slices through the
façade (pending while a batch runs, the current step's afterwards), a poll loop to completion, resetFields, and
the uninitialized façade."""
import time

import numpy as np
import pytest

from seiswavelab_curvilinear_amd import AsyncSimulation3D, BackendKind, Simulation3D
from test_oracle_3d import case

pytestmark = pytest.mark.gpu


def same_state(live, ref):
    for f in range(9):
        assert np.array_equal(live.field(f), ref.field(f)), f
    a, b = live.receiver_arrays(), ref.receiver_arrays()
    for k in ("vx", "vy", "vz"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [0, 1])
def test_background_batches_equal_direct_3d(precision):
    p = case("dem_layers_moment")
    p.maxSteps = 60
    print(f"[async3d] batches precision {precision}", flush=True)
    live, ref = AsyncSimulation3D(), Simulation3D(BackendKind.Hip)
    assert live.initialize(p, 0, precision)[0] and ref.initialize(p, 0, precision)[0]
    ok, err, planes = live.readDisplayFrame(5, 6, 3, 0)
    assert ok and planes[0].shape == (p.nz, p.nx)
    for batch in range(4):
        assert live.queue(8, 1000)
        if batch < 2:
            deadline = time.time() + 30
            while live.busy() and time.time() < deadline:
                assert live.poll()[0]
                time.sleep(0.001)
        live.synchronize()
        ok, err = live.poll()
        assert ok, err
    n = live.currentStep()
    assert 16 <= n <= 32
    assert ref.advance(n)[0]
    same_state(live, ref)
    ok, err, planes = live.readDisplayFrame(5, 6, 3, 0)
    ok_r, err_r, planes_r = ref.readDisplayFrame(5, 6, 3, 0)
    assert ok and ok_r and all(np.array_equal(x, y) for x, y in zip(planes, planes_r))
    d, dr = live.diagnostics(), ref.diagnostics()
    assert d["elasticEnergy"] == dr["elasticEnergy"] and d["peakEnergy"] == dr["peakEnergy"]
    t = live.backgroundTiming()
    assert t["batches"] == 4 and t["steps"] == n
    assert live.resetFields()[0] and live.currentStep() == 0
    live.close()


def test_poll_loop_to_completion_equals_direct_3d():
    p = case("analytic_sponge_force")
    print("[async3d] poll loop", flush=True)
    live, ref = AsyncSimulation3D(), Simulation3D(BackendKind.Hip)
    assert live.initialize(p, 0, 0)[0] and ref.initialize(p, 0, 0)[0]
    deadline = time.time() + 60
    while not live.completed() and time.time() < deadline:
        if not live.busy():
            live.queue(8, 50)
        ok, err = live.poll()
        assert ok, err
        time.sleep(0.001)
    live.synchronize()
    assert live.poll()[0] and live.completed() and live.currentStep() == p.maxSteps
    assert ref.advance(p.maxSteps)[0]
    same_state(live, ref)
    assert not live.queue(8, 50)  # a completed session queues nothing
    live.close()


def test_uninitialized_facade_3d():
    live = AsyncSimulation3D()
    assert not live.initialized() and not live.queue(8, 10)
    ok, err = live.advance(1)
    assert not ok and err
    live.close()
