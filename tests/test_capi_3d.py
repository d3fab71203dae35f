"""3D host construction (libswl.so, csrc/model3d.cpp) checked bit for bit against the oracle on the CPU."""
import ctypes as C

import numpy as np
import pytest

from cpu_sessions import CpuSession3D
from seiswavelab_curvilinear_amd import BackendKind, Simulation3D
from test_oracle_3d import CASES, case


@pytest.mark.parametrize("name", CASES)
def test_planned_3d_session_bitwise_equals_oracle(name):
    p = case(name)
    o = CpuSession3D("orc", p)
    s = Simulation3D(BackendKind.Cpu)
    ok, err = s.initialize(p)
    assert ok, err
    assert s.dt() == o.dt and s.totalSteps() == o.total
    L, Lo = s.layout(), o.layout()
    for k in ("nx", "ny", "nz", "receivers", "records"):
        assert L[k] == Lo[k], k
    for which in range(12):
        a, b = s.model_array(which), o.model(which)
        assert np.array_equal(a, b), which
    ro, rs = o.receivers(), s.receiver_arrays()
    assert np.array_equal(ro["index"], rs["index"]) and np.array_equal(ro["position"], rs["position"])
    do, ds = o.diagnostics(), s.diagnostics()
    for k in ("minJacobian", "maxMetricNorm", "minCellLength", "maxCellEdge", "waveDtLimit", "dampingDtLimit",
              "sourceDtLimit"):
        assert do[k] == ds[k], k


def test_3d_invalid_parameters_rejected_with_reference_messages():
    s = Simulation3D(BackendKind.Cpu)
    p = case("analytic_sponge_force")
    p.maxSteps = 30000
    ok, err = s.initialize(p)
    assert not ok and err == "Maximum steps must be 0 (duration) or 1..20,000."
    p = case("dem_layers_moment")
    p.layers[0].thickness = 0
    ok, err = s.initialize(p)
    assert not ok and err.startswith("Invalid layer: positive finite Vp/Vs/density")
    p = case("dem_layers_moment")
    p.surfaceGrid.nx = 10
    p.surfaceGrid.elevations = p.surfaceGrid.elevations[:, :10]
    ok, err = s.initialize(p)
    assert not ok and err.startswith("Imported DEM does not cover the staged X/Y model span")
    assert not s.initialized()


def test_3d_unplanned_session_cannot_step():
    s = Simulation3D(BackendKind.Cpu)
    ok, err = s.initialize(case("flat_datum_layers"))
    assert ok
    ok, err = s.advance(1)
    assert not ok and "planned without a device" in err
