"""C-ABI boundary and host-side session construction (CPU only: no compute calls).

* libswl.so loads and exports every function declared in include/swl.h;
* parameter blocks have the same layout in the header, the ctypes mirror and the oracle;
* the session planner (FP64 geometry, materials, absorbers, sources, dt plan, receivers,
  index maps) is bit-identical to the oracle on the same inputs.
"""
import ctypes as C
import math
import pathlib
import re

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, load
from seiswavelab_curvilinear_amd import BackendKind, Parameters, Precision, Simulation, native
from seiswavelab_curvilinear_amd.params import CParams2D, CParams3D
from test_oracle_2d import REF_CASES, precision_case, ref_case

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "swl.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(swl_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_parameter_block_layouts():
    lib = native.lib()
    assert lib.swl_params2d_size() == C.sizeof(CParams2D)
    assert lib.swl_params3d_size() == C.sizeof(CParams3D)
    orc = load("orc")
    orc.orc2d_params_size.restype = C.c_size_t
    orc.orc3d_params_size.restype = C.c_size_t
    assert orc.orc2d_params_size() == C.sizeof(CParams2D)
    assert orc.orc3d_params_size() == C.sizeof(CParams3D)


def test_invalid_parameters_rejected_with_reference_messages():
    s = Simulation()
    ok, err = s.initialize(Parameters(nx=7), BackendKind.Cpu)
    assert not ok and "8..16384" in err
    ok, err = s.initialize(Parameters(receiverCount=2000), BackendKind.Cpu)
    assert not ok and "1024" in err
    ok, err = s.initialize(Parameters(nz=30), BackendKind.Cpu)
    assert not ok and "two cells" in err
    assert not s.initialized()
    ok, _ = s.initialize(Parameters(nx=40, nz=40, sourceZ=5, receiverCount=4), BackendKind.Cpu)
    assert ok
    ok, err = s.initialize(Parameters(cfl=2), BackendKind.Cpu)
    assert not ok and s.initialized() and s.layout()["nx"] == 46  # transactional: old session kept
    ok, err = s.advance(1)
    assert not ok and "without a device" in err


PLAN_CASES = sorted(REF_CASES) + ["elastic_flat", "elastic_curved_moment", "gsls3_curved_pml", "dry_loess_gsls3_pml",
                                   "per_node_gsls2_pml"]


def plan_case(name):
    if name == "per_node_gsls2_pml":
        # per-node velocities and density (seeded 5 % perturbation) on the layered GSLS-2 model: every node gets
        # its own relaxed and unrelaxed moduli from its layer's fit
        p = ref_case("gsls2_layered_pml")
        rng = np.random.default_rng(11)
        depth = np.arange(p.nz)[:, None] * p.spacing * np.ones((1, p.nx))
        bounds = np.cumsum([l.thickness for l in p.layers])
        idx = np.minimum(np.searchsorted(bounds, depth, side="right"), len(p.layers) - 1)
        base = {k: np.array([getattr(l, k) for l in p.layers])[idx] for k in ("vp", "vs", "density")}
        scale = 1 + 0.05 * rng.standard_normal(depth.shape)  # one field for both speeds keeps Vp/Vs per layer
        p.nodeVp, p.nodeVs = (base["vp"] * scale).ravel(), (base["vs"] * scale).ravel()
        p.nodeDensity = (base["density"] * (1 + 0.02 * rng.standard_normal(depth.shape))).ravel()
        return p
    return ref_case(name) if name in REF_CASES else precision_case(name)


@pytest.mark.parametrize("name", PLAN_CASES)
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_planned_session_bitwise_equals_oracle(name, prec):
    p = plan_case(name)
    p.precision = prec
    s = Simulation()
    ok, err = s.initialize(p, BackendKind.Cpu)
    assert ok, err
    o = CpuSession2D("orc", p)
    assert s.dt() == o.dt and s.totalSteps() == o.total and s.recordDt() == o.record_dt
    L = s.layout()
    assert {k: v for k, v in L.items() if k != "records"} == {k: v for k, v in o.layout().items() if k != "records"}
    assert np.array_equal(s.probe_indices(), o.probe_indices())
    for w in range(22):
        a = o.model(w)
        if a is not None and (w < 16 or w - 16 < 3 * 2):
            if w >= 16 and L["mechanisms"] <= (w - 16) % 3:
                continue
            assert np.array_equal(s.model_array(w), a), w
    d1, d2 = s.diagnostics(), o.diagnostics()
    for k in ("waveDt", "dampingDt", "sourceDt", "cpmlDt", "relaxationDt", "ppw", "maxRelativeQError", "maxMetric",
              "minJacobian", "maxEdge", "flattenDepth", "sourceWeightSum", "sourceRmsRadiusM"):
        assert d1[k] == d2[k] or (math.isinf(d1[k]) and math.isinf(d2[k])), k


def test_pml_slot_map_equals_oracle():
    lib, orc = native.lib(), load("orc")
    for nx, nz, wx, wz in ((26, 20, 4, 3), (57, 41, 9, 12)):
        for q in range(nx * nz):
            i, k = q % nx, q // nx
            if 3 <= i < nx - 3 and 3 <= k < nz - 3:
                for axis in (0, 1):
                    assert lib.swl_pml_slot2d(nx, nz, wx, wz, q, axis) == orc.orc2d_pml_slot(nx, nz, wx, wz, q, axis)
