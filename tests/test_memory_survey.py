"""Session memory models and the automatic surface-receiver layout.

* 2D Detail::memoryBudget (src/two_d/MemoryBudget2D.hpp:12-38): the reference's own known answers
  (src/two_d/Validation2D.cpp:88-96, :111-127), and, where the reference build is present, equality with its
  memoryBudget over a sweep of parameter sets (CPU and CUDA backends).
* 3D estimateModelMemory3D (include/Seis3D/ModelPresets3D.hpp:38-49): the reference's known answer (a 300-node
  cube has 306^3 padded points and does not fit 8 GiB of VRAM, ConfigurationValidation.cpp:25-27) and equality
  with the reference build over a sweep.
* The HIP engine's device budget bounds what the engine really allocates (GPU): at least the engine's resident
  bytes and at most 35 % above them, for elastic, GSLS, sponge, CFS-PML, force and moment sessions, FP32 and FP64.
* setUniformSurfaceReceivers / synchronizeSurfaceReceivers (include/Seis2D/SurfaceSurvey2D.hpp:8-44): integer
  receiver layout, bit for bit against the reference build (receiver nodes, storage indices, the parameters the
  session reports), and the reference's refusals.
"""
import copy
import ctypes as C
import itertools

import numpy as np
import pytest

from cpu_sessions import CpuSession2D, have_ref, load
from seiswavelab_curvilinear_amd import (Absorber, BackendKind, GslsSettings, Precision, Simulation, Simulation3D,
                                         SourceKind, SourceMechanism, estimateModelMemory3D, memoryBudget)
from seiswavelab_curvilinear_amd.params import Parameters, SimulationParameters, to_c, to_c3d
from test_oracle_2d import ref_case
from test_oracle_3d import case as case3d

GIB = 1024.0 ** 3


def test_memory_budget_known_answers_2d():
    """Validation2D.cpp:88-96: an 8x8 CPU run with maxSteps 0 reserves records for 200001 samples."""
    p = Parameters(nx=8, nz=8, sourceZ=2, relief=0, receiverStart=2, receiverCount=2, receiverStride=1, maxSteps=0,
                   sampleIntervalMs=0)
    assert memoryBudget(p, BackendKind.Cpu)["hostBytes"] == 14.0 * 14 * 480 + 2.0 * 200001 * 2 * 8
    # Validation2D.cpp:111-127: 5001x2001 moment source, 3 s at 0.5 ms records, 250 receivers
    p = Parameters(nx=5001, nz=2001, spacing=2, source=SourceMechanism(SourceKind.IsotropicMoment), duration=3,
                   maxSteps=0, sampleIntervalMs=.5, receiverCount=250)
    nodes, records = 5007.0 * 2007.0, 250.0 * 6001 * 2 * 8
    cpu, gpu = memoryBudget(p, BackendKind.Cpu), memoryBudget(p, BackendKind.Cuda)
    assert cpu["hostBytes"] == nodes * 62 * 8 + records and cpu["deviceBytes"] == 0
    assert gpu["hostBytes"] == nodes * 20 * 8 + records and gpu["deviceBytes"] == nodes * 50 * 8
    p.source = SourceMechanism(SourceKind.DirectionalForce)
    force = memoryBudget(p, BackendKind.Cuda)
    assert gpu["hostBytes"] - force["hostBytes"] == nodes * 2 * 8
    assert gpu["deviceBytes"] - force["deviceBytes"] == nodes * 2 * 8


def budget_cases_2d():
    out = []
    for prec, absorber, external, kind, gsls, interval, steps in itertools.product(
            (Precision.Float64, Precision.Float32), (Absorber.Sponge, Absorber.CfsPml), (False, True),
            (SourceKind.DirectionalForce, SourceKind.IsotropicMoment), (0, 1, 3), (0.0, .25), (0, 700)):
        p = Parameters(precision=prec, nx=157, nz=83, absorber=absorber, externalAbsorption=external,
                       source=SourceMechanism(kind), sourceZ=9, receiverCount=37, sampleIntervalMs=interval,
                       maxSteps=steps, spongeX=13, spongeZ=11)
        if gsls:
            p.attenuation = GslsSettings(enabled=True, mechanisms=gsls)
        out.append(p)
    return out


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
def test_memory_budget_equals_reference_build_2d():
    lib = load("ref")
    lib.ref2d_memory_budget.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    for p in budget_cases_2d():
        c, keep = to_c(p)
        for backend in (BackendKind.Cpu, BackendKind.Cuda):
            want = np.zeros(2)
            lib.ref2d_memory_budget(C.byref(c), int(backend == BackendKind.Cuda), want.ctypes.data_as(C.c_void_p))
            got = memoryBudget(p, backend)
            assert (got["hostBytes"], got["deviceBytes"]) == (want[0], want[1]), (p, backend)


def test_memory_estimate_known_answer_3d():
    """ConfigurationValidation.cpp:25-27: the 300-node cube preset counts ghosts (306^3 points) and does not fit
    a GPU with 8 GiB free."""
    p = SimulationParameters(nx=300, ny=300, nz=300, sourceZ=50)
    m = estimateModelMemory3D(p)
    assert m["points"] == 306 ** 3 and m["withinBudget"]
    assert m["deviceMiB"] * 1024 * 1024 > .9 * 8 * GIB
    assert not estimateModelMemory3D(SimulationParameters(nx=301, ny=300, nz=300, sourceZ=50))["withinBudget"]


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
def test_memory_estimate_equals_reference_build_3d():
    lib = load("ref")
    lib.ref3d_memory_estimate.argtypes = [C.c_void_p, C.c_void_p]
    for name in ("analytic_sponge_force", "dem_layers_moment", "dem_gridindex_layers", "flat_datum_layers"):
        for steps in (0, 24, 777):
            p = case3d(name)
            p.maxSteps = steps
            c, keep = to_c3d(p)
            want = np.zeros(5)
            lib.ref3d_memory_estimate(C.byref(c), want.ctypes.data_as(C.c_void_p))
            m = estimateModelMemory3D(p)
            got = [m["points"], m["cpuMiB"], m["deviceMiB"], m["cudaHostMiB"], float(m["withinBudget"])]
            assert got == want.tolist(), (name, steps)


# ---- automatic surface receivers


def survey_cases():
    out = []
    for nx, sponge, external, absorbing, count, target in itertools.product(
            (40, 157, 500), (6, 20), (False, True), (True, False), (1, 2, 37, 96, 2000), (0, 1, 25, 1024)):
        p = Parameters(nx=nx, nz=60, spacing=2, relief=2, wavelength=90, sourceZ=6, spongeX=sponge, spongeZ=8,
                       externalAbsorption=external, absorbing=absorbing, receiverCount=count, receiverTargetCount=target,
                       autoSurfaceReceivers=True, receiverStart=3, receiverStride=7, receiverDepth=4, maxSteps=5)
        out.append(p)
    return out


def plan(p):
    s = Simulation()
    ok, err = s.initialize(copy.deepcopy(p), BackendKind.Cpu)
    return s, ok, err


@pytest.mark.skipif(not have_ref(), reason="reference build oracle/_ref not present")
def test_auto_surface_receivers_equal_reference_build():
    lib = load("ref")
    lib.ref2d_receiver_line.argtypes = [C.c_void_p, C.c_void_p]
    accepted = 0
    for p in survey_cases():
        try:
            r = CpuSession2D("ref", p)
        except ValueError as e:
            s, ok, err = plan(p)
            assert not ok and err == str(e), (p, err, e)
            continue
        s, ok, err = plan(p)
        assert ok, err
        line = np.zeros(5, np.int32)
        lib.ref2d_receiver_line(C.c_void_p(r.h), line.ctypes.data_as(C.c_void_p))
        q = s.parameters()
        assert [q.receiverStart, q.receiverCount, q.receiverStride, q.receiverDepth, q.receiverTargetCount] == \
            line.tolist()
        assert np.array_equal(s.probe_indices(), r.probe_indices())
        assert np.array_equal(s.receiver_arrays()["xz"], r.receivers()["xz"])
        accepted += 1
    assert accepted > 100


def test_auto_surface_receivers_rules():
    """SurfaceSurvey2D.hpp:8-44 on cases whose answer follows from the rules: an internal absorber keeps the
    line out of the side strips, the quantized count never replaces the target, and targets above 1024 are
    refused with the reference's message."""
    p = Parameters(nx=101, nz=60, sourceZ=6, spongeX=10, absorbing=True, autoSurfaceReceivers=True,
                   receiverTargetCount=11)
    s, ok, err = plan(p)
    assert ok, err
    q = s.parameters()
    # the span between the 10-node strips is 80 cells; stride round(80 / (11 - 1)) = 8 fits 11 receivers from node 10
    assert (q.receiverStart, q.receiverCount, q.receiverStride, q.receiverDepth, q.receiverTargetCount) == (10, 11, 8, 0, 11)
    p.receiverTargetCount = 1025
    s, ok, err = plan(p)
    assert not ok and err == "Desired surface receiver count must be 0..1024."


# ---- the HIP engine's device budget against its real allocations (GPU)


def resident_cases():
    a = ref_case("sine_terrain_small")                   # PML, force, layered
    b = ref_case("moment_external_pml")                  # external PML, moment
    c = ref_case("sponge_tilted_force")                  # sponge
    d = ref_case("gsls2_layered_pml")                    # GSLS-2
    e = ref_case("moment_external_pml")
    e.attenuation = GslsSettings(enabled=True, mechanisms=3)  # GSLS-3, moment
    return [a, b, c, d, e]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [Precision.Float64, Precision.Float32])
def test_hip_budget_bounds_engine_allocation_2d(prec):
    for n, p in enumerate(resident_cases()):
        p.precision = prec
        s = Simulation()
        ok, err = s.initialize(p, BackendKind.Hip, 0)
        assert ok, err
        assert s.advance(s.totalSteps())[0]  # an advance reserves result rows for up to 512 of its steps
        resident = s.diagnostics()["residentDeviceBytes"]
        budget = memoryBudget(p, BackendKind.Hip)["deviceBytes"]
        assert resident <= budget <= 1.35 * resident, (n, resident, budget)
        s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_hip_budget_bounds_engine_allocation_3d(precision):
    for name in ("analytic_sponge_force", "dem_layers_moment"):
        p = case3d(name)
        s = Simulation3D(BackendKind.Hip)
        ok, err = s.initialize(p, 0, precision)
        assert ok, err
        assert s.advance(p.maxSteps)[0]
        s.field(0)  # the download buffer is part of what a session holds once read
        resident = s.diagnostics()["residentDeviceBytes"]
        budget = estimateModelMemory3D(p, precision)["hipDeviceBytes"]
        assert resident <= budget, (name, resident, budget)
        s.close()
