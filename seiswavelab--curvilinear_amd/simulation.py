"""Host mirror of ``Seis2D::Curvilinear::Simulation`` (include/Seis2D/CurvilinearSolver2D.hpp:93-123).

Same method names and error behaviour as the reference: ``initialize``, ``advance`` and
``reset`` return ``(ok, error)``; a rejected step latches a fault until the session is
rebuilt; initialization is transactional (a failed call keeps the previous session).
All stepping runs in the HIP library; this class only marshals arguments.
"""
from __future__ import annotations

import copy
import ctypes as C
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import native
from .params import BackendKind, CParams2D, Parameters, to_c

DIAGNOSTIC_NAMES = ["energy", "peakEnergy", "maxVelocity", "tractionRelative", "finite", "waveDt", "dampingDt",
                    "sourceDt", "cpmlDt", "relaxationDt", "ppw", "maxRelativeQError", "maxMetric", "minJacobian",
                    "maxEdge", "flattenDepth", "sourceWeightSum", "sourceRmsRadiusM", "initializationMs",
                    "totalStepMs", "lastStepMs", "residentDeviceBytes"]
LAYOUT_NAMES = ["nx", "nz", "offsetX", "widthX", "widthZ", "fields", "cpml", "receivers", "records", "mechanisms"]
MODEL_ARRAYS = {"a": 0, "b": 1, "jacobian": 2, "lambda": 3, "mu": 4, "density": 5, "source": 6, "damping": 7,
                "gradientX": 8, "gradientZ": 9, "normals": 10, "coordinates": 11, "vp": 12, "vs": 13, "cpmlX": 14,
                "cpmlZ": 15}


@dataclass
class Receiver:
    x: int
    z: int
    position: Tuple[float, float]
    vx: np.ndarray
    vz: np.ndarray


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Simulation:
    """2D P-SV / Qp-Qs session on the HIP backend (or planned without a device)."""

    def __init__(self):
        self._h = None
        self._params = None

    # -- lifecycle -------------------------------------------------------------------------
    def initialize(self, params: Parameters, backend: BackendKind = BackendKind.Hip, device: int = 0):
        c, keep = to_c(params)
        out = C.c_void_p()
        lib = native.lib()
        if backend == BackendKind.Hip:
            rc = lib.swl_sim2d_create(C.byref(c), int(backend), device, C.byref(out))
        elif backend == BackendKind.Cpu:
            rc = lib.swl_sim2d_plan(C.byref(c), C.byref(out))  # device-free plan (no stepping)
        else:
            return False, "Unavailable backend: this build provides HIP (and device-free planning)."
        if rc != 0:
            return False, native.last_error()
        self.close()
        self._h = out
        self._params = params
        if params.autoSurfaceReceivers:  # the session's parameters carry the layout it chose
            e = CParams2D()
            lib.swl_sim2d_parameters(out, C.byref(e))
            q = self._params = copy.deepcopy(params)
            q.receiverStart, q.receiverCount, q.receiverStride = e.receiver_start, e.receiver_count, e.receiver_stride
            q.receiverDepth, q.receiverTargetCount = e.receiver_depth, e.receiver_target_count
        return True, ""

    def close(self):
        if self._h:
            native.lib().swl_sim2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, count: int):
        self._need()
        rc = native.lib().swl_sim2d_advance(self._h, int(count))
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def reset(self):
        self._need()
        rc = native.lib().swl_sim2d_reset(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def _need(self):
        if not self._h:
            raise native.NativeError("Initialize the 2D model first.")

    # -- accessors -------------------------------------------------------------------------
    def initialized(self) -> bool:
        return bool(self._h)

    def step(self) -> int:
        return native.lib().swl_sim2d_step(self._h)

    def totalSteps(self) -> int:
        return native.lib().swl_sim2d_total_steps(self._h)

    def completed(self) -> bool:
        return self.step() >= self.totalSteps()

    def dt(self) -> float:
        return native.lib().swl_sim2d_dt(self._h)

    def recordDt(self) -> float:
        return native.lib().swl_sim2d_record_dt(self._h)

    def parameters(self) -> Parameters:
        return self._params

    def layout(self) -> dict:
        self._need()
        o = np.zeros(10, np.int64)
        native.lib().swl_sim2d_layout(self._h, _dp(o))
        return dict(zip(LAYOUT_NAMES, (int(v) for v in o)))

    def diagnostics(self) -> dict:
        self._need()
        o = np.zeros(len(DIAGNOSTIC_NAMES))
        native.lib().swl_sim2d_diagnostics(self._h, _dp(o))
        d = dict(zip(DIAGNOSTIC_NAMES, o.tolist()))
        d["finite"] = bool(d["finite"])
        return d

    def receiver_arrays(self):
        self._need()
        L = self.layout()
        nr, n = L["receivers"], L["records"]
        vx, vz = np.zeros((nr, n)), np.zeros((nr, n))
        pos, xz = np.zeros((nr, 2)), np.zeros((nr, 2), np.int32)
        native.lib().swl_sim2d_receivers(self._h, _dp(vx), _dp(vz), _dp(pos), _dp(xz))
        return {"vx": vx, "vz": vz, "position": pos, "xz": xz}

    def receivers(self) -> List[Receiver]:
        a = self.receiver_arrays()
        return [Receiver(int(a["xz"][r, 0]), int(a["xz"][r, 1]), tuple(a["position"][r]), a["vx"][r], a["vz"][r])
                for r in range(a["vx"].shape[0])]

    def probe_indices(self) -> np.ndarray:
        self._need()
        o = np.zeros(self.layout()["receivers"], np.int64)
        native.lib().swl_sim2d_probe_indices(self._h, _dp(o))
        return o

    def model_array(self, name_or_index) -> np.ndarray:
        self._need()
        which = MODEL_ARRAYS.get(name_or_index, name_or_index)
        L = self.layout()
        n = L["nx"] * L["nz"]
        size = {10: 2 * n, 11: 2 * n, 14: 3 * (L["nx"] - 6), 15: 3 * (L["nz"] - 6)}.get(which, n)
        o = np.zeros(size)
        native.call("swl_sim2d_model", self._h, int(which), _dp(o))
        return o

    def snapshot_field(self, f: int) -> np.ndarray:
        self._need()
        L = self.layout()
        o = np.zeros(L["nx"] * L["nz"])
        native.call("swl_sim2d_snapshot", self._h, int(f), _dp(o))
        return o

    def snapshot(self, includeMemories: bool = False) -> dict:
        """Vx,Vz,Sxx,Szz,Sxz in padded storage (plus Syy, GSLS and CFS-PML memories on request)."""
        L = self.layout()
        out = {"fields": [self.snapshot_field(f) for f in range(5)]}
        if L["mechanisms"]:
            out["syy"] = self.snapshot_field(5)
            if includeMemories:
                out["memory"] = [self.snapshot_field(f) for f in range(6, L["fields"])]
        if includeMemories and L["cpml"]:
            out["boundaryMemory"] = self.pml()
        return out

    def pml(self) -> np.ndarray:
        self._need()
        L = self.layout()
        o = np.zeros(max(L["cpml"], 1))
        native.call("swl_sim2d_pml", self._h, _dp(o))
        return o[: L["cpml"]]

    def profile(self, enable: bool = True):
        """Per-launch HIP-event timing of the stage and finalize kernels (resets counters)."""
        self._need()
        native.call("swl_sim2d_profile", self._h, int(enable))

    def kernel_timing(self) -> dict:
        self._need()
        o = np.zeros(14)
        native.call("swl_sim2d_kernel_timing", self._h, _dp(o))
        return {"stage_ms": o[0], "stage_launches": int(o[1]), "finalize_ms": o[2], "finalize_launches": int(o[3]),
                "interior_tiles": int(o[4]), "vgprs": int(o[5]), "nodes_per_launch": int(o[6]),
                "border_tiles": int(o[7]), "metric_on_the_fly": bool(o[8]), "material_by_row": bool(o[9]),
                "pml_lanes": int(o[10]), "pml_blocks": int(o[11]), "latency_kernel": bool(o[12]),
                "lds_tiles": bool(o[13])}

    def fault(self) -> str:
        """The latched fault message ("" if none)."""
        self._need()
        return native.lib().swl_sim2d_fault(self._h).decode()

    def override_peak_energy(self, peak: float):
        """Replaces the accepted peak energy the late-growth fault rule compares against
        (CurvilinearSolver2D.cpp:576-578); the next advance, alone or in a batch, uses it."""
        self._need()
        native.call("swl_sim2d_override_peak_energy", self._h, float(peak))

    def set_option(self, name: str, value: int):
        """Engine tuning switch (swl_engine2d_set_option): "nontemporal_stores", "stored_metrics", "lds_tiles",
        "material_rows", "pml_vector", "latency_kernel", "end_in_stage", "split_border", "spread_border"."""
        self._need()
        native.call("swl_sim2d_set_option", self._h, name.encode(), int(value))

    def readFrame(self, component: int, includeAbsorber: bool = False):
        self._need()
        L = self.layout()
        p = self._params
        nx = L["nx"] - 6 if includeAbsorber else p.nx
        nz = L["nz"] - 6 if includeAbsorber else p.nz
        out = np.zeros(nx * nz, np.float32)
        rc = native.lib().swl_sim2d_read_frame(self._h, int(component), _dp(out), int(includeAbsorber))
        if rc != 0:
            return False, native.last_error(), None
        return True, "", out.reshape(nz, nx)

    def readProbes(self, nodes):
        self._need()
        xz = np.ascontiguousarray(np.asarray(nodes, np.int32).reshape(-1, 2))
        out = np.zeros((xz.shape[0], 5))
        rc = native.lib().swl_sim2d_read_probes(self._h, _dp(xz), xz.shape[0], _dp(out))
        if rc != 0:
            return False, native.last_error(), None
        return True, "", out


class _SessionView(Simulation):
    """A Simulation over a session owned by someone else (an AsyncSimulation): never destroys it."""

    def __init__(self, handle, params):
        super().__init__()
        self._h, self._params = handle, params

    def close(self):
        self._h = None

    def __del__(self):
        self._h = None


class AsyncSimulation:
    """Background stepping of a 2D session: the reference's AsyncSimulation (AsyncSimulation2D.hpp:5-22) over
    swl_async2d_*.  ``queue(steps, budgetMs)`` runs at most 8 steps on a worker thread; ``poll`` / ``synchronize``
    fold the published batch (step, diagnostics, receiver samples, a requested display frame) into the state
    the accessors return; ``pause`` stops a batch between steps."""

    def __init__(self):
        out = C.c_void_p()
        native.call("swl_async2d_create", C.byref(out))
        self._h = out
        self._params = None

    def close(self):
        if self._h:
            native.lib().swl_async2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _state(self):
        o = np.zeros(6, np.int64)
        native.lib().swl_async2d_state(self._h, _dp(o))
        return o

    def initialize(self, params: Parameters, backend: BackendKind = BackendKind.Hip, device: int = 0):
        c, keep = to_c(params)
        rc = native.lib().swl_async2d_initialize(self._h, C.byref(c), int(backend), int(device))
        if rc != 0:
            return False, native.last_error()
        self._params = params
        self._layout = self.session().layout()  # fixed for the session's lifetime
        return True, ""

    def queue(self, steps: int, budgetMs: float) -> bool:
        return native.lib().swl_async2d_queue(self._h, int(steps), float(budgetMs)) == 1

    def poll(self):
        rc = native.lib().swl_async2d_poll(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def pause(self):
        native.lib().swl_async2d_pause(self._h)

    def synchronize(self):
        if native.lib().swl_async2d_synchronize(self._h) != 0:
            raise RuntimeError(native.last_error())

    def busy(self) -> bool:
        return native.lib().swl_async2d_busy(self._h) == 1

    def backgroundTiming(self) -> dict:
        o = np.zeros(4)
        native.lib().swl_async2d_timing(self._h, _dp(o))
        return {"computeMs": o[0], "transferMs": o[1], "batches": int(o[2]), "steps": int(o[3])}

    def advance(self, count: int):
        rc = native.lib().swl_async2d_advance(self._h, int(count))
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def reset(self):
        rc = native.lib().swl_async2d_reset(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def initialized(self) -> bool:
        return bool(self._state()[0])

    def completed(self) -> bool:
        return bool(self._state()[1])

    def step(self) -> int:
        return int(self._state()[2])

    def totalSteps(self) -> int:
        return int(self._state()[3])

    def dt(self) -> float:
        o = np.zeros(2)
        native.lib().swl_async2d_times(self._h, _dp(o))
        return float(o[0])

    def recordDt(self) -> float:
        o = np.zeros(2)
        native.lib().swl_async2d_times(self._h, _dp(o))
        return float(o[1])

    def parameters(self) -> Parameters:
        return self._params

    def diagnostics(self) -> dict:
        o = np.zeros(len(DIAGNOSTIC_NAMES))
        native.lib().swl_async2d_diagnostics(self._h, _dp(o))
        d = dict(zip(DIAGNOSTIC_NAMES, o.tolist()))
        d["finite"] = bool(d["finite"])
        return d

    def receiver_arrays(self):
        s = self._state()
        nr, n = int(s[4]), int(s[5])
        vx, vz = np.zeros((nr, n)), np.zeros((nr, n))
        native.lib().swl_async2d_receivers(self._h, _dp(vx), _dp(vz))
        return {"vx": vx, "vz": vz}

    def readFrame(self, component: int, includeAbsorber: bool = False):
        """(ok, error, frame).  While a batch runs and the current step's frame is not yet published it returns
        (False, "", None) and remembers the request for the next batch, as the reference does."""
        L, p = self._layout, self._params
        nx = L["nx"] - 6 if includeAbsorber else p.nx
        nz = L["nz"] - 6 if includeAbsorber else p.nz
        out = np.zeros(nx * nz, np.float32)
        ready = C.c_int()
        rc = native.lib().swl_async2d_read_frame(self._h, int(component), _dp(out), int(includeAbsorber),
                                                  C.byref(ready))
        if rc != 0:
            return False, native.last_error(), None
        return (ready.value == 1), "", (out.reshape(nz, nx) if ready.value else None)

    def session(self) -> Simulation:
        """The session itself after the running batch has finished (operator const Simulation&)."""
        return _SessionView(native.lib().swl_async2d_session(self._h), self._params)

    def snapshot(self, includeMemories: bool = False) -> dict:
        return self.session().snapshot(includeMemories)


def memoryBudget(params: Parameters, backend: BackendKind = BackendKind.Hip) -> dict:
    """Detail::memoryBudget (src/two_d/MemoryBudget2D.hpp:12-38): the host and device bytes a 2D session adds, for
    the reference's CPU and CUDA backends as it computes them, and for this library's HIP engine (device bytes an
    upper bound of the engine's allocations)."""
    c, keep = to_c(params)
    o = np.zeros(2)
    native.call("swl_memory_budget2d", C.byref(c), int(backend), _dp(o))
    return {"hostBytes": float(o[0]), "deviceBytes": float(o[1])}


def estimateModelMemory3D(params, precision: int = 0) -> dict:
    """estimateModelMemory3D (include/Seis3D/ModelPresets3D.hpp:38-49), plus the HIP engine's device bytes at
    `precision` (0 FP64, 1 FP32)."""
    from .params import to_c3d
    c, keep = to_c3d(params)
    o = np.zeros(6)
    native.call("swl_memory_estimate3d", C.byref(c), int(precision), _dp(o))
    return {"points": int(o[0]), "cpuMiB": float(o[1]), "deviceMiB": float(o[2]), "cudaHostMiB": float(o[3]),
            "withinBudget": bool(o[4]), "hipDeviceBytes": float(o[5])}


def writeReceivers(sim: Simulation, path: str):
    """Receiver CSV with the reference's provenance header (CurvilinearSolver2D.cpp:811-854)."""
    rc = native.lib().swl_sim2d_write_receivers(sim._h, str(path).encode())
    return (rc == 0), ("" if rc == 0 else native.last_error())


# ---------------------------------------------------------------------------------------------
# 3D: mirror of ISimulationBackend3D / CpuSimulation3D (include/Seis3D/CurvilinearSolver.hpp:165-218)

DIAGNOSTIC_NAMES_3D = ["elasticEnergy", "peakEnergy", "sourcePeakEnergy", "maxAbsVelocity", "maxTractionRateRelative",
                       "allFinite", "minJacobian", "maxMetricNorm", "minCellLength", "maxCellEdge", "waveDtLimit",
                       "dampingDtLimit", "sourceDtLimit", "peakVz", "peakVx", "peakVy", "peakV", "peakPressure",
                       "initializationMs", "totalStepMs", "residentDeviceBytes"]
MODEL_ARRAYS_3D = {"metrics": 0, "jacobian": 1, "normals": 2, "coordinates": 3, "lambda": 4, "mu": 5, "density": 6,
                   "damping": 7, "source": 8, "gradientX": 9, "gradientY": 10, "gradientZ": 11}


@dataclass
class ReceiverTrace3D:
    position: Tuple[float, float, float]
    gridIndex: int
    samples: np.ndarray  # Vz (the reference keeps raw Vz in `samples`)
    samplesVx: np.ndarray
    samplesVy: np.ndarray


class Simulation3D:
    """3D nine-field elastic session on the HIP backend (FP64, or FP32 state with FP64 closure)."""

    def __init__(self, backend: BackendKind = BackendKind.Hip):
        self._h = None
        self._params = None
        self._backend = backend

    def initialize(self, params, device: int = 0, precision: int = 0):
        from .params import to_c3d
        c, keep = to_c3d(params)
        out = C.c_void_p()
        lib = native.lib()
        if self._backend == BackendKind.Hip:
            rc = lib.swl_sim3d_create(C.byref(c), int(self._backend), device, int(precision), C.byref(out))
        elif self._backend == BackendKind.Cpu:
            rc = lib.swl_sim3d_plan(C.byref(c), C.byref(out))  # device-free plan (no stepping)
        else:
            return False, "Unavailable backend: this build provides HIP (and device-free planning)."
        if rc != 0:
            return False, native.last_error()
        self.close()
        self._h, self._params = out, params
        return True, ""

    def close(self):
        if self._h:
            native.lib().swl_sim3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _need(self):
        if not self._h:
            raise native.NativeError("Initialize a model before stepping.")

    def advance(self, count: int):
        self._need()
        rc = native.lib().swl_sim3d_advance(self._h, int(count))
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def resetFields(self):
        self._need()
        rc = native.lib().swl_sim3d_reset(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def initialized(self) -> bool:
        return bool(self._h)

    def currentStep(self) -> int:
        return native.lib().swl_sim3d_step(self._h) if self._h else 0

    def totalSteps(self) -> int:
        return native.lib().swl_sim3d_total_steps(self._h) if self._h else 0

    def completed(self) -> bool:
        return bool(self._h) and self.currentStep() >= self.totalSteps()

    def dt(self) -> float:
        return native.lib().swl_sim3d_dt(self._h) if self._h else 0.0

    def parameters(self):
        return self._params

    def layout(self) -> dict:
        self._need()
        o = np.zeros(5, np.int64)
        native.lib().swl_sim3d_layout(self._h, _dp(o))
        return dict(zip(["nx", "ny", "nz", "receivers", "records"], (int(v) for v in o)))

    def diagnostics(self) -> dict:
        self._need()
        o = np.zeros(len(DIAGNOSTIC_NAMES_3D))
        native.lib().swl_sim3d_diagnostics(self._h, _dp(o))
        d = dict(zip(DIAGNOSTIC_NAMES_3D, o.tolist()))
        d["allFinite"] = bool(d["allFinite"])
        return d

    def receiver_arrays(self):
        self._need()
        L = self.layout()
        nr, n = L["receivers"], L["records"]
        vx, vy, vz = np.zeros((nr, n)), np.zeros((nr, n)), np.zeros((nr, n))
        pos, idx = np.zeros((nr, 3)), np.zeros(nr, np.int64)
        native.lib().swl_sim3d_receivers(self._h, *(_dp(a) for a in (vx, vy, vz, pos, idx)))
        return {"vx": vx, "vy": vy, "vz": vz, "position": pos, "index": idx}

    def receivers(self) -> List[ReceiverTrace3D]:
        a = self.receiver_arrays()
        return [ReceiverTrace3D(tuple(a["position"][r]), int(a["index"][r]), a["vz"][r], a["vx"][r], a["vy"][r])
                for r in range(a["vz"].shape[0])]

    def field(self, component: int) -> np.ndarray:
        self._need()
        L = self.layout()
        o = np.zeros(L["nx"] * L["ny"] * L["nz"])
        native.call("swl_sim3d_snapshot", self._h, int(component), _dp(o))
        return o

    def model_array(self, name_or_index) -> np.ndarray:
        self._need()
        which = MODEL_ARRAYS_3D.get(name_or_index, name_or_index)
        L = self.layout()
        n = L["nx"] * L["ny"] * L["nz"]
        o = np.zeros({0: 9 * n, 2: 3 * n, 3: 3 * n}.get(which, n))
        native.call("swl_sim3d_model", self._h, int(which), _dp(o))
        return o

    def readDisplayFrame(self, x: int, y: int, z: int, component: int):
        self._need()
        p = self._params
        out = np.zeros(p.nx * p.nz + p.ny * p.nz + p.nx * p.ny, np.float32)
        rc = native.lib().swl_sim3d_read_display_frame(self._h, int(x), int(y), int(z), int(component), _dp(out))
        if rc != 0:
            return False, native.last_error(), None
        a, b = p.nx * p.nz, p.ny * p.nz
        return True, "", [out[:a].reshape(p.nz, p.nx), out[a:a + b].reshape(p.nz, p.ny), out[a + b:].reshape(p.ny, p.nx)]

    def profile(self, enable: bool = True):
        self._need()
        native.call("swl_sim3d_profile", self._h, int(enable))

    def set_option(self, name: str, value: int):
        """Engine tuning switches: "vector_terrain", "vector_flat", "overlap_bands", "split_surface", "lds_tiles",
        "brick_x", "brick_y", "brick_z" (results identical; see swl_engine3d_set_option)."""
        self._need()
        native.call("swl_sim3d_set_option", self._h, name.encode(), int(value))

    def override_source_peak(self, peak: float):
        """Replaces the source-window peak energy of the late-growth fault rule (CurvilinearSolver.cpp:477-488)."""
        self._need()
        native.call("swl_sim3d_override_source_peak", self._h, float(peak))

    def kernel_timing(self) -> dict:
        self._need()
        o = np.zeros(7)
        native.call("swl_sim3d_kernel_timing", self._h, _dp(o))
        return {"stage_ms": o[0], "stage_launches": int(o[1]), "nodes_per_launch": int(o[2]), "vgprs": int(o[3]),
                "metric_planes": int(o[4]), "material_by_depth": bool(o[5]), "z_row_metrics": bool(o[6])}


class AsyncSimulation3D:
    """Background stepping of a 3D session: the reference's AsyncSimulation3D (AsyncSimulation3D.hpp:5-33) over
    swl_async3d_*.  ``queue(maxSteps, budgetMs)`` runs at most 8 steps on a worker thread; ``poll`` /
    ``synchronize`` fold the published batch (step, diagnostics, Vx / Vy / Vz samples, requested display slices)
    into the state the accessors return; ``pause`` stops a batch between steps."""

    def __init__(self, backend: BackendKind = BackendKind.Hip):
        out = C.c_void_p()
        native.call("swl_async3d_create", C.byref(out))
        self._h = out
        self._params = None
        self._backend = backend

    def close(self):
        if self._h:
            native.lib().swl_async3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _state(self):
        o = np.zeros(6, np.int64)
        native.lib().swl_async3d_state(self._h, _dp(o))
        return o

    def initialize(self, params, device: int = 0, precision: int = 0):
        from .params import to_c3d
        c, keep = to_c3d(params)
        rc = native.lib().swl_async3d_initialize(self._h, C.byref(c), int(self._backend), int(device), int(precision))
        if rc != 0:
            return False, native.last_error()
        self._params = params
        return True, ""

    def queue(self, maxSteps: int, budgetMs: float) -> bool:
        return native.lib().swl_async3d_queue(self._h, int(maxSteps), float(budgetMs)) == 1

    def poll(self):
        rc = native.lib().swl_async3d_poll(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def pause(self):
        native.lib().swl_async3d_pause(self._h)

    def synchronize(self):
        if native.lib().swl_async3d_synchronize(self._h) != 0:
            raise RuntimeError(native.last_error())

    def busy(self) -> bool:
        return native.lib().swl_async3d_busy(self._h) == 1

    def backgroundTiming(self) -> dict:
        o = np.zeros(4)
        native.lib().swl_async3d_timing(self._h, _dp(o))
        return {"computeMs": o[0], "transferMs": o[1], "batches": int(o[2]), "steps": int(o[3])}

    def advance(self, count: int):
        rc = native.lib().swl_async3d_advance(self._h, int(count))
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def resetFields(self):
        rc = native.lib().swl_async3d_reset(self._h)
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def initialized(self) -> bool:
        return bool(self._state()[0])

    def completed(self) -> bool:
        return bool(self._state()[1])

    def currentStep(self) -> int:
        return int(self._state()[2])

    def totalSteps(self) -> int:
        return int(self._state()[3])

    def dt(self) -> float:
        return float(native.lib().swl_async3d_dt(self._h))

    def parameters(self):
        return self._params

    def diagnostics(self) -> dict:
        o = np.zeros(len(DIAGNOSTIC_NAMES_3D))
        native.lib().swl_async3d_diagnostics(self._h, _dp(o))
        d = dict(zip(DIAGNOSTIC_NAMES_3D, o.tolist()))
        d["allFinite"] = bool(d["allFinite"])
        return d

    def receiver_arrays(self):
        s = self._state()
        nr, n = int(s[4]), int(s[5])
        vx, vy, vz = np.zeros((nr, n)), np.zeros((nr, n)), np.zeros((nr, n))
        native.lib().swl_async3d_receivers(self._h, _dp(vx), _dp(vy), _dp(vz))
        return {"vx": vx, "vy": vy, "vz": vz}

    def readDisplayFrame(self, x: int, y: int, z: int, component: int):
        """(ok, error, planes).  While a batch runs and the current step's slices are not yet published it returns
        (False, "", None) and keeps the request for the next batch, as the reference does."""
        p = self._params
        out = np.zeros(p.nx * p.nz + p.ny * p.nz + p.nx * p.ny, np.float32)
        ready = C.c_int()
        rc = native.lib().swl_async3d_read_display_frame(self._h, int(x), int(y), int(z), int(component), _dp(out),
                                                          C.byref(ready))
        if rc != 0:
            return False, native.last_error(), None
        if not ready.value:
            return False, "", None
        a, b = p.nx * p.nz, p.ny * p.nz
        return True, "", [out[:a].reshape(p.nz, p.nx), out[a:a + b].reshape(p.nz, p.ny), out[a + b:].reshape(p.ny, p.nx)]

    def field(self, component: int) -> np.ndarray:
        """One field of the accepted state (padded storage), after the running batch finishes (snapshot())."""
        h = native.lib().swl_async3d_session(self._h)
        o = np.zeros(int(np.prod([self._params.nx + 6, self._params.ny + 6, self._params.nz + 6])))
        native.call("swl_sim3d_snapshot", h, int(component), _dp(o))
        return o


class Batch2D:
    """Steps many 2D sessions together (one launch per RK stage for all of them; graph-captured chunks).

    Sessions keep their own semantics: ``advance`` moves every member up to ``count`` steps, each
    stops at its own total, and a member's fault latches on that member only (``Simulation.fault``)."""

    STAT_NAMES = ["grid", "captures", "replays", "direct_chunks", "chunks", "tune_ms_0", "tune_ms_1", "tune_ms_2",
                  "tune_ms_3", "chunk", "capture_failures", "fused", "lanes", "lane_runs"]

    def __init__(self, device: int = 0):
        out = C.c_void_p()
        native.call("swl_batch2d_create", int(device), C.byref(out))
        self._h = out
        self._members = []

    def close(self):
        if self._h:
            native.lib().swl_batch2d_destroy(self._h)
            self._h = None
            self._members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, sim: "Simulation"):
        sim._need()
        native.call("swl_batch2d_add", self._h, sim._h)
        self._members.append(sim)  # keeps the session alive while it is a member

    def remove(self, sim: "Simulation"):
        native.call("swl_batch2d_remove", self._h, sim._h)
        self._members.remove(sim)

    def __len__(self):
        return native.lib().swl_batch2d_size(self._h)

    def advance(self, count: int):
        rc = native.lib().swl_batch2d_advance(self._h, int(count))
        return (rc == 0), ("" if rc == 0 else native.last_error())

    def set_option(self, name: str, value: int):
        native.call("swl_batch2d_set_option", self._h, name.encode(), int(value))

    def stats(self) -> dict:
        o = np.zeros(len(self.STAT_NAMES))
        native.lib().swl_batch2d_stats(self._h, _dp(o))
        return dict(zip(self.STAT_NAMES, o.tolist()))

    @staticmethod
    def build(params, device: int = 0, workers: int = 8):
        """Creates one HIP session per Parameters on `workers` threads (the native session builds run outside
        the GIL and overlap each other's host work and uploads) and a batch of them in the given order.
        Returns (batch, sessions); on a failed build the sessions already made are closed and RuntimeError names
        the shot."""
        import concurrent.futures as cf

        def one(p):
            s = Simulation()
            ok, err = s.initialize(p, BackendKind.Hip, device)
            return s, ("" if ok else err)

        with cf.ThreadPoolExecutor(max_workers=max(1, int(workers))) as pool:
            built = list(pool.map(one, params))
        failed = [(i, e) for i, (_, e) in enumerate(built) if e]
        if failed:
            for s, _ in built:
                s.close()
            raise RuntimeError(f"shot {failed[0][0]}: {failed[0][1]}")
        batch = Batch2D(device)
        sims = [s for s, _ in built]
        for s in sims:
            batch.add(s)
        return batch, sims
