"""Host mirror of ``Seis2D::Curvilinear::Simulation`` (include/Seis2D/CurvilinearSolver2D.hpp:93-123).

Same method names and error behaviour as the reference: ``initialize``, ``advance`` and
``reset`` return ``(ok, error)``; a rejected step latches a fault until the session is
rebuilt; initialization is transactional (a failed call keeps the previous session).
All stepping runs in the HIP library; this class only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import native
from .params import BackendKind, Parameters, to_c

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
        o = np.zeros(10, np.int64)
        native.lib().swl_sim2d_layout(self._h, _dp(o))
        return dict(zip(LAYOUT_NAMES, (int(v) for v in o)))

    def diagnostics(self) -> dict:
        o = np.zeros(len(DIAGNOSTIC_NAMES))
        native.lib().swl_sim2d_diagnostics(self._h, _dp(o))
        d = dict(zip(DIAGNOSTIC_NAMES, o.tolist()))
        d["finite"] = bool(d["finite"])
        return d

    def receiver_arrays(self):
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
        o = np.zeros(self.layout()["receivers"], np.int64)
        native.lib().swl_sim2d_probe_indices(self._h, _dp(o))
        return o

    def model_array(self, name_or_index) -> np.ndarray:
        which = MODEL_ARRAYS.get(name_or_index, name_or_index)
        L = self.layout()
        n = L["nx"] * L["nz"]
        size = {10: 2 * n, 11: 2 * n, 14: 3 * (L["nx"] - 6), 15: 3 * (L["nz"] - 6)}.get(which, n)
        o = np.zeros(size)
        native.call("swl_sim2d_model", self._h, int(which), _dp(o))
        return o

    def snapshot_field(self, f: int) -> np.ndarray:
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
        L = self.layout()
        o = np.zeros(max(L["cpml"], 1))
        native.call("swl_sim2d_pml", self._h, _dp(o))
        return o[: L["cpml"]]

    def profile(self, enable: bool = True):
        """Per-launch HIP-event timing of the stage and finalize kernels (resets counters)."""
        native.call("swl_sim2d_profile", self._h, int(enable))

    def kernel_timing(self) -> dict:
        o = np.zeros(9)
        native.call("swl_sim2d_kernel_timing", self._h, _dp(o))
        return {"stage_ms": o[0], "stage_launches": int(o[1]), "finalize_ms": o[2], "finalize_launches": int(o[3]),
                "interior_tiles": int(o[4]), "vgprs": int(o[5]), "nodes_per_launch": int(o[6]),
                "border_tiles": int(o[7]), "metric_on_the_fly": bool(o[8])}

    def set_option(self, name: str, value: int):
        """Engine tuning switch (currently "nontemporal_stores")."""
        native.call("swl_sim2d_set_option", self._h, name.encode(), int(value))

    def readFrame(self, component: int, includeAbsorber: bool = False):
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
        xz = np.ascontiguousarray(np.asarray(nodes, np.int32).reshape(-1, 2))
        out = np.zeros((xz.shape[0], 5))
        rc = native.lib().swl_sim2d_read_probes(self._h, _dp(xz), xz.shape[0], _dp(out))
        if rc != 0:
            return False, native.last_error(), None
        return True, "", out


def writeReceivers(sim: Simulation, path: str):
    """Receiver CSV with the reference's provenance header (CurvilinearSolver2D.cpp:811-854)."""
    rc = native.lib().swl_sim2d_write_receivers(sim._h, str(path).encode())
    return (rc == 0), ("" if rc == 0 else native.last_error())
