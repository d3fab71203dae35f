"""The 2D engine seam: the HIP engine that replaces the reference's CudaEngine
(src/two_d/CudaEngine2D.hpp:5-24), driven through the engine-level C-ABI of include/swl.h.

The reference's façade builds an FP64 ``Point::MeshT`` (src/two_d/PointKernels2D.hpp:14-27) in padded
storage ``q = i + nx*k`` and hands it to ``CudaEngine(mesh, receivers, device)``; its ``advance()`` then
calls ``step(time, dt, frequency, scale, stepNumber)``, ``diagnostics``, ``rollback`` and ``sample`` once
per step (src/two_d/CurvilinearSolver2D.cpp:553-597).  :class:`Engine2D` takes the same mesh as a dict of
FP64 arrays and offers the same methods, so a caller holding the reference's mesh can swap engines.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np

from . import native

_dp = C.POINTER(C.c_double)


class Mesh2DDesc(C.Structure):
    """swl_mesh2d_desc (include/swl.h)."""
    _fields_ = [("precision", C.c_int), ("nx", C.c_int), ("nz", C.c_int), ("h", C.c_double),
                ("a", _dp), ("b", _dp), ("jac", _dp), ("normals", _dp), ("lambda_", _dp), ("mu", _dp), ("rho", _dp),
                ("source", _dp), ("damping", _dp), ("source_direction", C.c_double * 3),
                ("source_gradient", _dp * 2), ("mechanisms", C.c_int), ("field_count", C.c_int),
                ("cpml_x", _dp), ("cpml_z", _dp), ("cpml_width_x", C.c_int), ("cpml_width_z", C.c_int),
                ("cpml_count", C.c_int64), ("rates", C.c_double * 3), ("delta_p", _dp * 3), ("delta_s", _dp * 3),
                ("surface_depth", _dp), ("surface_slope", _dp), ("fade", _dp), ("fade_derivative", _dp)]


class Diag2D(C.Structure):
    """swl_diag2d (include/swl.h): CudaEngine::Diagnostics."""
    _fields_ = [("energy", C.c_double), ("max_velocity", C.c_double), ("traction_relative", C.c_double),
                ("finite", C.c_int), ("rhs_registers", C.c_int), ("rhs_local_bytes", C.c_uint64)]


MESH_ARRAYS = ("a", "b", "jac", "normals", "lambda", "mu", "rho", "source", "damping")


class Engine2D:
    """CudaEngine over the C-ABI.

    ``mesh`` keys: precision (0 FP64, 1 FP32), nx, nz (padded), h, the FP64 arrays a, b, jac, normals (2 per
    node), lambda, mu, rho, source, damping (padded storage), direction (3), gradient (pair of arrays for an
    explosive source, else None), mechanisms, fields, and for a CFS-PML cpml_x, cpml_z (3 per computational
    node), width_x, width_z, cpml_count; for GSLS rates (3) and delta_p, delta_s (lists of arrays).
    ``receivers``: padded storage index of each probe, ``(x + 3 + offsetX) + nx*(z + 3)``."""

    def __init__(self, mesh: Dict, receivers: Sequence[int], device: int = 0):
        self._h = None
        d = Mesh2DDesc()
        keep = []

        def arr(a, n=None):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if n is not None and a.size != n:
                raise ValueError(f"mesh array of {a.size} values where {n} are needed")
            keep.append(a)
            return a.ctypes.data_as(_dp)

        nx, nz = int(mesh["nx"]), int(mesh["nz"])
        n = nx * nz
        d.precision, d.nx, d.nz, d.h = int(mesh.get("precision", 0)), nx, nz, float(mesh["h"])
        for name in MESH_ARRAYS:
            setattr(d, "lambda_" if name == "lambda" else name, arr(mesh[name], 2 * n if name == "normals" else n))
        d.source_direction = (C.c_double * 3)(*mesh.get("direction", (0.0, 0.0, 1.0)))
        g = mesh.get("gradient")
        if g is not None:
            d.source_gradient[0], d.source_gradient[1] = arr(g[0], n), arr(g[1], n)
        d.mechanisms = int(mesh.get("mechanisms", 0))
        d.field_count = int(mesh.get("fields", 5))
        if mesh.get("cpml_count", 0):
            d.cpml_x, d.cpml_z = arr(mesh["cpml_x"], 3 * (nx - 6)), arr(mesh["cpml_z"], 3 * (nz - 6))
            d.cpml_width_x, d.cpml_width_z = int(mesh["width_x"]), int(mesh["width_z"])
            d.cpml_count = int(mesh["cpml_count"])
        d.rates = (C.c_double * 3)(*(list(mesh.get("rates", ())) + [0.0, 0.0, 0.0])[:3])
        for l in range(d.mechanisms):
            d.delta_p[l], d.delta_s[l] = arr(mesh["delta_p"][l], n), arr(mesh["delta_s"][l], n)
        rec = np.ascontiguousarray(receivers, dtype=np.int64)
        h = C.c_void_p()
        native.call("swl_engine2d_create", C.byref(d), rec.ctypes.data_as(C.c_void_p), int(rec.size), int(device),
                    C.byref(h))
        self._h, self.nx, self.nz, self.fields, self.cpml = h, nx, nz, d.field_count, int(d.cpml_count)
        self.receivers = int(rec.size)

    def close(self):
        if self._h:
            native.lib().swl_engine2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # CudaEngine::step / rollback / reset (CudaEngine2D.cu:143-154)
    def step(self, time: float, dt: float, frequency: float, scale: float, step: int):
        native.call("swl_engine2d_step", self._h, float(time), float(dt), float(frequency), float(scale), int(step))

    def rollback(self):
        native.call("swl_engine2d_rollback", self._h)

    def reset(self):
        native.call("swl_engine2d_reset", self._h)

    @staticmethod
    def _diag(o: Diag2D) -> dict:
        return {"energy": o.energy, "maxVelocity": o.max_velocity, "tractionRelative": o.traction_relative,
                "finite": bool(o.finite), "rhsRegisters": o.rhs_registers, "rhsLocalBytes": o.rhs_local_bytes}

    def diagnostics(self) -> dict:
        """CudaEngine::diagnostics (CudaEngine2D.cu:155-158)."""
        o = Diag2D()
        native.call("swl_engine2d_diagnostics", self._h, C.byref(o))
        return self._diag(o)

    def diagnostics_noalloc(self) -> dict:
        """diagnostics() without any device allocation: the inspection's per-tile partials go to a fixed
        device array (no counterpart in the reference; for callers that cannot allocate)."""
        o = Diag2D()
        native.call("swl_engine2d_diagnostics_noalloc", self._h, C.byref(o))
        return self._diag(o)

    def sample(self) -> np.ndarray:
        """CudaEngine::sample (CudaEngine2D.cu:159-162): (receivers, 2) of Vx, Vz."""
        o = np.zeros(2 * max(self.receivers, 1))
        native.call("swl_engine2d_sample", self._h, o.ctypes.data_as(C.c_void_p))
        return o[: 2 * self.receivers].reshape(-1, 2)

    def snapshot(self, field: int) -> np.ndarray:
        o = np.zeros(self.nx * self.nz)
        native.call("swl_engine2d_snapshot", self._h, int(field), o.ctypes.data_as(C.c_void_p))
        return o

    def pml(self) -> np.ndarray:
        o = np.zeros(max(self.cpml, 1))
        native.call("swl_engine2d_pml", self._h, o.ctypes.data_as(C.c_void_p))
        return o[: self.cpml]

    def probes(self, indices: Sequence[int]) -> np.ndarray:
        """CudaEngine::probes: (n, 5) of Vx, Vz, Sxx, Szz, Sxz at padded indices."""
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        o = np.zeros(5 * max(idx.size, 1))
        native.call("swl_engine2d_probes", self._h, idx.ctypes.data_as(C.c_void_p), int(idx.size),
                    o.ctypes.data_as(C.c_void_p))
        return o[: 5 * idx.size].reshape(-1, 5)

    def frame(self, component: int, nx: int, nz: int, offset_x: int = 0) -> np.ndarray:
        o = np.zeros(nx * nz, np.float32)
        native.call("swl_engine2d_frame", self._h, int(component), o.ctypes.data_as(C.c_void_p), int(nx), int(nz),
                    int(offset_x))
        return o.reshape(nz, nx)

    def residentBytes(self) -> int:
        return int(native.lib().swl_engine2d_resident_bytes(self._h))


def mesh_from_arrays(model, layout: dict, params, precision: int = 0) -> Dict:
    """The engine mesh from FP64 model arrays numbered as the oracle's and the session's model queries
    (0 a, 1 b, 2 J, 3 lambda, 4 mu, 5 rho, 6 source, 7 damping, 8/9 moment gradient, 10 normals,
    14/15 CFS-PML profiles): ``model(which)`` returns the array.  Elastic models only (the GSLS rates are
    not among those arrays)."""
    from .params import Absorber, SourceKind
    explosive = params.source.kind == SourceKind.IsotropicMoment
    mesh = {"precision": precision, "nx": layout["nx"], "nz": layout["nz"], "h": params.spacing,
            "a": model(0), "b": model(1), "jac": model(2), "lambda": model(3), "mu": model(4), "rho": model(5),
            "source": model(6), "damping": model(7), "normals": model(10),
            "direction": (0.0, 0.0, 0.0) if explosive else tuple(np.asarray(params.source.direction, float) /
                                                                  np.linalg.norm(params.source.direction)),
            "gradient": (model(8), model(9)) if explosive else None,
            "mechanisms": 0, "fields": 5}
    if layout["cpml"] and params.absorber == Absorber.CfsPml:
        mesh.update(cpml_x=model(14), cpml_z=model(15), width_x=layout["widthX"], width_z=layout["widthZ"],
                    cpml_count=layout["cpml"])
    return mesh


def source_scale(params) -> float:
    """The stepping scale the reference's advance() passes: amplitude / h^2 (CurvilinearSolver2D.cpp:557-561)."""
    from .params import SourceKind
    amplitude = params.momentPerLength if params.source.kind == SourceKind.IsotropicMoment else params.forcePerLength
    return amplitude / (params.spacing * params.spacing)
