"""2D model documents and wavefield snapshots as ASCII VTS files.

Mirror of ``writeModelVts`` / ``readModelVts`` (include/Seis2D/ModelIo2D.hpp:4-13) and ``writeSnapshot`` /
``displayedSnapshotOptions`` (include/Seis2D/CurvilinearSolver2D.hpp:124-140): same names, the same
``(ok, error)`` results and messages.  Formatting, parsing and the FNV-1a checksum run in the native
library (csrc/vts2d.cpp); files are byte-compatible with the reference's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import native
from .params import BackendKind, CParams2D, Parameters, from_c


@dataclass
class ComputeSettings:
    """ComputeSettings3D (include/SeisCommon/ComputeDevices.hpp:11-15): CPU 0 = auto; device -1 = largest."""
    cpuThreads: int = 0
    cudaDevice: int = -1


@dataclass
class ModelDocument:
    """ModelDocument (include/Seis2D/ModelIo2D.hpp:4-9) plus the schema version the file carried."""
    parameters: Parameters = field(default_factory=Parameters)
    compute: ComputeSettings = field(default_factory=ComputeSettings)
    backend: BackendKind = BackendKind.Cpu
    referenceDt: float = 0.0
    version: int = 0


@dataclass
class SnapshotExportOptions:
    """Vx, Vz, velocity magnitude, pressure, Sxx, Szz, Sxz; defaults are the reference's legacy export."""
    components: List[bool] = field(default_factory=lambda: [True, True, False, False, True, True, True])
    includeMaterial: bool = True


def displayedSnapshotOptions(component: int) -> SnapshotExportOptions:
    """Only the displayed component (0 Vz, 1 Vx, 2 |V|, 3 pressure), no materials."""
    o = SnapshotExportOptions([False] * 7, False)
    if 0 <= component < 4:
        o.components[(1, 0, 2, 3)[component]] = True
    return o


def _handle(sim):
    return getattr(sim, "_h", None) if sim is not None else None


def writeModelVts(sim, path, compute: Optional[ComputeSettings] = None) -> Tuple[bool, str]:
    """Self-contained model (not a wavefield checkpoint) of an initialized session; the document records
    BackendKind 1 for a HIP session and 0 for a device-free plan."""
    compute = compute or ComputeSettings()
    rc = native.lib().swl_sim2d_write_model_vts(_handle(sim), str(path).encode(), int(compute.cpuThreads),
                                                 int(compute.cudaDevice))
    return (rc == 0), ("" if rc == 0 else native.last_error())


def readModelVts(path) -> Tuple[bool, str, Optional[ModelDocument]]:
    """Reads this schema (versions 1..7), validating shape, finite data and the checksum."""
    lib = native.lib()
    h = C.c_void_p()
    if lib.swl_model_vts_read(str(path).encode(), C.byref(h)) != 0:
        return False, native.last_error(), None
    try:
        c = CParams2D()
        native.call("swl_model_doc2d_params", h, C.byref(c))
        info = np.zeros(5)
        native.call("swl_model_doc2d_info", h, info.ctypes.data_as(C.c_void_p))
        params = from_c(c)  # copies the surface samples out of the document
    finally:
        lib.swl_model_doc2d_destroy(h)
    params.cpuThreads = int(info[1])
    doc = ModelDocument(params, ComputeSettings(int(info[1]), int(info[2])),
                        BackendKind.Hip if int(info[0]) == 1 else BackendKind.Cpu, float(info[3]), int(info[4]))
    return True, "", doc


def writeSnapshot(sim, path, options: Optional[SnapshotExportOptions] = None) -> Tuple[bool, str]:
    """Wavefield VTS of the current state: selected components, derived speed and plane-strain pressure,
    materials on request; values in the session's precision, coordinates in FP64."""
    options = options or SnapshotExportOptions()
    if len(options.components) != 7:
        return False, "Snapshot export needs seven component flags."
    comps = (C.c_int32 * 7)(*[int(bool(v)) for v in options.components])
    rc = native.lib().swl_sim2d_write_snapshot(_handle(sim), str(path).encode(), comps, int(bool(options.includeMaterial)))
    return (rc == 0), ("" if rc == 0 else native.last_error())


# ---------------------------------------------------------------------------------------------------------
# 3D: receiver CSV (include/Seis3D/ReceiverIo.hpp) and wavefield VTS (include/Seis3D/VtsWriter.hpp)

@dataclass
class VtsExportOptions3D:
    """Vx, Vy, Vz, |V|, mean normal stress, Sxx, Syy, Szz, Sxy, Sxz, Syz; materials; the velocity vector."""
    fields: List[bool] = field(default_factory=lambda: [True] * 11)
    includeMaterial: bool = True
    includeVelocityVector: bool = True


def writeReceiversCsv(sim, path, components=None, allVelocityComponents: bool = False) -> Tuple[bool, str]:
    """Both overloads of the reference: `components` = (Vx, Vy, Vz) flags, or (without it) Vz with the legacy
    column names, or all three with `allVelocityComponents`."""
    lib = native.lib()
    if components is not None:
        if len(components) != 3:
            return False, "Receiver export needs three component flags."
        sel = (C.c_int32 * 3)(*[int(bool(v)) for v in components])
        rc = lib.swl_sim3d_write_receivers_csv(_handle(sim), str(path).encode(), sel)
    else:
        rc = lib.swl_sim3d_write_receivers_vz_csv(_handle(sim), str(path).encode(), int(bool(allVelocityComponents)))
    return (rc == 0), ("" if rc == 0 else native.last_error())


def writeVts(sim, path, options: Optional[VtsExportOptions3D] = None) -> Tuple[bool, str]:
    """Snapshot of the current 3D state as a VTK StructuredGrid of the physical nodes."""
    options = options or VtsExportOptions3D()
    if len(options.fields) != 11:
        return False, "VTS export needs eleven field flags."
    fields = (C.c_int32 * 11)(*[int(bool(v)) for v in options.fields])
    rc = native.lib().swl_sim3d_write_vts(_handle(sim), str(path).encode(), fields, int(bool(options.includeMaterial)),
                                          int(bool(options.includeVelocityVector)))
    return (rc == 0), ("" if rc == 0 else native.last_error())
