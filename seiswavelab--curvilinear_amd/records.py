"""Receiver records as SEG-Y: the reference's SeisWave::writeSegy / readSegy (include/SeisCommon/SegyIo.hpp:5-19,
src/common/SegyIo.cpp) over the C-ABI of libswl (swl_segy_*).

``write_segy`` takes traces of one component; ``read_segy`` returns ``ImportedRecords`` as the reference's
reader fills it (SeismicRecords.hpp ImportedRecords / ImportedTrace).  ``write_session_segy`` exports a 2D
session's receivers the way the reference's experiment driver assembles them (Experiment2D.cpp:44-71).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import native


class SegyExportProfile(enum.IntEnum):
    """SegyIo.hpp:5: rev 2 IEEE32 / IEEE64 with the exact interval; rev 1 IBM32 / IEEE32 (whole microseconds)."""
    ExactIeee32 = 0
    ExactIeee64 = 1
    LegacyIbm32 = 2
    CompatibleIeee32 = 3


class _Trace(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("source", C.c_double * 3), ("start_time", C.c_double),
                ("samples", C.c_void_p)]


@dataclass
class TraceRecord:
    position: Sequence[float]
    source: Sequence[float]
    samples: np.ndarray
    start_time: float = 0.0


@dataclass
class ImportedTrace:
    position: tuple
    source: tuple
    start_time: float
    samples: np.ndarray


@dataclass
class ImportedRecords:
    dt: float = 0.0
    component: str = "Amplitude"
    units: str = "raw"
    description: str = ""
    traces: List[ImportedTrace] = field(default_factory=list)


def write_segy(path, traces: Sequence[TraceRecord], dt: float, component: str, units: str = "m/s",
               profile: SegyExportProfile = SegyExportProfile.ExactIeee32, gain: float = 1.0,
               sample_interval: float = 0.0):
    """Returns (ok, error) like the reference (a failed write leaves the destination unchanged)."""
    arrays = [np.ascontiguousarray(t.samples, dtype=np.float64) for t in traces]
    count = len(arrays[0]) if arrays else 0
    if any(len(a) != count for a in arrays):
        return False, "SEG-Y export requires equal trace lengths."
    cs = (_Trace * max(len(arrays), 1))()
    for r, (t, a) in enumerate(zip(traces, arrays)):
        cs[r].position[:] = [float(v) for v in t.position]
        cs[r].source[:] = [float(v) for v in t.source]
        cs[r].start_time = float(t.start_time)
        cs[r].samples = a.ctypes.data
    rc = native.lib().swl_segy_write(str(path).encode(), C.cast(cs, C.c_void_p), len(arrays), count, float(dt),
                                     component.encode(), units.encode(), int(profile), float(gain),
                                     float(sample_interval))
    return (rc == 0), ("" if rc == 0 else native.last_error())


def read_segy(path):
    """Returns (ok, error, ImportedRecords or None)."""
    lib = native.lib()
    h = C.c_void_p()
    if lib.swl_segy_read(str(path).encode(), C.byref(h)) != 0:
        return False, native.last_error(), None
    try:
        n, dt = C.c_int(), C.c_double()
        comp, units, desc = (C.create_string_buffer(256) for _ in range(3))
        native.call("swl_segy_records_info", h, C.byref(n), C.byref(dt), comp, units, desc, 256)
        out = ImportedRecords(dt.value, comp.value.decode(), units.value.decode(), desc.value.decode())
        for r in range(n.value):
            pos, src = np.zeros(3), np.zeros(3)
            start, count = C.c_double(), C.c_int64()
            native.call("swl_segy_records_trace", h, r, pos.ctypes.data, src.ctypes.data, C.byref(start),
                        C.byref(count))
            s = np.zeros(count.value)
            native.call("swl_segy_records_samples", h, r, s.ctypes.data)
            out.traces.append(ImportedTrace(tuple(pos), tuple(src), start.value, s))
        return True, "", out
    finally:
        lib.swl_segy_records_destroy(h)


def write_session_segy(sim, path, component: int, profile: SegyExportProfile = SegyExportProfile.ExactIeee32,
                       gain: float = 1.0, sample_interval: float = 0.0):
    """A 2D Simulation's receivers, component 0 (Vx) or 1 (Vz), one file.  Returns (ok, error)."""
    sim._need()
    rc = native.lib().swl_sim2d_write_segy(sim._h, str(path).encode(), int(component), int(profile), float(gain),
                                           float(sample_interval))
    return (rc == 0), ("" if rc == 0 else native.last_error())
