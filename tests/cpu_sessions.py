"""ctypes wrappers for the CPU checkers: the oracle restatement (oracle/liboracle.so)
and the reference's own CPU core (oracle/_ref/libswlref.so, present only where the
reference tree was available at build time).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libswlref.so"

_libs = {}


def have_ref() -> bool:
    """True where the reference build exists (this container only; it never travels to the GPU box).
    Checks the file without loading it, so merely collecting a test module loads nothing."""
    return REF_SO.exists()


def load(kind: str):
    """kind = 'orc' (oracle) or 'ref' (reference build)."""
    if kind in _libs:
        return _libs[kind]
    path = ORACLE_SO if kind == "orc" else REF_SO
    if not path.exists():
        return None
    lib = C.CDLL(str(path))
    for dim in ("2d", "3d"):
        pre = f"{kind}{dim}_"
        if not hasattr(lib, pre + "create"):
            continue
        getattr(lib, pre + "create").restype = C.c_void_p
        getattr(lib, pre + "create").argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        getattr(lib, pre + "destroy").argtypes = [C.c_void_p]
        getattr(lib, pre + "advance").argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        getattr(lib, pre + "dt").restype = C.c_double
        getattr(lib, pre + "dt").argtypes = [C.c_void_p]
        for n in ("step", "total"):
            getattr(lib, pre + n).argtypes = [C.c_void_p]
        for n in ("layout", "diagnostics", "receivers", "field", "model", "probe_indices", "pml", "reset",
                  "record_dt"):
            if hasattr(lib, pre + n):
                getattr(lib, pre + n).argtypes = None
        if hasattr(lib, pre + "record_dt"):
            getattr(lib, pre + "record_dt").restype = C.c_double
            getattr(lib, pre + "record_dt").argtypes = [C.c_void_p]
    for name in (f"{kind}2d_set_peak", f"{kind}3d_set_source_peak"):
        if hasattr(lib, name):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [C.c_void_p, C.c_double]
    if hasattr(lib, f"{kind}2d_pml_slot"):
        f = getattr(lib, f"{kind}2d_pml_slot")
        f.restype = C.c_int64
        f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int]
    _libs[kind] = lib
    return lib


DIAG2D = ["energy", "peakEnergy", "maxVelocity", "tractionRelative", "finite", "waveDt", "dampingDt", "sourceDt",
          "cpmlDt", "relaxationDt", "ppw", "maxRelativeQError", "maxMetric", "minJacobian", "maxEdge",
          "flattenDepth", "sourceWeightSum", "sourceRmsRadiusM"]


class CpuSession2D:
    """Drives orc2d_* / ref2d_* through the same calls as the reference's Simulation."""

    def __init__(self, kind: str, params):
        from seiswavelab_curvilinear_amd.params import to_c
        self.lib = load(kind)
        if self.lib is None:
            raise FileNotFoundError(kind)
        self.pre = f"{kind}2d_"
        c, self._keep = to_c(params)
        self._c = c
        err = C.create_string_buffer(512)
        self.h = getattr(self.lib, self.pre + "create")(C.byref(c), err, 512)
        self.error = err.value.decode()
        if not self.h:
            raise ValueError(self.error)

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    def close(self):
        if self.h:
            self._f("destroy")(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, n):
        err = C.create_string_buffer(512)
        ok = self._f("advance")(C.c_void_p(self.h), int(n), err, 512)
        return bool(ok), err.value.decode()

    def reset(self):
        return bool(self._f("reset")(C.c_void_p(self.h)))

    def set_peak(self, peak):
        """Oracle test hook: the accepted peak energy of the late-growth fault rule."""
        self._f("set_peak")(C.c_void_p(self.h), float(peak))

    @property
    def step(self):
        return self._f("step")(C.c_void_p(self.h))

    @property
    def total(self):
        return self._f("total")(C.c_void_p(self.h))

    @property
    def dt(self):
        return self._f("dt")(C.c_void_p(self.h))

    @property
    def record_dt(self):
        return self._f("record_dt")(C.c_void_p(self.h))

    def layout(self):
        o = np.zeros(10, np.int64)
        self._f("layout")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        keys = ["nx", "nz", "offsetX", "widthX", "widthZ", "fields", "cpml", "receivers", "records", "mechanisms"]
        return dict(zip(keys, (int(v) for v in o)))

    def diagnostics(self):
        o = np.zeros(len(DIAG2D))
        self._f("diagnostics")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        return dict(zip(DIAG2D, o.tolist()))

    def receivers(self):
        L = self.layout()
        nr, n = L["receivers"], L["records"]
        vx, vz = np.zeros((nr, n)), np.zeros((nr, n))
        pos, xz = np.zeros((nr, 2)), np.zeros((nr, 2), np.int32)
        self._f("receivers")(C.c_void_p(self.h), vx.ctypes.data_as(C.c_void_p), vz.ctypes.data_as(C.c_void_p),
                             pos.ctypes.data_as(C.c_void_p), xz.ctypes.data_as(C.c_void_p))
        return {"vx": vx, "vz": vz, "position": pos, "xz": xz}

    def probe_indices(self):
        L = self.layout()
        o = np.zeros(L["receivers"], np.int64)
        self._f("probe_indices")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        return o

    def field(self, f):
        L = self.layout()
        o = np.zeros(L["nx"] * L["nz"])
        self._f("field")(C.c_void_p(self.h), int(f), o.ctypes.data_as(C.c_void_p))
        return o

    def pml(self):
        L = self.layout()
        o = np.zeros(max(L["cpml"], 1))
        self._f("pml")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        return o[: L["cpml"]]

    def model(self, which):
        L = self.layout()
        n = L["nx"] * L["nz"]
        size = {10: 2 * n, 11: 2 * n, 14: 3 * (L["nx"] - 6), 15: 3 * (L["nz"] - 6)}.get(which, n)
        o = np.zeros(size)
        ok = self._f("model")(C.c_void_p(self.h), int(which), o.ctypes.data_as(C.c_void_p))
        return o if ok else None


def rel_l2(ref, val):
    ref = np.asarray(ref, dtype=np.float64).ravel()
    val = np.asarray(val, dtype=np.float64).ravel()
    return float(np.sqrt(np.sum((ref - val) ** 2) / max(np.sum(ref * ref), 1e-300)))


DIAG3D = ["energy", "peakEnergy", "sourcePeakEnergy", "maxAbsVelocity", "maxTractionRateRelative", "allFinite",
          "minJacobian", "maxMetricNorm", "minCellLength", "maxCellEdge", "waveDtLimit", "dampingDtLimit",
          "sourceDtLimit", "peakVz", "peakVx", "peakVy", "peakV", "peakPressure"]


class CpuSession3D:
    """Drives orc3d_* / ref3d_* (CpuSimulation3D on the CPU backend)."""

    def __init__(self, kind: str, params):
        from seiswavelab_curvilinear_amd.params import to_c3d
        self.lib = load(kind)
        if self.lib is None or not hasattr(self.lib, f"{kind}3d_create"):
            raise FileNotFoundError(kind)
        self.pre = f"{kind}3d_"
        c, self._keep = to_c3d(params)
        self._c = c
        err = C.create_string_buffer(512)
        self.h = getattr(self.lib, self.pre + "create")(C.byref(c), err, 512)
        self.error = err.value.decode()
        if not self.h:
            raise ValueError(self.error)

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    def close(self):
        if self.h:
            self._f("destroy")(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, n):
        err = C.create_string_buffer(512)
        ok = self._f("advance")(C.c_void_p(self.h), int(n), err, 512)
        return bool(ok), err.value.decode()

    @property
    def step(self):
        return self._f("step")(C.c_void_p(self.h))

    @property
    def total(self):
        return self._f("total")(C.c_void_p(self.h))

    @property
    def dt(self):
        return self._f("dt")(C.c_void_p(self.h))

    def set_source_peak(self, peak):
        """Oracle test hook: the source-window peak energy of the late-growth fault rule."""
        self._f("set_source_peak")(C.c_void_p(self.h), float(peak))

    def layout(self):
        o = np.zeros(5, np.int64)
        self._f("layout")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        return dict(zip(["nx", "ny", "nz", "receivers", "records"], (int(v) for v in o)))

    def count(self):
        L = self.layout()
        return L["nx"] * L["ny"] * L["nz"]

    def diagnostics(self):
        o = np.zeros(len(DIAG3D))
        self._f("diagnostics")(C.c_void_p(self.h), o.ctypes.data_as(C.c_void_p))
        return dict(zip(DIAG3D, o.tolist()))

    def receivers(self):
        L = self.layout()
        nr, n = L["receivers"], L["records"]
        vx, vy, vz = np.zeros((nr, n)), np.zeros((nr, n)), np.zeros((nr, n))
        pos, idx = np.zeros((nr, 3)), np.zeros(nr, np.int64)
        self._f("receivers")(C.c_void_p(self.h), *(a.ctypes.data_as(C.c_void_p) for a in (vx, vy, vz, pos, idx)))
        return {"vx": vx, "vy": vy, "vz": vz, "position": pos, "index": idx}

    def field(self, f):
        o = np.zeros(self.count())
        self._f("field")(C.c_void_p(self.h), int(f), o.ctypes.data_as(C.c_void_p))
        return o

    def model(self, which):
        n = self.count()
        size = {0: 9 * n, 2: 3 * n, 3: 3 * n}.get(which, n)
        o = np.zeros(size)
        ok = self._f("model")(C.c_void_p(self.h), int(which), o.ctypes.data_as(C.c_void_p))
        return o if ok else None
