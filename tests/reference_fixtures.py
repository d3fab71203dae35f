"""Reference-build fixtures (tests/golden/reference_*.npz, written by
tests/golden/make_reference_fixtures.py from oracle/_ref): what they hold and how a run is reduced to
the same record, for the oracle (CPU sessions) and for the HIP path (Simulation / Simulation3D)."""
import pathlib

import numpy as np

from seiswavelab_curvilinear_amd.params import Precision

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
CASES_2D = ["gsls2_layered_pml", "moment_external_pml", "sine_terrain_small", "sponge_tilted_force"]
FIXTURES_2D = [(name, Precision.Float64) for name in CASES_2D] + [("sine_terrain_small", Precision.Float32)]


def fixture_path(dim, name, prec=Precision.Float64):
    tag = "f64" if prec == Precision.Float64 else "f32"
    return GOLDEN / f"reference_{dim}_{name}_{tag}.npz"


def load_fixture(dim, name, prec=Precision.Float64):
    with np.load(fixture_path(dim, name, prec), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def field_stats(fields):
    """Per field: L2 norm, sum, max |value|."""
    return np.array([[np.sqrt(np.sum(f * f)), np.sum(f), np.max(np.abs(f))] for f in fields])


def record_2d(s, total, energy_of, fields):
    """Steps a 2D session one step at a time; s has advance(n) -> (ok, err)."""
    energy, peak = [], []
    for _ in range(total):
        ok, err = s.advance(1)
        assert ok, err
        e, pk = energy_of(s)
        energy.append(e), peak.append(pk)
    return np.array(energy), np.array(peak), field_stats(fields(s))


def run_oracle_2d(kind, p):
    from cpu_sessions import CpuSession2D
    s = CpuSession2D(kind, p)
    energy, peak, stats = record_2d(s, s.total, lambda x: (x.diagnostics()["energy"], x.diagnostics()["peakEnergy"]),
                                    lambda x: [x.field(f) for f in range(x.layout()["fields"])])
    r = s.receivers()
    out = {"dt": np.array(s.dt), "total": np.array(s.total), "record_dt": np.array(s.record_dt), "energy": energy,
           "peak_energy": peak, "vx": r["vx"], "vz": r["vz"], "field_stats": stats}
    s.close()
    return out


def run_oracle_3d(kind, p):
    from cpu_sessions import CpuSession3D
    s = CpuSession3D(kind, p)
    energy, peak, stats = record_2d(s, p.maxSteps,
                                    lambda x: (x.diagnostics()["energy"], x.diagnostics()["peakEnergy"]),
                                    lambda x: [x.field(f) for f in range(9)])
    r = s.receivers()
    out = {"dt": np.array(s.dt), "total": np.array(s.total), "energy": energy, "peak_energy": peak,
           "vx": r["vx"], "vy": r["vy"], "vz": r["vz"], "field_stats": stats}
    s.close()
    return out
