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


# ---- SEG-Y: one seeded record set written in five export variants by the reference's writeSegy
SEGY_VARIANTS = {  # name: (profile, gain, sample interval s, source dt s, component, units)
    "exact32": (0, 1.0, 0.0, 1.9310594e-4, "Vz", "m/s"),
    "exact64": (1, 1.0, 0.0, 1.9310594e-4, "Vx", "m/s"),
    "exact32_gain_resampled": (0, 2.5, 2.5e-4, 1.9310594e-4, "Vz", "m/s"),
    "legacy_ibm": (2, 1.0, 2e-4, 1.9310594e-4, "Vx", "m/s"),
    "compatible32": (3, 1.0, 0.0, 2.5e-4, "P", "Pa"),
}


def segy_records(seed=17, n=5, count=97):
    """Positions, sources, start times and samples (n x count) of the synthetic record set."""
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, 900, n), rng.uniform(-50, 50, n), rng.uniform(0, 12, n)])
    pos[0, 1] = 0.0
    src = np.tile([450.25, 0.0, 10.125], (n, 1))
    start = np.zeros(n)
    start[2] = 0.002
    s = rng.standard_normal((n, count)) * 1e-6
    s[0, :5] = 0.0
    s[1, 3], s[1, 4], s[1, 5] = 1e-46, -3e-39, 7.5  # float32 underflow, IBM-representable small value, large
    s[3, 10] = -0.5                                   # IBM fraction rounding at a hex-digit boundary
    return pos, src, start, s


def segy_fixture_path(name):
    return GOLDEN / f"segy_{name}.sgy"
