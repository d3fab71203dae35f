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


# ---- the 500x200 sine-terrain workload at full size (seed 1, FP64, all 2000 steps), run by the reference build
FULL_2D = "sine_terrain_500x200"
FULL_2D_TRACE_STRIDE = 10  # full traces of every 10th receiver are kept; the L2 norm of every receiver's traces


def reduce_full_2d(run):
    """The stored form of a full-size run: every field of run_oracle_2d except the traces, which are kept in full
    for every FULL_2D_TRACE_STRIDE-th receiver and as per-receiver L2 norms (Vx, Vz) for all."""
    out = {k: v for k, v in run.items() if k not in ("vx", "vz")}
    out["vx_kept"] = run["vx"][::FULL_2D_TRACE_STRIDE]
    out["vz_kept"] = run["vz"][::FULL_2D_TRACE_STRIDE]
    out["trace_norms"] = np.stack([np.sqrt(np.sum(run["vx"] ** 2, axis=1)), np.sqrt(np.sum(run["vz"] ** 2, axis=1))], 1)
    return out


# ---- 2D receiver CSV (writeReceivers, CurvilinearSolver2D.cpp:811-854) written by the reference build after
#      the case's full run
RECEIVERS2D_CASES = ["sponge_tilted_force", "moment_external_pml"]


def receivers2d_fixture_path(name):
    return GOLDEN / f"receivers2d_{name}.csv.gz"


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


# ---- VTS model documents and wavefield snapshots (tests/golden/vts_*.vts.gz, the reference's own writers) ----
VTS_MODEL_VARIANTS = ["v2_sponge", "v2_gsls_layered_surface", "v3_pml", "v4_external", "v5_fp32_gsls",
                      "v6_physical", "v7_target_moment"]
VTS_SNAPSHOT_CASES = {"v2_gsls_layered_surface": 30, "v4_external": 30, "v5_fp32_gsls": 30}  # name: steps
VTS_ALL_COMPONENTS = [1, 1, 1, 1, 1, 1, 1]


def vts_case(name):
    """Small (24 x 20) sessions, one per schema feature of the model document."""
    from seiswavelab_curvilinear_amd.params import Absorber, Layer, Parameters, SourceKind
    p = Parameters(nx=24, nz=20, spacing=2.5, relief=2, wavelength=40, sourceZ=8, receiverStart=2,
                   receiverCount=4, receiverStride=4, receiverDepth=2, duration=.01, maxSteps=40, spongeX=4,
                   spongeZ=4, cpuThreads=1)
    p.source.direction = (1, 0, 2)
    if name == "v2_gsls_layered_surface":
        p.attenuation.enabled, p.attenuation.mechanisms = True, 2
        p.layered = True
        p.layers = [Layer(12, 1800, 900, 2000, 40, 25), Layer(0, 3000, 1700, 2300, 80, 50)]
        p.importedSurface, p.surfaceDatum = True, 500
        p.surfaceSamples = [(0, 499), (20, 501.5), (40, 500.25), (60, 499.5)]
        p.sampleIntervalMs = .25
    elif name == "v3_pml":
        p.absorber = Absorber.CfsPml
    elif name == "v4_external":
        p.absorber, p.externalAbsorption = Absorber.CfsPml, True
    elif name == "v5_fp32_gsls":
        p.precision, p.absorber = Precision.Float32, Absorber.CfsPml
        p.attenuation.enabled = True
    elif name == "v6_physical":
        p.sourceSigmaM, p.sourceRadiusM = 2.0, 5.0
    elif name == "v7_target_moment":
        p.receiverTargetCount = 12
        p.source.kind = SourceKind.IsotropicMoment
    return p


def vts_fixture_path(name, step=None):
    return GOLDEN / (f"vts_model_{name}.vts.gz" if step is None else f"vts_snapshot_{name}_step{step}.vts.gz")


# ---- 3D receiver CSV and wavefield VTS (tests/golden/vts3d_*, receivers3d_*: the reference's own writers) ----
VTS3D_CASES = ["analytic_sponge_force", "dem_layers_moment"]  # cases of tests/test_oracle_3d.py
RECEIVERS3D_EXPORTS = {"analytic_sponge_force": "vz", "dem_layers_moment": "x_z"}  # legacy Vz / Vx+Vz columns


def vts3d_fixture_path(name):
    """writeVts of the step-0 state: all eleven fields, the materials and the velocity vector."""
    return GOLDEN / f"vts3d_{name}_step0.vts.gz"


def receivers3d_fixture_path(name):
    """writeReceiversCsv after the case's maxSteps, in the RECEIVERS3D_EXPORTS form."""
    return GOLDEN / f"receivers3d_{name}_{RECEIVERS3D_EXPORTS[name]}.csv.gz"


def vts_text(path):
    import gzip
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt", encoding="ascii", newline="") as f:
        return f.read()


def parse_vts(text):
    """Independent reading of a VTS file (2D documents and snapshots, 3D snapshots): {"field": {name: [values]},
    "point": {name: array; the 3D velocity vector as (n, 3)}, "coordinates": (n, 3) array, "order": [field array
    names in file order]}."""
    import re
    out = {"field": {}, "point": {}, "coordinates": None, "order": []}
    field_end = text.index("</FieldData>")
    for m in re.finditer(r"<DataArray ([^>]*)>([^<]*)</DataArray>", text):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        values = np.array([float(v) for v in m.group(2).split()])
        if m.start() < field_end:
            out["field"][attrs["Name"]] = values
            out["order"].append(attrs["Name"])
        elif attrs.get("NumberOfComponents") == "3" and attrs.get("Name") != "Velocity":
            out["coordinates"] = values.reshape(-1, 3)
        else:
            out["point"][attrs["Name"]] = values.reshape(-1, 3) if attrs.get("NumberOfComponents") == "3" else values
    return out


def fnv1a64(data: bytes, h=0xcbf29ce484222325):
    """FNV-1a 64-bit, byte by byte (Fowler-Noll-Vo)."""
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def model_checksum(doc):
    """The document checksum recomputed from the text: names (NUL-terminated) and values (IEEE-754 little
    endian, -0 as +0) of every metadata array in file order except ModelChecksum, then the six point arrays
    in file order, then Coordinates (x, 0, z per node)."""
    import struct
    h = 0xcbf29ce484222325
    number = lambda v: struct.pack("<d", 0.0 if v == 0 else float(v))  # noqa: E731
    for name in doc["order"]:
        if name == "ModelChecksum":
            continue
        h = fnv1a64(name.encode() + b"\0" + b"".join(number(v) for v in doc["field"][name]), h)
    for name, values in doc["point"].items():
        h = fnv1a64(name.encode() + b"\0" + b"".join(number(v) for v in values), h)
    return fnv1a64(b"Coordinates\0" + b"".join(number(v) for v in doc["coordinates"].ravel()), h)
