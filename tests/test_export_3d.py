"""3D receiver CSV and wavefield VTS export (csrc/vts3d.cpp, modelio.py) against the reference's own writers.

Pinned by: files written by the reference build (tests/golden/vts3d_*_step0.vts.gz: writeVts of the step-0
state with every field, the materials and the velocity vector; tests/golden/receivers3d_*.csv.gz:
writeReceiversCsv after the case's steps), which ours must reproduce byte for byte at t = 0 (state exactly
zero on both sides) and, after stepping, line for line in the header with the values at the FP64 bar
(1e-10 relative L2); the reference geometry in those files against the oracle's exactly; stepped exports
against our own raw state exactly and against the oracle at the path's tolerance (FP64 1e-10, FP32 1e-4);
and the error paths of src/three_d/ReceiverIo.cpp:12-19 and src/three_d/VtsWriter.cpp:19-35.
"""
import numpy as np
import pytest

from cpu_sessions import CpuSession3D
from reference_fixtures import (RECEIVERS3D_EXPORTS, VTS3D_CASES, parse_vts, receivers3d_fixture_path,
                                vts3d_fixture_path, vts_text)
from seiswavelab_curvilinear_amd import BackendKind, Simulation3D, VtsExportOptions3D, writeReceiversCsv, writeVts
from test_oracle_3d import case

FIELDS = ["Vx", "Vy", "Vz", "VelocityMagnitude", "MeanNormalStress", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"]


def physical(nx, ny, nz):
    """Padded indices (ghost width 3) of the physical nodes, x fastest, then y, then z: the VTS point order."""
    k, j, i = np.meshgrid(np.arange(nz) + 3, np.arange(ny) + 3, np.arange(nx) + 3, indexing="ij")
    return ((k * (ny + 6) + j) * (nx + 6) + i).ravel()


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _csv(text):
    """(comment and header lines, value rows as an array)."""
    lines = text.splitlines()
    head = [l for l in lines if l.startswith("#")] + [lines[sum(1 for l in lines if l.startswith("#"))]]
    rows = np.array([[float(v) for v in l.split(",")] for l in lines[len(head):]])
    return head, rows


# ---------------------------------------------------------------------------------------------------------
# CPU: the reference's files, and argument checks that need no session

@pytest.mark.parametrize("name", VTS3D_CASES)
def test_reference_vts3d_geometry_and_materials_equal_oracle(name):
    """The reference's step-0 file against the oracle's model: point coordinates, Vp / Vs / density per node
    (from the moduli for layered models) exactly, eleven zero fields, the provenance scalars."""
    p = case(name)
    o = CpuSession3D("orc", p)
    doc = parse_vts(vts_text(vts3d_fixture_path(name)))
    idx = physical(p.nx, p.ny, p.nz)
    assert np.array_equal(doc["coordinates"], o.model(3).reshape(-1, 3)[idx])
    assert doc["field"]["SimulationStep"][0] == 0 and doc["field"]["SimulationDt"][0] == o.dt
    assert doc["field"]["GridSpacing"][0] == p.spacing
    for f in FIELDS:
        assert doc["point"][f].shape == (idx.size,) and not doc["point"][f].any(), f
    assert doc["point"]["Velocity"].shape == (idx.size, 3) and not doc["point"]["Velocity"].any()
    lam, mu, rho = (o.model(w)[idx] for w in (4, 5, 6))
    if p.layeredMaterial:
        assert np.array_equal(doc["point"]["Vp"], np.sqrt((lam + 2 * mu) / rho))
        assert np.array_equal(doc["point"]["Vs"], np.sqrt(mu / rho))
        assert np.array_equal(doc["point"]["Density"], rho)
        ids = doc["point"]["LayerId"]
        assert set(np.unique(ids)) <= set(range(len(p.layers))) and len(np.unique(ids)) > 1
        # layer ids follow the materials: every node of one layer carries that layer's density
        for n, layer in enumerate(p.layers):
            assert np.all(rho[ids == n] == layer.density), n
    else:
        assert "LayerId" not in doc["point"]
        for k, v in (("Vp", p.vp), ("Vs", p.vs), ("Density", p.density)):
            assert np.all(doc["point"][k] == v), k


def test_export_arguments_without_session(tmp_path):
    ok, err = writeReceiversCsv(None, tmp_path / "r.csv", (False, False, False))
    assert not ok and err == "Select at least one receiver component."
    ok, err = writeReceiversCsv(None, tmp_path / "r.csv")
    assert not ok and err == "Initialize a simulation before exporting traces."
    ok, err = writeVts(None, tmp_path / "s.vts", VtsExportOptions3D([False] * 11, True, False))
    assert not ok and err == "Select at least one wavefield component."
    ok, err = writeVts(None, tmp_path / "s.vts")
    assert not ok and err == "Uninitialized session"
    ok, err = writeVts(None, tmp_path / "s.vts", VtsExportOptions3D([True] * 3))
    assert not ok and err == "VTS export needs eleven field flags."
    assert not list(tmp_path.iterdir())


# ---------------------------------------------------------------------------------------------------------
# GPU: exports of HIP sessions

def _hip(p, precision=0):
    print(f"[export3d] nx={p.nx} ny={p.ny} nz={p.nz} precision={precision}", flush=True)
    s = Simulation3D(BackendKind.Hip)
    ok, err = s.initialize(p, 0, precision)
    assert ok, err
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision", [(n, 0) for n in VTS3D_CASES] + [("analytic_sponge_force", 1)])
def test_vts3d_step0_byte_identical_to_reference(name, precision, tmp_path):
    s = _hip(case(name), precision)
    ok, err = writeVts(s, tmp_path / "s.vts")
    assert ok, err
    assert (tmp_path / "s.vts").read_bytes() == vts_text(vts3d_fixture_path(name)).encode()
    assert [f.name for f in tmp_path.iterdir()] == ["s.vts"]
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision", [(n, 0) for n in VTS3D_CASES] + [("dem_layers_moment", 1)])
def test_vts3d_after_steps(name, precision, tmp_path):
    """Every exported array equals our state exactly (|V| to the last bit of hypot), the materials and the
    coordinates equal the reference's step-0 file, and the fields match the oracle at the path tolerance."""
    p = case(name)
    s = _hip(p, precision)
    ok, err = s.advance(p.maxSteps)
    assert ok, err
    ok, err = writeVts(s, tmp_path / "s.vts")
    assert ok, err
    ours = parse_vts((tmp_path / "s.vts").read_text())
    ref0 = parse_vts(vts_text(vts3d_fixture_path(name)))
    assert ours["order"] == ref0["order"] and list(ours["point"]) == list(ref0["point"])
    assert ours["field"]["SimulationStep"][0] == p.maxSteps
    assert ours["field"]["TIME_VALUE"][0] == p.maxSteps * s.dt()
    for k in ours["field"]:
        if k not in ("SimulationStep", "TIME_VALUE"):
            assert np.array_equal(ours["field"][k], ref0["field"][k]), k
    assert np.array_equal(ours["coordinates"], ref0["coordinates"])
    for k in ("Vp", "Vs", "Density", "LayerId"):
        if k in ref0["point"]:
            assert np.array_equal(ours["point"][k], ref0["point"][k]), k
    idx = physical(p.nx, p.ny, p.nz)
    raw = [s.field(f)[idx] for f in range(9)]
    exact = dict(zip(["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"], raw))
    exact["MeanNormalStress"] = -(raw[3] + raw[4] + raw[5]) / 3.0
    for k, v in exact.items():
        assert np.array_equal(ours["point"][k], v), k
    assert np.array_equal(ours["point"]["Velocity"], np.stack(raw[:3], axis=1))
    big = np.maximum(np.maximum(np.abs(raw[0]), np.abs(raw[1])), np.abs(raw[2]))
    unit = np.where(big > 0, big, 1.0)
    mag = big * np.sqrt((raw[0] / unit) ** 2 + (raw[1] / unit) ** 2 + (raw[2] / unit) ** 2)  # no underflow
    assert np.all(np.abs(ours["point"]["VelocityMagnitude"] - mag) <= 1e-15 * mag)
    o = CpuSession3D("orc", p)
    assert o.advance(p.maxSteps)[0]
    tol = 1e-4 if precision else 1e-10
    for n, k in enumerate(["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"]):
        assert _rel_l2(ours["point"][k], o.field(n)[idx]) <= tol, k
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", VTS3D_CASES)
def test_receivers_csv_matches_reference(name, tmp_path):
    """Header (provenance, source position and indices, receiver positions, column names) line for line; the
    time column exactly; the traces at the FP64 bar."""
    p = case(name)
    s = _hip(p)
    ok, err = s.advance(p.maxSteps)
    assert ok, err
    path = tmp_path / "sub" / "r.csv"
    if RECEIVERS3D_EXPORTS[name] == "vz":
        ok, err = writeReceiversCsv(s, path)
    else:
        ok, err = writeReceiversCsv(s, path, (True, False, True))
    assert ok, err
    head, rows = _csv(path.read_text())
    want_head, want_rows = _csv(vts_text(receivers3d_fixture_path(name)))
    assert head == want_head
    assert rows.shape == want_rows.shape == (p.maxSteps + 1, want_rows.shape[1])
    assert np.array_equal(rows[:, 0], want_rows[:, 0])
    assert np.abs(want_rows[:, 1:]).max() > 0
    assert _rel_l2(rows[:, 1:], want_rows[:, 1:]) <= 1e-10
    s.close()


@pytest.mark.gpu
def test_receivers_csv_all_components_are_the_recorded_traces(tmp_path):
    """allVelocityComponents: Vx, Vy, Vz per receiver with suffixed column names; %.17g reads back exactly."""
    p = case("dem_layers_moment")
    s = _hip(p)
    ok, err = s.advance(7)
    assert ok, err
    ok, err = writeReceiversCsv(s, tmp_path / "r.csv", allVelocityComponents=True)
    assert ok, err
    head, rows = _csv((tmp_path / "r.csv").read_text())
    a = s.receiver_arrays()
    nr = a["vz"].shape[0]
    assert head[0].startswith("# Vx,Vy,Vz in m/s")
    assert head[-1] == "time_s," + ",".join(f"receiver_{r + 1}_{c}" for r in range(nr) for c in ("Vx", "Vy", "Vz"))
    assert len(head) == 3 + nr + 1
    for r in range(nr):
        x, y, z = (float(v) for v in head[3 + r].split(": ")[1].split(","))
        assert (x, y, z) == tuple(a["position"][r])
    assert rows.shape == (8, 1 + 3 * nr)
    assert np.array_equal(rows[:, 0], np.arange(8) * s.dt())
    for c, k in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(rows[:, 1 + c::3].T, a[k][:, :8]), k
    s.close()


@pytest.mark.gpu
def test_vts3d_selection(tmp_path):
    """Only the selected arrays, in the reference's order; the Scalars attribute names the first selected
    field; the velocity vector alone is a valid export; folders are created; nothing is left behind."""
    p = case("analytic_sponge_force")
    s = _hip(p)
    assert s.advance(5)[0]
    sel = [False] * 11
    sel[4] = sel[9] = True
    ok, err = writeVts(s, tmp_path / "a" / "b.vts", VtsExportOptions3D(sel, False, False))
    assert ok, err
    text = (tmp_path / "a" / "b.vts").read_text()
    assert '<PointData Scalars="MeanNormalStress">' in text
    assert list(parse_vts(text)["point"]) == ["MeanNormalStress", "Sxz"]
    ok, err = writeVts(s, tmp_path / "v.vts", VtsExportOptions3D([False] * 11, False, True))
    assert ok, err
    text = (tmp_path / "v.vts").read_text()
    assert '<PointData Vectors="Velocity">' in text and list(parse_vts(text)["point"]) == ["Velocity"]
    ok, err = writeVts(s, tmp_path / "n.vts", VtsExportOptions3D([False] * 11, True, False))
    assert not ok and err == "Select at least one wavefield component."
    assert sorted(f.name for f in tmp_path.iterdir()) == ["a", "v.vts"]
    s.close()
