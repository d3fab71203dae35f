"""SEG-Y records (csrc/segy.cpp through the C-ABI) against the reference's writeSegy / readSegy.

* Writer: byte-for-byte equal to the files the reference build wrote for the same record set
  (tests/golden/segy_*.sgy, five export variants), and to a live reference write where oracle/_ref exists.
* Reader: decodes those files to the values the writer was given (exact for IEEE64, float32 / IBM32
  rounding otherwise), and to what the reference reader returns where oracle/_ref exists; hand-built files
  cover the other formats (integer, IBM), little-endian order, rev 0 and trace weighting against their
  closed-form values.
* Errors: the reference's messages for invalid exports.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from cpu_sessions import have_ref, load
from reference_fixtures import SEGY_VARIANTS, segy_fixture_path, segy_records
from seiswavelab_curvilinear_amd.records import SegyExportProfile, TraceRecord, read_segy, write_segy

HAVE_REF = have_ref()


def traces():
    pos, src, start, s = segy_records()
    return [TraceRecord(pos[r], src[r], s[r], start[r]) for r in range(len(s))], s


def write_variant(path, name):
    profile, gain, interval, dt, comp, units = SEGY_VARIANTS[name]
    tr, _ = traces()
    return write_segy(path, tr, dt, comp, units, SegyExportProfile(profile), gain, interval)


def ref_lib():
    lib = load("ref")
    lib.ref_segy_write.restype = C.c_int
    lib.ref_segy_write.argtypes = [C.c_char_p] + [C.c_void_p] * 4 + [C.c_int, C.c_int64, C.c_double, C.c_char_p,
                                                                     C.c_char_p, C.c_int, C.c_double, C.c_double,
                                                                     C.c_char_p, C.c_int]
    lib.ref_segy_read.restype = C.c_void_p
    lib.ref_segy_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    lib.ref_segy_free.argtypes = [C.c_void_p]
    lib.ref_segy_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p, C.c_char_p,
                                  C.c_char_p, C.c_int]
    lib.ref_segy_trace.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int64), C.c_void_p]
    return lib


def ref_read(path):
    lib = ref_lib()
    err = C.create_string_buffer(256)
    h = lib.ref_segy_read(str(path).encode(), err, 256)
    if not h:
        return None, err.value.decode()
    n, dt = C.c_int(), C.c_double()
    comp, units, desc = (C.create_string_buffer(256) for _ in range(3))
    lib.ref_segy_info(h, C.byref(n), C.byref(dt), comp, units, desc, 256)
    out = {"dt": dt.value, "component": comp.value.decode(), "units": units.value.decode(),
           "description": desc.value.decode(), "traces": []}
    for r in range(n.value):
        pos, src, start, cnt = np.zeros(3), np.zeros(3), C.c_double(), C.c_int64()
        lib.ref_segy_trace(h, r, pos.ctypes.data, src.ctypes.data, C.byref(start), C.byref(cnt), None)
        s = np.zeros(cnt.value)
        lib.ref_segy_trace(h, r, pos.ctypes.data, src.ctypes.data, C.byref(start), C.byref(cnt), s.ctypes.data)
        out["traces"].append((pos, src, start.value, s))
    lib.ref_segy_free(h)
    return out, ""


def same_records(ours, ref):
    assert ours.dt == ref["dt"] and ours.component == ref["component"] and ours.units == ref["units"]
    assert ours.description == ref["description"]
    assert len(ours.traces) == len(ref["traces"])
    for t, (pos, src, start, s) in zip(ours.traces, ref["traces"]):
        assert np.array_equal(t.position, pos) and np.array_equal(t.source, src) and t.start_time == start
        assert np.array_equal(t.samples, s)


@pytest.mark.parametrize("name", list(SEGY_VARIANTS))
def test_writer_bytes_equal_reference_files(name, tmp_path):
    path = tmp_path / f"{name}.sgy"
    ok, err = write_variant(path, name)
    assert ok, err
    assert path.read_bytes() == segy_fixture_path(name).read_bytes()
    assert not list(tmp_path.glob("*.partial*"))


@pytest.mark.skipif(not HAVE_REF, reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("name", list(SEGY_VARIANTS))
def test_writer_bytes_equal_live_reference(name, tmp_path):
    profile, gain, interval, dt, comp, units = SEGY_VARIANTS[name]
    pos, src, start, s = (np.ascontiguousarray(a) for a in segy_records(seed=99, n=9, count=131))
    lib = ref_lib()
    err = C.create_string_buffer(256)
    assert lib.ref_segy_write(str(tmp_path / "ref.sgy").encode(), pos.ctypes.data, src.ctypes.data, start.ctypes.data,
                              s.ctypes.data, s.shape[0], s.shape[1], dt, comp.encode(), units.encode(), profile, gain,
                              interval, err, 256), err.value
    tr = [TraceRecord(pos[r], src[r], s[r], start[r]) for r in range(len(s))]
    ok, e = write_segy(tmp_path / "ours.sgy", tr, dt, comp, units, SegyExportProfile(profile), gain, interval)
    assert ok, e
    assert (tmp_path / "ours.sgy").read_bytes() == (tmp_path / "ref.sgy").read_bytes()


@pytest.mark.parametrize("name", list(SEGY_VARIANTS))
def test_reader_decodes_reference_files(name):
    profile, gain, interval, dt, comp, units = SEGY_VARIANTS[name]
    ok, err, rec = read_segy(segy_fixture_path(name))
    assert ok, err
    pos, src, start, s = segy_records()
    odt = interval or dt
    assert rec.component == comp and rec.units == (units if gain == 1 else f"scaled({units})")
    assert len(rec.traces) == len(s)
    if profile == SegyExportProfile.ExactIeee64 or profile == SegyExportProfile.ExactIeee32:
        assert rec.dt == odt * 1e6 * 1e-6  # the exact interval, via its IEEE64 microseconds
    else:
        assert rec.dt == round(odt * 1e6) * 1e-6
    for r, t in enumerate(rec.traces):
        np.testing.assert_array_equal(np.round(np.asarray(t.position) * 1000), np.round(pos[r] * 1000))
        assert t.start_time == start[r]
        if interval:
            continue
        if profile == SegyExportProfile.ExactIeee64:
            assert np.array_equal(t.samples, s[r] * gain)
        elif profile in (SegyExportProfile.ExactIeee32, SegyExportProfile.CompatibleIeee32):
            assert np.array_equal(t.samples, (s[r] * gain).astype(np.float32).astype(np.float64))


@pytest.mark.skipif(not HAVE_REF, reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("name", list(SEGY_VARIANTS))
def test_reader_equals_reference_reader(name):
    ok, err, ours = read_segy(segy_fixture_path(name))
    assert ok, err
    ref, rerr = ref_read(segy_fixture_path(name))
    assert ref is not None, rerr
    same_records(ours, ref)


def handmade(path, fmt, values, little=False, revision=1, weighting=0, marker=True):
    """A minimal SEG-Y file: blank textual header, one trace of `values` in format `fmt`."""
    e = "<" if little else ">"
    text = b" " * 3200
    binary = bytearray(400)
    struct.pack_into(e + "H", binary, 16, 1000)        # 1 ms
    struct.pack_into(e + "H", binary, 20, len(values))
    struct.pack_into(e + "H", binary, 24, fmt)
    if revision >= 2 and marker:
        struct.pack_into(e + "I", binary, 96, 0x01020304)
    binary[300] = revision
    header = bytearray(240)
    struct.pack_into(e + "h", header, 68, -100), struct.pack_into(e + "h", header, 70, 10)
    struct.pack_into(e + "i", header, 80, 123), struct.pack_into(e + "i", header, 84, -7)
    struct.pack_into(e + "i", header, 40, 250)
    struct.pack_into(e + "H", header, 114, len(values))
    struct.pack_into(e + "h", header, 168, weighting)
    struct.pack_into(e + "h", header, 108, 4)
    code = {2: "i", 3: "h", 8: "b", 5: "f", 6: "d", 4: "i"}
    if fmt == 1:
        data = b"".join(struct.pack(e + "I", v) for v in values)
    else:
        data = b"".join(struct.pack(e + code[fmt], v) for v in values)
    path.write_bytes(text + bytes(binary) + bytes(header) + data)


@pytest.mark.parametrize("fmt,values,expected", [
    (2, [7, -3, 2 ** 30], [7, -3, 2 ** 30]),
    (3, [-32768, 5, 32767], [-32768, 5, 32767]),
    (8, [-128, 1, 127], [-128, 1, 127]),
    (5, [0.5, -1.25, 3e-3], np.float32([0.5, -1.25, 3e-3]).astype(np.float64)),
    # IBM: 0x41100000 = +1.0 (16^1 * 1/16), 0xC276A000 = -118.625 (the textbook example), 0 = 0
    (1, [0x41100000, 0xC276A000, 0], [1.0, -118.625, 0.0]),
])
@pytest.mark.parametrize("little", [False, True])
@pytest.mark.parametrize("weighting", [0, 3])
def test_reader_formats_orders_and_weighting(fmt, values, expected, little, weighting, tmp_path):
    path = tmp_path / "h.sgy"
    handmade(path, fmt, values, little=little, revision=1, weighting=weighting)
    ok, err, rec = read_segy(path)
    assert ok, err
    t = rec.traces[0]
    np.testing.assert_array_equal(t.samples, np.ldexp(np.asarray(expected, dtype=np.float64), -weighting))
    assert rec.dt == 1e-3
    # header scalars: xy 10 (multiply), elevation -100 (divide); z positive down
    assert t.position == (1230.0, -70.0, -2.5)
    assert t.start_time == 0.004
    assert rec.description == f"SEG-Y rev 1 | format {fmt} | {'little' if little else 'big'} endian"
    if HAVE_REF:
        ref, rerr = ref_read(path)
        assert ref is not None, rerr
        same_records(rec, ref)


def test_reader_rev0_and_errors(tmp_path):
    path = tmp_path / "r0.sgy"
    handmade(path, 5, [1.0, 2.0], revision=0)
    ok, err, rec = read_segy(path)
    assert ok, err
    assert rec.traces[0].start_time == 0.004 and rec.units == "raw"
    handmade(path, 4, [1, 2])  # format 4 (fixed point with gain) is not supported
    ok, err, _ = read_segy(path)
    assert not ok and err == "Unsupported SEG-Y sample format. Supported codes: 1,2,3,5,6,8."
    path.write_bytes(b"x" * 100)
    ok, err, _ = read_segy(path)
    assert not ok and err == "SEG-Y input is shorter than its file headers."


def test_write_errors_use_reference_messages(tmp_path):
    tr, _ = traces()
    path = tmp_path / "bad.sgy"
    ok, err = write_segy(path, tr, 1.9310594e-4, "Vz", "m/s", SegyExportProfile.ExactIeee32, gain=0.0)
    assert not ok and err == "SEG-Y export multiplier must be finite and positive."
    ok, err = write_segy(path, tr, 1.9310594e-4, "Vz", "m/s", SegyExportProfile.LegacyIbm32)
    assert not ok and err == "Legacy SEG-Y requires whole-microsecond dt; select 0.5 or 1 ms output sampling."
    ok, err = write_segy(path, [], 1e-3, "Vz")
    assert not ok and err == "Select an existing raw component before SEG-Y export."
    tr[1].start_time = 0.0015
    ok, err = write_segy(path, tr, 1e-3, "Vz")
    assert not ok and err == "SEG-Y export delay must be representable as whole milliseconds."
    assert not path.exists()


def test_compatible_round_trip_is_float32_of_the_input(tmp_path):
    """Experiment2D.cpp:56-70: CompatibleIeee32 written, read back, every sample == float32(raw)."""
    tr, s = traces()
    path = tmp_path / "Vx.sgy"
    ok, err = write_segy(path, tr, 2.5e-4, "Vx", "m/s", SegyExportProfile.CompatibleIeee32)
    assert ok, err
    ok, err, rec = read_segy(path)
    assert ok, err
    assert rec.dt == 2.5e-4 and len(rec.traces) == len(s)
    for r, t in enumerate(rec.traces):
        assert np.array_equal(t.samples, s[r].astype(np.float32).astype(np.float64))
