"""The reference's file formats, host side (no GPU): Fortran unformatted
sequential records (output.f90:103-166, input.f90:106-170) and the
ener_time.data / spectrum.data text of spectral_dns_dealiased.f90:120-171.

Pins: records written by the library are read by a Fortran program built
here with amdflang (tests/fortran/record_io.f90) and the reverse; short,
truncated, corrupted and over-long files are rejected with SCL_EINVAL; a
record split into subrecords (gfortran's convention for records over 2 GiB,
built here with small subrecords) reads back whole; ener_time lines are
byte-identical to the ones the Fortran runtime writes with '(I6,20E12.4)'."""
import os
import struct
import subprocess

import numpy as np
import pytest

import speccodelib_amd as scl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RIO = os.path.join(ROOT, "speccodelib_amd", "fortran", "_build", "record_io")
needs_fortran = pytest.mark.skipif(not os.path.exists(RIO), reason="Fortran program not built (amdflang)")


def _field(n):
    i = np.arange(1, n + 1, dtype=np.float32)
    return (i - 0.5j * i).astype(np.complex64)


def test_record_round_trip(tmp_path):
    a = (np.random.default_rng(0).standard_normal((4, 8, 5)) + 1j).astype(np.complex64)
    p = tmp_path / "ux3.dat"
    scl.record_write(p, a)
    raw = p.read_bytes()
    assert struct.unpack("<i", raw[:4])[0] == a.nbytes and struct.unpack("<i", raw[-4:])[0] == a.nbytes
    assert len(raw) == a.nbytes + 8
    np.testing.assert_array_equal(scl.record_read(p, a.shape, a.dtype), a)


@needs_fortran
def test_fortran_reads_library_record_and_back(tmp_path):
    n = 1000
    p = tmp_path / "lib.dat"
    scl.record_write(p, _field(n))
    r = subprocess.run([RIO, "r", str(p), str(n)], capture_output=True, text=True, timeout=60)
    assert r.stdout.strip() == "ok", r.stdout + r.stderr
    q = tmp_path / "f.dat"
    subprocess.run([RIO, "w", str(q), str(n)], check=True, timeout=60)
    np.testing.assert_array_equal(scl.record_read(q, (n,), np.complex64), _field(n))


def _rejected(path, shape=(100,)):
    with pytest.raises(scl.SclError) as e:
        scl.record_read(path, shape, np.complex64)
    assert e.value.status == -1
    return str(e.value)


def test_bad_records_rejected(tmp_path):
    good = tmp_path / "good.dat"
    scl.record_write(good, _field(100))
    raw = good.read_bytes()
    cases = {
        "truncated data": raw[:500],
        "no trailing marker": raw[:-4],
        "empty": b"",
        "disagreeing markers": raw[:-4] + struct.pack("<i", 799),
        "trailing bytes": raw + b"\0",
        "another length": struct.pack("<i", 792) + raw[4:796] + struct.pack("<i", 792),
        "a second record": raw + raw,
    }
    for name, data in cases.items():
        p = tmp_path / "bad.dat"
        p.write_bytes(data)
        msg = _rejected(p)
        print(name, "->", msg)
    _rejected(tmp_path / "missing.dat")


def test_subrecords_read_whole(tmp_path):
    a = _field(300)
    b = a.tobytes()
    parts = [b[:1000], b[1000:1700], b[1700:]]
    out = b""
    for k, part in enumerate(parts):
        lead = -len(part) if k < len(parts) - 1 else len(part)
        trail = len(part) if k == 0 else -len(part)
        out += struct.pack("<i", lead) + part + struct.pack("<i", trail)
    p = tmp_path / "split.dat"
    p.write_bytes(out)
    np.testing.assert_array_equal(scl.record_read(p, a.shape, a.dtype), a)
    # a first subrecord whose trailing marker claims a predecessor is rejected
    bad = struct.pack("<i", -1000) + b[:1000] + struct.pack("<i", -1000) + out[1008:]
    p.write_bytes(bad)
    _rejected(p, a.shape)


@needs_fortran
def test_ener_time_lines_match_fortran(tmp_path):
    f = tmp_path / "fortran.data"
    subprocess.run([RIO, "e", str(f)], check=True, timeout=60)
    vals = np.array([[0.0, 1.0e-3, 1.5, -0.25], [0.123456, 9.99951e-1, 1234.5678, -6.02e-23],
                     [12.0, 0.002, 1.4541583, 3.3e7], [-1.0, 5.0e-5, 99999.0, 0.5],
                     [7.25, -3.0e-9, 0.1, 2.0]], np.float32)
    g = tmp_path / "lib.data"
    for i, v in enumerate(vals, 1):
        scl.write_ener_time(g, 10 * i - 7, *[float(x) for x in v])
    assert g.read_text() == f.read_text(), (g.read_text(), f.read_text())


def test_spectrum_block_reads_back(tmp_path):
    p = tmp_path / "spectrum.data"
    ek = np.array([0.5, 1.25e-3, 3.0e-7])
    scl.write_spectrum(p, 0.25, ek)
    scl.write_spectrum(p, 0.5, ek * 2)
    lines = p.read_text().split("\n")
    assert float(lines[0]) == 0.25 and float(lines[4]) == 0.5
    ll, e = lines[1].split()
    assert int(ll) == 1 and float(e) == float(np.float32(0.5))
    assert [int(x.split()[0]) for x in lines[5:8]] == [1, 2, 3]
    np.testing.assert_allclose([float(x.split()[1]) for x in lines[5:8]], ek * 2, rtol=1e-7)
