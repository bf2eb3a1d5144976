"""Host I/O of the path: the photon-map text dump and restore
(PhotonMap::dump / restore, src/pmap.cc:180-251), the PFM framebuffer
writer, and the 32-byte records of the backend status protocol
(emit_status, src/render.cc:279-289).

Pins: the dump text is compared with C printf "%f" output produced
independently (Python's %-formatting follows C's rounding); restored values
with libc strtof of the same text and the reference's float quotient /
double product for the colour scale; the PFM bytes with the PF layout
(bottom row first); the status records with "%c %6d %6d %6d %6d\\n" padded
with zeros to 32 bytes.
"""
import ctypes as C
import os

import numpy as np
import pytest

import sray_amd as S

_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.c_void_p]


def _strtof(s):
    return np.float32(_libc.strtof(s.encode(), None))


def _photons(n, seed=7):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-50, 50, size=(n, 12)).astype(np.float32)
    p[:, 9:] = rng.uniform(0, 3, size=(n, 3)).astype(np.float32)
    p[0, :] = [0.5, -1.25, 3.75, 0.0, -0.0, 1.0, 1e-7, -1e-7, 123456.789, 1e-3, 2.5e-6, 0.1]
    return p


def test_photon_dump_text(tmp_path):
    p = _photons(5)
    path = tmp_path / "pm.dump"
    S.photon_dump(path, p, 300.0)
    want = "energy %f\n" % 300.0
    for r in p.astype(np.float64):
        want += "<%f %f %f> <%f %f %f> <%f %f %f> <%f %f %f>\n" % tuple(r)
    assert path.read_text() == want


def test_photon_restore_scales_colours(tmp_path):
    p = _photons(40)
    path = tmp_path / "pm.dump"
    S.photon_dump(path, p, 300.0)
    got = S.photon_restore(path, 200.0)
    text = path.read_text().splitlines()[1:]
    scale = np.float32(200.0) / np.float32(300.0)
    want = np.empty_like(got)
    for i, line in enumerate(text):
        vals = [_strtof(v) for v in line.replace("<", " ").replace(">", " ").split()]
        want[i, :9] = vals[:9]
        want[i, 9:] = [np.float32(np.float64(v) * np.float64(scale)) for v in vals[9:]]
    assert got.shape == (40, 12)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_photon_round_trip_same_energy(tmp_path):
    p = _photons(100, seed=3)
    path = tmp_path / "pm.dump"
    S.photon_dump(path, p, 250.0)
    got = S.photon_restore(path, 250.0)
    want = np.vectorize(lambda v: _strtof("%f" % v), otypes=[np.float32])(p.astype(np.float64))
    assert np.array_equal(got, want)
    assert np.abs(got - p).max() <= 5e-7 + 1e-6 * np.abs(p).max()


def test_photon_restore_errors(tmp_path):
    bad = tmp_path / "bad.dump"
    bad.write_text("<1 2 3> <1 2 3> <1 2 3> <1 2 3>\n")
    with pytest.raises(S.SrayError):
        S.photon_restore(bad, 300.0)
    short = tmp_path / "short.dump"
    short.write_text("energy 300.000000\n<1 2 3> <1 2 3> <1 2 3>\n")
    with pytest.raises(S.SrayError):
        S.photon_restore(short, 300.0)
    with pytest.raises(S.SrayError):
        S.photon_restore(tmp_path / "missing.dump", 300.0)
    empty = tmp_path / "empty.dump"
    empty.write_text("energy 300.000000\n")
    assert S.photon_restore(empty, 300.0).shape == (0, 12)


def test_pfm_layout(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.uniform(-1, 4, size=(3, 5, 4)).astype(np.float32)
    path = tmp_path / "fb.pfm"
    S.write_pfm(path, img)
    data = path.read_bytes()
    head = b"PF\n5 3\n-1.0\n"
    assert data[:len(head)] == head
    body = np.frombuffer(data[len(head):], "<f4").reshape(3, 5, 3)
    assert np.array_equal(body, img[::-1, :, :3])


def test_status_records():
    for st, args in (("f", (40, 23, 48, 16)), ("s", (0, 48, 48, 48)), ("e", (1872, 1056, 48, 24)),
                     ("d", (123456, 0, 0, 0)), ("s", (-5, -123456, 0, 7))):
        rec = S.status_record(st, *args)
        want = ("%c %6d %6d %6d %6d\n" % ((st,) + args)).encode()
        assert len(rec) == 32 and rec == want + b"\0" * (32 - len(want))
    with pytest.raises(S.SrayError):
        S.status_record("d", 10 ** 9, 10 ** 9, 0, 0)
