"""The remote-control session (asc_remote_*) in the reference's socket
protocol, version 2 (Socket.cpp:53-444), on the CPU: the messages a session
answers without a device, the pose curves against the reference's own
KeyCurve.hpp (compiled into oracle/_ref), .tra export and import through
the session, and the wire format over a loopback TCP connection, packed as
PyAutoscoper packs it (little-endian 4-byte ints and 8-byte doubles after
a type byte).
"""
import ctypes as C
import socket
import struct
from pathlib import Path

import numpy as np
import pytest

from test_io import _trial_dir

GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


def msg(t, fmt="", *vals, tail=b""):
    return bytes([t]) + struct.pack("<" + fmt, *vals) + tail


@pytest.fixture()
def remote(asc):
    r = asc.Remote(None)
    yield r
    r.close()


def test_session_messages_without_a_device(asc, remote):
    assert remote.handle(msg(0)) == b"\x00"
    assert remote.handle(msg(16)) == msg(16, "i", 2)
    assert remote.handle(msg(14)) == msg(14, "i", 0)
    assert remote.handle(msg(1, tail=b"/no/such/trial.cfg")) == b"\x00"
    assert remote.handle(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    assert remote.handle(msg(14)) == msg(14, "i", 4)
    assert remote.handle(msg(15)) == msg(15, "i", 1)
    assert remote.handle(msg(5, "i", 0)) == b"\x05"
    assert remote.handle(msg(17, "ii", 2, 0)) == b"\x11"
    assert remote.handle(msg(17, "ii", 9, 0)) == b"\x00"
    vie = str(GOLD / "xParameters" / "control_settings.vie").encode()
    assert remote.handle(msg(4, "i", -1, tail=vie)) == b"\x04"
    assert remote.handle(msg(4, "i", 1, tail=vie)) == b"\x04"
    assert remote.handle(msg(4, "i", 5, tail=vie)) == b"\x00"
    assert remote.handle(msg(12)) == b"\x0c"
    # device requests fail on a session without a device
    assert remote.handle(msg(8, "i6d", 0, 0, 0, 0, 0, 0, 0)) == b"\x00"
    assert remote.handle(msg(9, "d", 0.5)) == b"\x09"  # answered, threshold unchanged
    assert remote.handle(msg(99)) == b"\x00"
    assert remote.handle(msg(13)) == b"" and remote.closed  # message 13 closes the connection


def _curves(orc):
    lib = orc.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref is not built (the reference sources are absent)")
    for name in ("ref_float_curve", "ref_quat_curve"):
        getattr(lib, name).restype = None
    return lib


def test_pose_curves_match_reference_keycurve(asc, orc, remote):
    """getPose between, before and after keys: x, y, z through the
    reference's KeyCurve<float> (Hermite, smooth tangents), the orientation
    through KeyCurve<Quatf> (slerp), keys inserted in the same order."""
    lib = _curves(orc)
    assert remote.handle(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    rng = np.random.default_rng(13)
    times = np.array([7, 2, 11, 4, 20, 11], np.int32)  # out of order, one key replaced
    poses = np.concatenate([rng.uniform(-40, 40, (6, 3)), rng.uniform(-170, 170, (6, 3))], 1)
    stores = []
    for t, p in zip(times, poses):
        assert remote.handle(msg(7, "ii6d", 1, int(t), *p)) == b"\x07"
        stores.append(asc.geometry.pose_store_set(p))
    stores = np.array(stores, np.float32)
    at = np.arange(-2, 24, dtype=np.float32)
    m = at.size
    ref = np.zeros((m, 7), np.float32)
    out = np.zeros(m, np.float32)
    for k in range(3):
        vals = np.ascontiguousarray(stores[:, k])
        lib.ref_float_curve(6, times.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p), m,
                            at.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        ref[:, k] = out
    q = np.ascontiguousarray(stores[:, 3:7])
    qo = np.zeros((m, 4), np.float32)
    lib.ref_quat_curve(6, times.ctypes.data_as(C.c_void_p), q.ctypes.data_as(C.c_void_p), m,
                       at.ctypes.data_as(C.c_void_p), qo.ctypes.data_as(C.c_void_p))
    ref[:, 3:] = qo
    for i, f in enumerate(at.astype(int)):
        got = struct.unpack("<6d", remote.handle(msg(6, "ii", 1, int(f)))[1:])
        assert np.array_equal(got, asc.geometry.pose_store_get(ref[i])), f
    # another volume's curves are untouched: the identity at the origin
    assert struct.unpack("<6d", remote.handle(msg(6, "ii", 0, 3))[1:]) == (0.0,) * 6


def test_tracking_export_and_import(asc, remote, tmp_path):
    root, _, _ = _trial_dir(tmp_path, asc, 5)
    cfg = root / "t.cfg"
    cfg.write_text("Version 1.1\nmayaCam_csv cams/c1.txt\nCameraRootDir frames\nVolumeFile vol/v.tif\n"
                   "VoxelSize 1 1 1\nVolumeFile vol/v.tif\nVoxelSize 1 1 1\n")
    assert remote.handle(msg(1, tail=str(cfg).encode())) == b"\x01"
    assert remote.handle(msg(15)) == msg(15, "i", 5)
    keys = {1: (1.5, -2.0, 3.25, 10.0, -20.0, 30.0), 3: (2.5, -1.0, 4.0, 12.0, -18.0, 33.0)}
    for f, p in keys.items():
        assert remote.handle(msg(7, "ii6d", 0, f, *p)) == b"\x07"
    out = tmp_path / "deep" / "dir" / "r.tra"  # the parent directory is created
    for interpolate in (0, 1):
        # volume 0, xyzypr rows, comma separated, mm, degrees
        assert remote.handle(msg(3, "7i", 0, 0, 0, 1, 0, 0, interpolate, tail=str(out).encode())) == b"\x03"
        rows = [ln.split(",") for ln in out.read_text().splitlines()]
        assert len(rows) == 5
        for f in range(5):
            got = struct.unpack("<6d", remote.handle(msg(6, "ii", 0, f))[1:])
            if f in keys or interpolate:
                assert [float(v) for v in rows[f]] == pytest.approx(got, abs=1e-12)
            else:
                assert rows[f] == ["NaN"] * 6
    # all volumes as matrices; import into a fresh session
    assert remote.handle(msg(3, "7i", -1, 1, 1, 1, 0, 0, 0, tail=str(out).encode())) == b"\x03"
    before = [struct.unpack("<6d", remote.handle(msg(6, "ii", 0, f))[1:]) for f in (1, 3)]
    r2 = asc.Remote(None)
    assert r2.handle(msg(1, tail=str(cfg).encode())) == b"\x01"
    assert r2.handle(msg(2, "7i", -1, 1, 1, 1, 0, 0, 0, tail=str(out).encode())) == b"\x02"
    after = [struct.unpack("<6d", r2.handle(msg(6, "ii", 0, f))[1:]) for f in (1, 3)]
    assert np.allclose(after, before, atol=1e-4)
    assert r2.handle(msg(2, "7i", -1, 1, 1, 1, 0, 0, 0, tail=b"/no/such.tra")) == b"\x00"
    r2.close()


def test_loopback_socket(asc, remote):
    port = remote.listen(0)
    s = socket.create_connection(("127.0.0.1", port), timeout=10)

    def send(m):
        s.sendall(m)
        return s.recv(1 << 16)

    assert send(msg(16)) == msg(16, "i", 2)  # the version check PyAutoscoper makes first
    assert send(msg(0)) == b"\x00"
    assert send(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    assert send(msg(7, "ii6d", 3, 0, 1.0, 2.0, 3.0, 0.0, 0.0, 0.0)) == b"\x07"
    got = struct.unpack("<6d", send(msg(6, "ii", 3, 0))[1:])
    assert got == pytest.approx((1.0, 2.0, 3.0, 0.0, 0.0, 0.0), abs=1e-5)
    s.sendall(msg(13))  # close: the server drops the connection
    assert s.recv(16) == b""
    s.close()
    s2 = socket.create_connection(("127.0.0.1", port), timeout=10)  # a new client is served
    s2.sendall(msg(14))
    assert s2.recv(16) == msg(14, "i", 4)
    s2.close()
