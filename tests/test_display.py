"""Headless display output on the CPU: the PNG writer (decoded here with
Python's zlib and checked chunk by chunk against the PNG specification:
signature, IHDR fields, CRC-32 of every chunk, the zlib stream and its
Adler-32), the view composite's conversion (GL's clamp to [0, 1] and
rounding to 8 bits, rows flipped from GL's bottom-up order), and
save_full_drr's PGM (Tracker.cpp:155-204) on hand-computed values.
"""
import struct
import zlib

import numpy as np
import pytest


def read_png(path):
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, chunks = 8, []
    while at < len(data):
        n, = struct.unpack(">I", data[at:at + 4])
        kind = data[at + 4:at + 8]
        body = data[at + 8:at + 8 + n]
        crc, = struct.unpack(">I", data[at + 8 + n:at + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        chunks.append((kind, body))
        at += 12 + n
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"] and chunks[2][1] == b""
    w, h, depth, ctype, comp, filt, lace = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, comp, filt, lace) == (8, 0, 0, 0) and ctype in (0, 2)
    ch = 3 if ctype == 2 else 1
    raw = zlib.decompress(chunks[1][1])  # checks the Adler-32 too
    assert len(raw) == h * (1 + w * ch)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * ch)
    assert (rows[:, 0] == 0).all()  # filter type 0 on every scanline
    img = rows[:, 1:].reshape(h, w, ch) if ch == 3 else rows[:, 1:].reshape(h, w)
    return img


@pytest.mark.parametrize("shape", [(5, 7), (3, 4, 3), (301, 257, 3)])  # the last spans several stored blocks
def test_png_round_trip(asc, tmp_path, shape):
    img = np.random.default_rng(1).integers(0, 256, shape, dtype=np.uint8)
    p = tmp_path / "a.png"
    asc.io.write_png(p, img)
    assert np.array_equal(read_png(p), img)


def test_view_png_conversion(asc, tmp_path):
    h, w = 4, 6
    rgb = np.linspace(-0.5, 1.5, h * w * 3, dtype=np.float32).reshape(h, w, 3)
    rgb[0, 0] = [np.nan, 0.5 / 255, 254.5 / 255]  # NaN -> 0; ties round up
    p = tmp_path / "v.png"
    asc.io.write_view_png(p, rgb)
    got = read_png(p)
    v = np.where(np.isnan(rgb), 0.0, np.clip(rgb, 0, 1)).astype(np.float32)
    want = np.floor(v * np.float32(255) + np.float32(0.5)).astype(np.uint8)[::-1]
    assert np.array_equal(got, want)
    assert tuple(got[-1, 0]) == (0, 1, 255)  # row 0 of the composite is the PNG's last row


def read_pgm(path):
    t = path.read_text()
    head, rest = t.split("\n", 3)[:3], t.split("\n", 3)[3]
    assert head[0] == "P2" and head[2] == "255"
    w, h = map(int, head[1].split())
    assert rest.endswith(" ") and "\n" not in rest
    return np.array(rest.split(), int).reshape(h, w)


def test_full_drr_pgm_known_values(asc, tmp_path):
    p = tmp_path / "d.pgm"
    # stretch over [0, 1]: 255 * 0.5 = 127.5 -> 127 -> 128 once inverted
    asc.io.write_full_drr_pgm(p, np.array([[0.0, 0.5, 1.0], [0.25, 0.75, 1.0]], np.float32))
    assert read_pgm(p).tolist() == [[255, 128, 0], [192, 64, 0]]
    # the maximum starts at the smallest positive float: an all-negative
    # image stretches over [min, ~0], so -1 of [-2, -1] is 127 (stored 128)
    asc.io.write_full_drr_pgm(p, np.array([[-2.0, -1.0]], np.float32))
    assert read_pgm(p).tolist() == [[255, 128]]
    # no range: 255 everywhere
    asc.io.write_full_drr_pgm(p, np.full((2, 2), 0.3, np.float32))
    assert (read_pgm(p) == 255).all()


def test_remote_full_drr_without_device(asc, tmp_path):
    r = asc.Remote(None)
    r.set_output_dir(tmp_path)
    assert r.handle(b"\x0c") == b"\x0c"  # answered as the reference does, nothing written
    assert list(tmp_path.iterdir()) == []
    r.close()
