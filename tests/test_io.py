"""On-disk inputs and outputs (asc_tiff_*, asc_read_camera, .vie, .cfg,
.tra) against independent decoders, published examples, closed forms and
the reference's sample data (tests/golden/sample_data, copied from the
reference's sample_data by tests/golden/copy_sample_data.py).

How each part is pinned:
  * TIFF: round trips through encoders written here (PackBits, TIFF LZW with
    its early code-width change, predictor 2, both byte orders, several
    pages and strips); the PackBits example of the TIFF 6.0 specification
    (section 9); the sample volume and the first 200 rows of a sample
    radiograph decoded by a pure-Python decoder written here.
  * MayaCam 2.0: equal (bit for bit) to the oracle's restatement of
    Camera::loadMayaCam2 on the sample calibrations.
  * MayaCam 1.0 / VTKCam 1.0: the reference's loader needs jsoncpp, which is
    absent, so they are pinned by the published formulas
    (camera-calibration.md) and identities: the frame is
    CoordFrame::from_xyzypr of the file's values (bit-exact against the
    oracle's frame math), the film rectangle of the image plane equals the
    viewport, the optical axis passes through the principal point, and a
    VTKCam view angle of 2a spans a film width of 4 tan(a) on both axes.
  * .vie: the effective values the GUI code (FilterTreeWidget*.cpp) gives for
    sample_data/xParameters/control_settings.vie and empty_filter.vie, and
    spin-box rounding and bounds. Rounding of an exact binary tie (e.g.
    0.125) follows Qt's fixed formatting (away from zero): parity unpinned
    for that case.
  * .cfg / .tra: the rules of Trial.cpp and AutoscoperMainWindow.cpp on
    files written here; .tra matrices against the oracle's frame math.
"""
import math
import struct
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


# ---------------------------------------------------------------------------
# TIFF encoders and decoders written for the tests
# ---------------------------------------------------------------------------
def packbits_encode(data: bytes) -> bytes:
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i
        while j + 1 < n and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i:  # a run of j - i + 1 equal bytes
            out += bytes([(257 - (j - i + 1)) & 0xFF, data[i]])
            i = j + 1
            continue
        k = i
        while k + 1 < n and k - i < 127 and not (k + 2 < n and data[k + 1] == data[k + 2]):
            k += 1
        out += bytes([k - i]) + data[i:k + 1]
        i = k + 1
    return bytes(out)


def packbits_decode(data: bytes, want: int) -> bytes:
    out, i = bytearray(), 0
    while len(out) < want and i < len(data):
        h = data[i] - 256 if data[i] > 127 else data[i]
        i += 1
        if h >= 0:
            out += data[i:i + h + 1]
            i += h + 1
        elif h != -128:
            out += bytes([data[i]]) * (1 - h)
            i += 1
    return bytes(out)


def lzw_encode(data: bytes) -> bytes:
    """TIFF LZW as libtiff writes it: MSB-first codes, Clear first, the width
    grows once entry 2^w - 1 has been assigned (the decoder, one entry
    behind, grows after entry 2^w - 2), Clear again at entry 4094."""
    bits = []

    def put(code, width):
        bits.extend((code >> (width - 1 - k)) & 1 for k in range(width))

    table = {bytes([c]): c for c in range(256)}
    nxt, width = 258, 9
    put(256, width)
    w = b""
    for ch in data:
        wc = w + bytes([ch])
        if wc in table:
            w = wc
            continue
        put(table[w], width)
        table[wc] = nxt
        nxt += 1
        if nxt >= (1 << width) and width < 12:
            width += 1
        if nxt >= 4094:
            put(256, width)
            table = {bytes([c]): c for c in range(256)}
            nxt, width = 258, 9
        w = bytes([ch])
    if w:
        put(table[w], width)
        nxt += 1
        if nxt >= (1 << width) and width < 12:
            width += 1
    put(257, width)
    while len(bits) % 8:
        bits.append(0)
    return bytes(int("".join(map(str, bits[i:i + 8])), 2) for i in range(0, len(bits), 8))


def lzw_decode(data: bytes, want: int) -> bytes:
    out = bytearray()
    pos, nbits = 0, 8 * len(data)
    table = [bytes([c]) for c in range(256)] + [b"", b""]
    width, prev = 9, None
    while len(out) < want and pos + width <= nbits:
        code = 0
        for k in range(width):
            b = pos + k
            code = code << 1 | ((data[b >> 3] >> (7 - (b & 7))) & 1)
        pos += width
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width, prev = 9, None
            continue
        if prev is None:
            out += table[code]
            prev = code
            continue
        entry = table[code] if code < len(table) else table[prev] + table[prev][:1]
        out += entry
        table.append(table[prev] + entry[:1])
        prev = code
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out)


def write_tiff(path, pages, compression=1, predictor=1, order="<", rows_per_strip=None, extra=()):
    """Baseline grey TIFF, one page per 2-D array (uint8/16/32), in strips."""
    pages = [np.asarray(p) for p in pages]
    bps = pages[0].dtype.itemsize * 8
    blobs = []
    for p in pages:
        h, w = p.shape
        rps = rows_per_strip or h
        strips = []
        for y0 in range(0, h, rps):
            rows = p[y0:y0 + rps].astype(p.dtype.newbyteorder(order), copy=True)
            if predictor == 2:
                a = rows.astype(np.int64)
                d = a.copy()
                d[:, 1:] = a[:, 1:] - a[:, :-1]
                rows = (d % (1 << bps)).astype(p.dtype).astype(p.dtype.newbyteorder(order))
            raw = rows.tobytes()
            if compression == 32773:
                raw = packbits_encode(raw)
            elif compression == 5:
                raw = lzw_encode(raw)
            strips.append(raw)
        blobs.append((p.shape, rps, strips))
    out = bytearray(b"II*\0" if order == "<" else b"MM\0*")
    out += b"\0\0\0\0"
    prev_link = 4
    for (h, w), rps, strips in blobs:
        offs = []
        for s in strips:
            offs.append(len(out))
            out += s
            if len(out) % 2:
                out += b"\0"
        tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bps]), (259, 3, [compression]), (262, 3, [1]),
                (273, 4, offs), (277, 3, [1]), (278, 4, [rps]), (279, 4, [len(s) for s in strips]),
                (284, 3, [1])]
        if predictor != 1:
            tags.append((317, 3, [predictor]))
        tags += list(extra)
        tags.sort()
        arrays = {}
        for tag, typ, vals in tags:
            size = 2 if typ == 3 else 4
            if size * len(vals) > 4:
                arrays[tag] = len(out)
                out += struct.pack(order + ("H" if typ == 3 else "I") * len(vals), *vals)
                if len(out) % 2:
                    out += b"\0"
        ifd = len(out)
        struct.pack_into(order + "I", out, prev_link, ifd)
        out += struct.pack(order + "H", len(tags))
        for tag, typ, vals in tags:
            if tag in arrays:
                out += struct.pack(order + "HHII", tag, typ, len(vals), arrays[tag])
            elif typ == 3:
                v = list(vals) + [0] * (2 - len(vals))
                out += struct.pack(order + "HHIHH", tag, typ, len(vals), *v)
            else:
                out += struct.pack(order + "HHII", tag, typ, len(vals), vals[0])
        prev_link = len(out)
        out += b"\0\0\0\0"
    Path(path).write_bytes(bytes(out))


# ---------------------------------------------------------------------------
# TIFF
# ---------------------------------------------------------------------------
def test_packbits_specification_example():
    """TIFF 6.0 section 9 example: packed and unpacked bytes."""
    packed = bytes.fromhex("FE AA 02 80 00 2A FD AA 03 80 00 2A 22 F7 AA".replace(" ", ""))
    unpacked = bytes.fromhex(("AA AA AA 80 00 2A AA AA AA AA 80 00 2A 22" + " AA" * 10).replace(" ", ""))
    assert packbits_decode(packed, len(unpacked)) == unpacked
    assert packbits_decode(packbits_encode(unpacked), len(unpacked)) == unpacked
    d = np.frombuffer(unpacked, np.uint8).reshape(4, 6)
    assert d.size == len(unpacked)


def test_lzw_encoder_and_decoder_agree():
    rng = np.random.default_rng(1)
    for data in (bytes(rng.integers(0, 256, 5000, dtype=np.uint8)), b"ab" * 3000,
                 bytes(rng.integers(0, 4, 20000, dtype=np.uint8)), b"x" * 9000):
        assert lzw_decode(lzw_encode(data), len(data)) == data


@pytest.mark.parametrize("compression", [1, 32773, 5])
@pytest.mark.parametrize("dtype,order,predictor", [(np.uint8, "<", 1), (np.uint8, "<", 2), (np.uint16, "<", 1),
                                                   (np.uint16, ">", 2), (np.uint16, ">", 1),
                                                   (np.uint32, "<", 1)])
def test_tiff_round_trip(asc, tmp_path, compression, dtype, order, predictor):
    rng = np.random.default_rng(7)
    info = np.iinfo(dtype)
    top = min(info.max, 60000)
    pages = []
    for k in range(3):  # smooth and random parts, so every code path runs
        yy, xx = np.mgrid[0:37, 0:29]
        smooth = ((xx * 7 + yy * 3 + k * 11) % 50).astype(np.int64)
        noise = rng.integers(0, top, (37, 29))
        pages.append(np.where(xx < 15, smooth, noise).astype(dtype))
    path = tmp_path / "t.tif"
    write_tiff(path, pages, compression=compression, predictor=predictor, order=order, rows_per_strip=5)
    assert asc.io.tiff_info(path) == (29, 37, 3, 8 * np.dtype(dtype).itemsize)
    got = asc.io.read_tiff(path)
    assert got.dtype == np.dtype(dtype) and np.array_equal(got, np.stack(pages))


def test_tiff_errors(asc, tmp_path):
    p = tmp_path / "t.tif"
    write_tiff(p, [np.zeros((4, 4), np.uint8)], extra=[(322, 3, [16])])  # a tile width: tiled image
    with pytest.raises(asc.AscError, match="Tiled"):
        asc.io.read_tiff(p)
    write_tiff(p, [np.zeros((4, 4), np.uint8)], compression=8)
    with pytest.raises(asc.AscError, match="compression"):
        asc.io.read_tiff(p)
    write_tiff(p, [np.zeros((4, 4), np.uint16)], extra=[(339, 3, [2])])  # signed samples
    with pytest.raises(asc.AscError, match="SampleFormat"):
        asc.io.read_tiff(p)
    write_tiff(p, [np.zeros((4, 4), np.uint8), np.zeros((5, 4), np.uint8)])
    with pytest.raises(asc.AscError, match="Non uniform"):
        asc.io.read_tiff(p)
    good = tmp_path / "g.tif"
    write_tiff(good, [np.arange(64, dtype=np.uint8).reshape(8, 8)], compression=5)
    b = good.read_bytes()
    (tmp_path / "cut.tif").write_bytes(b[:20])
    with pytest.raises(asc.AscError):
        asc.io.read_tiff(tmp_path / "cut.tif")
    with pytest.raises(asc.AscError, match="Unable to open"):
        asc.io.read_tiff(tmp_path / "missing.tif")


def _tiff_strips(path):
    """(width, height, bps, predictor, compression, rows per strip, strips) of page 0."""
    b = Path(path).read_bytes()
    o = "<" if b[:2] == b"II" else ">"
    ifd = struct.unpack_from(o + "I", b, 4)[0]
    n = struct.unpack_from(o + "H", b, ifd)[0]
    tags = {}
    for e in range(n):
        tag, typ, cnt, val = struct.unpack_from(o + "HHII", b, ifd + 2 + 12 * e)
        size = 2 if typ == 3 else 4
        if size * cnt <= 4:
            vals = list(struct.unpack_from(o + ("H" if typ == 3 else "I") * cnt, b, ifd + 2 + 12 * e + 8))
        else:
            vals = list(struct.unpack_from(o + ("H" if typ == 3 else "I") * cnt, b, val))
        tags[tag] = vals
    strips = [b[a:a + c] for a, c in zip(tags[273], tags[279])]
    return (tags[256][0], tags[257][0], tags[258][0], tags.get(317, [1])[0], tags[259][0], tags[278][0], strips)


def test_sample_volume_matches_independent_decoder(asc):
    path = GOLD / "Models" / "rad_dcm_cropped.tif"
    got = asc.io.read_tiff(path)
    assert got.shape == (128, 86, 68) and got.dtype == np.uint16
    # page 0 through the decoder written here; every page through its strip layout
    w, h, bps, pred, comp, rps, strips = _tiff_strips(path)
    assert (comp, bps, pred) == (32773, 16, 1)
    raw = b"".join(packbits_decode(s, min(rps, h - k * rps) * w * 2) for k, s in enumerate(strips))
    assert np.array_equal(np.frombuffer(raw, "<u2").reshape(h, w), got[0])
    assert got.max() > 1000 and (got > 0).mean() > 0.05  # a segmented bone in a zero background


def test_sample_radiograph_matches_independent_decoder(asc):
    path = GOLD / "XMA_UND" / "xr_data_wrist_cam01" / "xr_rad_uln_cam01_UND.0099.tif"
    got = asc.io.read_tiff(path)
    assert got.shape == (1, 1760, 1760) and got.dtype == np.uint8
    w, h, bps, pred, comp, rps, strips = _tiff_strips(path)
    assert (comp, bps, pred, w, h) == (5, 8, 2, 1760, 1760)
    rows = []
    for s in strips[:50]:  # the first 200 rows
        r = np.frombuffer(lzw_decode(s, rps * w), np.uint8).reshape(rps, w).astype(np.int64)
        rows.append((np.cumsum(r, axis=1) % 256).astype(np.uint8))  # undo predictor 2
    ref = np.concatenate(rows)
    assert np.array_equal(got[0, :ref.shape[0]], ref)


# ---------------------------------------------------------------------------
# Cameras
# ---------------------------------------------------------------------------
def _film_rect(cam):
    """The radiograph's film rectangle as calculate_viewport derives it from
    image-plane corners 3 and 9 (Tracker.cpp:923-1032)."""
    a = cam.to_array()
    R, t, ip = a[:9].reshape(3, 3), a[9:12], a[12:24].reshape(4, 3)
    out = []
    for corner in (ip[1], ip[3]):
        q = R @ (corner - t)  # camera coordinates: rows of R^T are the axis columns
        out += [-2 * q[0] / q[2], -2 * q[1] / q[2]]
    return np.array(out)


@pytest.mark.parametrize("cam", [1, 2])
def test_mayacam2_sample_matches_oracle(asc, orc, cam):
    path = GOLD / "Calibration" / f"xr_calib_wrist_cam0{cam}.txt"
    got = asc.io.read_camera(path).to_array()
    lines = path.read_text().splitlines()
    size = [float(v) for v in lines[1].split(",")]
    K = [float(v) for ln in lines[4:7] for v in ln.split(",")]
    R = [float(v) for ln in lines[9:12] for v in ln.split(",")]
    t = [float(lines[k]) for k in (14, 15, 16)]
    ref = orc.camera_from_mayacam2(np.array(K).reshape(3, 3), np.array(R).reshape(3, 3), np.array(t), size)
    assert np.array_equal(got, ref)
    vp = got[24:28]
    rect = _film_rect(asc.io.read_camera(path))
    # fx != fy by a few 1e-4: the image plane uses their mean
    assert np.allclose(rect, [vp[0] - vp[2] / 2, vp[1] - vp[3] / 2, vp[0] + vp[2] / 2, vp[1] + vp[3] / 2],
                       rtol=0, atol=2e-3 * abs(vp[2]))


MC1_DOC = "-737.5,-514.9,94.3\n89.7,-2.0,-55.2\n3.9,4.1,-632.5\n840.2,839.2,-6325.8\n0.1,1760,1760\n"


def test_mayacam1_published_example(asc, orc, tmp_path):
    p = tmp_path / "cam.csv"
    p.write_text(MC1_DOC)
    c = asc.io.read_camera(p)
    a = c.to_array()
    # frame: location and CoordFrame::from_xyzypr(x, y, z, rz, ry, rx)
    f = orc.frame_from_xyzypr([-737.5, -514.9, 94.3, -55.2, -2.0, 89.7])
    assert np.array_equal(a[:12], f)
    # viewport (camera-calibration.md): c = (u0, v0) = file values - 1, f = |Z|
    u0, v0, z, W, H = 840.2 - 1, 839.2 - 1, 6325.8, 1760.0, 1760.0
    assert np.allclose(a[24:28], [-(2 * u0 - W) / z, -(2 * v0 - H) / z, 2 * W / z, 2 * H / z], rtol=1e-15)
    assert np.array_equal(a[28:30], [W, H])
    vp = a[24:28]
    assert np.allclose(_film_rect(c), [vp[0] - vp[2] / 2, vp[1] - vp[3] / 2, vp[0] + vp[2] / 2, vp[1] + vp[3] / 2],
                       rtol=1e-12, atol=1e-12)
    # the image plane sits 1.5 camera distances beyond the camera, centred on
    # the point the principal ray meets at pixel (W/2, H/2) offset
    R, t = a[:9].reshape(3, 3), a[9:12]
    centre = a[12:24].reshape(4, 3).mean(0)
    q = R @ (centre - t)
    assert np.isclose(q[2], -1.5 * np.linalg.norm(t), rtol=1e-12)
    # image size 0 reads as 1024
    p.write_text(MC1_DOC.replace("0.1,1760,1760", "0.1,0,0"))
    assert np.array_equal(asc.io.read_camera(p).to_array()[28:30], [1024, 1024])
    p.write_text(MC1_DOC.replace("89.7,-2.0,-55.2", "89.7,-2.0"))
    with pytest.raises(asc.AscError, match="MayaCam 1.0"):
        asc.io.read_camera(p)


def test_vtkcam_identities(asc, tmp_path):
    import json
    doc = {"@schema": "https://autoscoperm.slicer.org/vtkCamera-schema-1.0.json", "version": 1.0,
           "image-width": 1000, "image-height": 800, "camera-position": [400.0, -300.0, 250.0],
           "focal-point": [10.0, 20.0, -5.0], "view-up": [0.0, 0.0, 1.0], "view-angle": 30.0}
    p = tmp_path / "cam.json"
    p.write_text(json.dumps(doc, indent=2))
    c = asc.io.read_camera(p)
    a = c.to_array()
    R = a[:9].reshape(3, 3)  # rows: x, y, z axes
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
    fwd = np.array(doc["focal-point"]) - np.array(doc["camera-position"])
    assert np.allclose(-R[2], fwd / np.linalg.norm(fwd), atol=1e-14)  # looks down -z
    assert abs(R[0] @ np.array(doc["view-up"])) < 1e-14  # x axis is horizontal
    assert R[1] @ np.array(doc["view-up"]) > 0
    assert np.array_equal(a[9:12], doc["camera-position"])
    # principal point at the image centre; f_x = W / (2 tan(angle / 2)) and
    # f_y = H / (2 tan(angle / 2)) (camera-calibration.md), so the film
    # spans 4 tan(angle / 2) on both axes at z = -2
    assert np.allclose(a[24:28], [0, 0, 4 * math.tan(math.radians(15.0)), 4 * math.tan(math.radians(15.0))],
                       atol=1e-15)
    for bad in ({"version": 2.0}, {"@schema": "x"}):
        p.write_text(json.dumps({**doc, **bad}))
        with pytest.raises(asc.AscError, match="VTKCam"):
            asc.io.read_camera(p)


# ---------------------------------------------------------------------------
# Filter settings (.vie)
# ---------------------------------------------------------------------------
def _filters(fs):
    return [(f.type, [round(float(x), 6) for x in f.p[:3]]) for f in fs]


def test_vie_sample_files(asc):
    s = asc.io.read_filter_settings(GOLD / "xParameters" / "control_settings.vie")
    # rad: Sobel(blend 0.43, scale 3) -> Contrast(alpha 2.5, beta 0.9); DRR
    # step exp(7*0.62-5) = 0.5169 mm, intensity exp(15*0.35-5) = 1.284, cutoff 0;
    # DRR filter Sobel(blend 0.1, scale 1.7)
    assert _filters(s["rad_filters"]) == [(asc.FILTER_SOBEL, [3.0, 0.43, 0.0]), (asc.FILTER_CONTRAST, [2.5, 0.9, 3.0])]
    assert _filters(s["drr_filters"]) == [(asc.FILTER_SOBEL, [1.7, 0.1, 0.0])]
    assert s["drr_step"] == np.float32(math.exp(7 * 0.62 - 5)) and round(s["drr_step"], 4) == 0.5169
    assert s["drr_intensity"] == np.float32(math.exp(15 * 0.35 - 5)) and round(s["drr_intensity"], 3) == 1.284
    assert s["cutoff_fraction"] == 0.0 and s["sections"] == (True, True, True)
    e = asc.io.read_filter_settings(GOLD / "xParameters" / "empty_filter.vie")
    assert e["rad_filters"] == [] and e["drr_filters"] == []
    assert round(e["drr_intensity"], 2) == 10.49  # RI 0.49


def test_vie_spinbox_bounds_and_rounding(asc, tmp_path):
    p = tmp_path / "s.vie"
    p.write_text("\n".join([
        "DrrRenderer_begin", "SampleDistance 0.004", "RayIntensity 0.125", "Cutoff 0.994", "DrrRenderer_end",
        "DrrFilters_begin",
        "SobelFilter_begin", "Scale 1.7", "Blend 12", "SobelFilter_end",  # blend clamps to 1, scale to 10
        "GaussianFilter_begin", "Radius 2.345", "GaussianFilter_end",
        "DrrFilters_end",
        "RadFilters_begin",
        "ContrastFilter_begin", "Alpha 0.625", "ContrastFilter_end",  # beta keeps its default 1
        "SharpenFilter_begin", "Radius 3", "Contrast 0.333", "SharpenFilter_end",
        "SobelFilter_begin", "SobelFilter_end",  # defaults: blend 0.5, scale 1
        "RadFilters_end", ""]))
    s = asc.io.read_filter_settings(p)
    assert s["sample_distance"] == 0.01 and s["ray_intensity"] == 0.13 and s["cutoff_fraction"] == 0.99
    assert _filters(s["drr_filters"]) == [(asc.FILTER_SOBEL, [10.0, 1.0, 0.0]), (asc.FILTER_GAUSSIAN, [2.35, 0, 0])]
    assert _filters(s["rad_filters"]) == [(asc.FILTER_CONTRAST, [0.63, 1.0, 3.0]),
                                          (asc.FILTER_SHARPEN, [3.0, 0.33, 0.0]),
                                          (asc.FILTER_SOBEL, [1.0, 0.5, 0.0])]
    assert s["drr_step"] == np.float32(math.exp(7 * 0.01 - 5))
    desc = asc.AscVolumeDesc()
    desc.min_value, desc.max_value = 0.25, 0.75
    assert asc.io.drr_cutoff(0.5, desc) == np.float32(0.5)


# ---------------------------------------------------------------------------
# Trial configuration (.cfg)
# ---------------------------------------------------------------------------
def test_golden_wrist_trial(asc):
    t = asc.Trial(GOLD / "wrist.cfg")
    assert t.version == (1, 1) and (t.n_cameras, t.n_volumes, t.n_frames) == (2, 4, 1)
    assert t.render == (512, 512) and np.array_equal(t.offsets, [0.1] * 6)
    for i in range(2):
        cam = asc.io.read_camera(GOLD / "Calibration" / f"xr_calib_wrist_cam0{i + 1}.txt")
        assert np.array_equal(t.cameras[i].to_array(), cam.to_array())
    names = ["rad", "mc2_mc3", "mc3", "uln"]
    for i, n in enumerate(names):
        v = t.volume(i)
        assert np.array_equal(v["data"], asc.io.read_tiff(GOLD / "Models" / f"{n}_dcm_cropped.tif"))
        assert np.array_equal(v["voxel_mm"], np.float32([0.39625, 0.39625, 0.625]).astype(np.float64))
        assert v["flip"] == (0, 0, 0)
    f = t.frame(1, 0)
    assert f.shape == (1760, 1760) and f.dtype == np.uint8
    assert t.frame_path(0, 5).endswith("xr_rad_uln_cam01_UND.0099.tif")  # past the end: the last frame
    t.close()


def _trial_dir(tmp_path, asc, nframes=3):
    root = tmp_path / "trial"
    (root / "cams").mkdir(parents=True)
    (root / "cams" / "c1.txt").write_text((GOLD / "Calibration" / "xr_calib_wrist_cam01.txt").read_text())
    (root / "vol").mkdir()
    rng = np.random.default_rng(3)
    vol = rng.integers(0, 4000, (6, 5, 4)).astype(np.uint16)
    write_tiff(root / "vol" / "v.tif", list(vol), compression=32773)
    (root / "frames").mkdir()
    frames = []
    for k in range(nframes):
        fr = rng.integers(0, 255, (6, 7)).astype(np.uint8)
        frames.append(fr)
        write_tiff(root / "frames" / f"f{k:03d}.tif{'f' if k == 1 else ''}", [fr], compression=5, predictor=2)
    (root / "frames" / ".hidden.tif").write_bytes(b"")
    (root / "frames" / "notes.txt").write_text("not a frame")
    return root, vol, frames


def test_trial_paths_defaults_and_frames(asc, tmp_path):
    root, vol, frames = _trial_dir(tmp_path, asc)
    cfg = root / "t.cfg"
    cfg.write_text("# comment\n\nVersion 1.2\nmayaCam_csv cams\\c1.txt\nCameraRootDir frames//\n"
                   "VolumeFile vol/v.tif\nVoxelSize 0.5 0.25 1\nVolumeFlip 1 0 1\n")
    t = asc.Trial(cfg)
    assert t.version == (1, 2) and t.render == (512, 512) and np.array_equal(t.offsets, [0.1] * 6)
    v = t.volume(0)
    assert np.array_equal(v["data"], vol) and v["flip"] == (1, 0, 1)
    assert np.array_equal(v["voxel_mm"], [0.5, 0.25, 1.0])
    assert t.n_frames == 3
    for k in range(3):  # f000.tif, f001.tiff, f002.tif in order; hidden and other files skipped
        assert np.array_equal(t.frame(0, k), frames[k])
    assert np.array_equal(t.frame(0, 9), frames[2])
    bg = t.background(0)
    assert np.array_equal(bg, np.max([f.astype(np.float32) / np.float32(255) for f in frames], 0))
    t.close()
    # version 1.0: paths are taken as given (absolute here)
    cfg.write_text(f"mayaCam_csv {root}/cams/c1.txt\nCameraRootDir {root}/frames\nVolumeFile {root}/vol/v.tif\n"
                   "VoxelSize 1 1 1\nRenderResolution 256 128\nOptimizationOffsets 1 2 3 4 5 6\n")
    t = asc.Trial(cfg)
    assert t.version == (1, 0) and t.render == (256, 128) and np.array_equal(t.offsets, [1, 2, 3, 4, 5, 6])
    t.close()


@pytest.mark.parametrize("text,msg", [
    ("VolumeFile vol/v.tif\nVoxelSize 1 1 1\n", "at least one mayacam"),
    ("mayaCam_csv cams/c1.txt\nVolumeFile vol/v.tif\nVoxelSize 1 1 1\n", "cameras and videos must match"),
    ("mayaCam_csv cams/c1.txt\nCameraRootDir frames\n", "at least one volume"),
    ("mayaCam_csv cams/c1.txt\nCameraRootDir frames\nVolumeFile vol/v.tif\n", "voxel sizes"),
    ("mayaCam_csv cams/c1.txt\nCameraRootDir frames\nVolumeFile vol/v.tif\nVoxelSize 1 1 1\nMeshFile a.stl\n"
     "MeshFile b.stl\n", "mesh file"),
    ("mayaCam_csv cams/c1.txt\nCameraRootDir nowhere\nVolumeFile vol/v.tif\nVoxelSize 1 1 1\n", "Unable to open"),
])
def test_trial_validation(asc, tmp_path, text, msg):
    root, _, _ = _trial_dir(tmp_path, asc, 1)
    cfg = root / "bad.cfg"
    cfg.write_text("Version 1.1\n" + text)
    with pytest.raises(asc.AscError, match=msg):
        asc.Trial(cfg)


# ---------------------------------------------------------------------------
# Tracking results (.tra)
# ---------------------------------------------------------------------------
def test_tra_text_and_round_trips(asc, orc, tmp_path):
    rng = np.random.default_rng(5)
    poses = np.concatenate([rng.uniform(-50, 50, (4, 2, 3)), rng.uniform(-80, 80, (4, 2, 3))], 2)
    poses[2, 1] = np.nan  # an untracked frame of volume 1
    p = tmp_path / "r.tra"
    asc.io.write_tra(p, poses[:1, :1], flags=asc.io.TRA_COMMAS)
    assert p.read_text() == ",".join(f"{v:.16f}" for v in poses[0, 0]) + "\n"
    for flags in (0, 4, 4 | 8, 4 | 16, 4 | 8 | 16):
        asc.io.write_tra(p, poses, flags=flags)
        back = asc.io.read_tra(p, 4, 2, flags=flags)
        assert np.isnan(back[2, 1]).all() and np.allclose(back[~np.isnan(poses)], poses[~np.isnan(poses)],
                                                          rtol=1e-13, atol=1e-12)
    line = (tmp_path / "r.tra").read_text().splitlines()[2]
    assert line.count("NaN") == 6
    # matrices: CoordFrame::to_matrix, columns (or rows) of the 4x4
    for flags in (1 | 4, 1 | 2 | 4, 1 | 2 | 4 | 8):
        asc.io.write_tra(p, poses, flags=flags)
        vals = [float(v) for v in p.read_text().splitlines()[0].split(",")[:16]]
        f = orc.frame_from_xyzypr(poses[0, 0])
        col = np.zeros(16)
        col[[0, 1, 2, 4, 5, 6, 8, 9, 10]] = f[:9]
        col[12:15] = f[9:12] / (10.0 if flags & 8 else 1.0)
        col[15] = 1.0
        expect = col.reshape(4, 4).T.ravel() if flags & 2 else col
        assert np.allclose(vals, expect, rtol=0, atol=1e-15 + 1e-16 * np.abs(expect).max())
        back = asc.io.read_tra(p, 4, 2, flags=flags)
        ok = ~np.isnan(poses)
        assert np.allclose(back[ok], poses[ok], rtol=1e-10, atol=1e-9)
    # fewer lines than frames: the rest stays NaN
    p.write_text("1,2,3,4,5,6\n")
    back = asc.io.read_tra(p, 3, 1, flags=4)
    assert np.array_equal(back[0, 0], [1, 2, 3, 4, 5, 6]) and np.isnan(back[1:]).all()
