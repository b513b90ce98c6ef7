"""The remote-control session's device messages on the wrist sample trial
(tests/golden/sample_data), in the wire format PyAutoscoper sends:

  * getNCC (8): the per-camera 1 - NCC of a pose equals the fused scene's
    for the same trial, filters and frame, bit for bit;
  * optimizeFrame (11) with Autoscoper_RANDOM_SEED set: the tracked pose
    equals asc_optimize_frame's (reference-order PSO, then the simplex)
    from the same start and seed;
  * getImageCropped (10): the three 8-bit images of the whole view (the
    radiograph, the DRR and the mask, Tracker::getImageData) against the
    oracle's pipeline: equal but for a truncation step in a few pixels
    (DRR values agree within 1e-4 relative);
  * setBackground (9) answers, and costs stay defined.
"""
import os
import struct
from pathlib import Path

import numpy as np
import pytest

from test_gpu_sample_data import base_pose, build_scene, wrist  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


def msg(t, fmt="", *vals, tail=b""):
    return bytes([t]) + struct.pack("<" + fmt, *vals) + tail


@pytest.fixture()
def remote(asc, ctx):
    r = asc.Remote(ctx)
    assert r.handle(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    vie = str(GOLD / "xParameters" / "control_settings.vie").encode()
    assert r.handle(msg(4, "i", -1, tail=vie)) == b"\x04"
    assert r.handle(msg(5, "i", 0)) == b"\x05"
    yield r
    r.close()


def test_ncc_matches_scene(asc, ctx, wrist, remote):
    sc, vols, _, _ = build_scene(asc, ctx, wrist)
    rng = np.random.default_rng(8)
    for b in (0, 3):
        for _ in range(3):
            p = base_pose(wrist, b) + np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-5, 5, 3)])
            rep = remote.handle(msg(8, "i6d", b, *p))
            assert rep[0] == 8 and rep[1] == 2
            got = struct.unpack("<2d", rep[2:])
            _, per = sc.eval_poses(p[None], volumes=[b])
            assert np.array_equal(got, per[0])
    sc.close()
    for v in vols:
        v.close()


def test_optimize_matches_optimize_frame(asc, ctx, wrist, remote):
    sc, vols, _, _ = build_scene(asc, ctx, wrist)
    start = base_pose(wrist, 0) + np.array([0.8, -0.5, 0.4, 1.5, -1.0, 0.8])
    assert remote.handle(msg(7, "ii6d", 0, 0, *start)) == b"\x07"
    os.environ["Autoscoper_RANDOM_SEED"] = "5"
    try:
        # volume 0, frame 0, 1 repeat, 6 epochs, limits -3/3, stall 3, dframe 1,
        # PSO + simplex, NCC, current-frame guess
        rep = remote.handle(msg(11, "iiiiddiiiii", 0, 0, 1, 6, -3.0, 3.0, 3, 1, 0, 0, 0))
    finally:
        del os.environ["Autoscoper_RANDOM_SEED"]
    assert rep == b"\x0b"
    got = struct.unpack("<6d", remote.handle(msg(6, "ii", 0, 0))[1:])
    store0 = asc.geometry.pose_store_set(start)
    s, st = sc.optimize_frame(store0, method=0, max_iter=6, max_stall=3, min_limit=-3.0, max_limit=3.0,
                              offsets=(0.1,) * 6, seed=5, particles=100, synchronous=False, volume=0)
    ref = asc.geometry.pose_store_get(s)
    print(f"remote optimize: {np.round(got, 4)} ({st.evaluations} evaluations)")
    assert np.array_equal(got, ref)
    # the SAD cost of implant models is not built: the request fails
    assert remote.handle(msg(11, "iiiiddiiiii", 0, 0, 1, 6, -3.0, 3.0, 3, 1, 0, 1, 0)) == b"\x00"
    sc.close()
    for v in vols:
        v.close()


def test_image_matches_oracle(asc, ctx, orc, wrist, remote):
    b, c = 0, 1
    pose = base_pose(wrist, b)
    rep = remote.handle(msg(10, "ii6d", b, c, *pose))
    assert rep[0] == 10
    w, h = struct.unpack("<ii", rep[1:9])
    assert (w, h) == tuple(wrist["render"])
    imgs = np.frombuffer(rep[9:], np.uint8).reshape(3, h, w)
    # the oracle's pipeline over the whole view
    s = wrist["settings"]
    prep, cam = wrist["preps"][b], wrist["cams"][c]
    _, imv, _, _ = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], w, h)
    vp = np.float32([-cam[26] / 2, -cam[27] / 2, cam[26], cam[27]])
    v = wrist["vols"][b]
    vol = asc.Volume(ctx, v["data"], scale=tuple(v["voxel_mm"]), flip=v["flip"])
    cutoff = asc.io.drr_cutoff(s["cutoff_fraction"], vol.desc)
    vol.close()
    drr, _ = orc.drr(prep["volume"], imv[:12], vp, s["drr_step"], s["drr_intensity"], cutoff, w, h)
    drr = orc.filter_chain(asc.filters.to_records(s["drr_filters"]), drr)
    rad = orc.rad(wrist["frames"][c], cam[24:28], vp, w, h)
    rad = orc.filter_chain(asc.filters.to_records(s["rad_filters"]), rad)
    m = (drr != 0).astype(np.float32)
    ref = [rad * m, drr * m, m]
    for k in range(3):
        r8 = (np.trunc(255 * ref[k].astype(np.float64)).astype(np.int64) & 0xFF).astype(np.uint8)
        diff = imgs[k].astype(np.int64) - r8
        print(f"image {k}: {np.count_nonzero(diff)} of {diff.size} pixels differ")
        assert (np.abs(diff) <= 1).mean() >= 0.999 and np.count_nonzero(diff) <= 0.002 * diff.size
    assert m.mean() > 0.001


def test_background_threshold(asc, ctx, wrist, remote, tmp_path):
    """With one frame per camera no background image can be made: the
    session answers and keeps costs unmasked, as the GUI does. With two
    frames the background mask (threshold <= the frames' maximum) applies."""
    pose = base_pose(wrist, 0)
    before = remote.handle(msg(8, "i6d", 0, *pose))
    assert remote.handle(msg(9, "d", 0.1)) == b"\x09"
    assert remote.handle(msg(8, "i6d", 0, *pose)) == before
    # the same trial with every frame listed twice
    import shutil
    root = tmp_path / "w2"
    shutil.copytree(GOLD, root)
    for cam in ("cam01", "cam02"):
        d = root / "XMA_UND" / f"xr_data_wrist_{cam}"
        f = next(d.iterdir())
        shutil.copyfile(f, d / (f.stem + "b.tif"))
    r2 = asc.Remote(ctx)
    assert r2.handle(msg(1, tail=str(root / "wrist.cfg").encode())) == b"\x01"
    assert r2.handle(msg(4, "i", -1, tail=str(root / "xParameters" / "control_settings.vie").encode())) == b"\x04"
    unmasked = r2.handle(msg(8, "i6d", 0, *pose))
    assert unmasked == before
    assert r2.handle(msg(9, "d", 0.6)) == b"\x09"
    masked = r2.handle(msg(8, "i6d", 0, *pose))
    assert masked[0] == 8 and masked != unmasked
    r2.close()


def _open(asc, ctx, server=None):
    r = asc.Remote(ctx, server=server)
    assert r.handle(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    assert r.handle(msg(4, "i", -1, tail=str(GOLD / "xParameters" / "control_settings.vie").encode())) == b"\x04"
    assert r.handle(msg(5, "i", 0)) == b"\x05"
    return r


def _track(r, volume, start):
    """getNCC at the start, optimizeFrame (PSO 6 epochs + simplex, seed from
    the environment), then the tracked pose and its getNCC."""
    assert r.handle(msg(7, "ii6d", volume, 0, *start)) == b"\x07"
    before = r.handle(msg(8, "i6d", volume, *start))
    rep = r.handle(msg(11, "iiiiddiiiii", volume, 0, 1, 6, -3.0, 3.0, 3, 1, 0, 0, 0))
    assert rep == b"\x0b"
    pose = struct.unpack("<6d", r.handle(msg(6, "ii", volume, 0))[1:])
    return before, pose, r.handle(msg(8, "i6d", volume, *pose))


def test_sessions_share_the_servers_launches(asc, ctx, wrist):
    """Two socket sessions on one registration server (asc_remote_create_shared)
    optimise different bones at the same time: their messages ride in shared
    fused launches, and every answer is the same bits as a session of its own."""
    import threading

    starts = {0: base_pose(wrist, 0) + np.array([0.8, -0.5, 0.4, 1.5, -1.0, 0.8]),
              3: base_pose(wrist, 3) + np.array([-0.6, 0.7, -0.3, -1.0, 1.2, 0.5])}
    os.environ["Autoscoper_RANDOM_SEED"] = "5"
    try:
        alone = {}
        for v, start in starts.items():
            r = _open(asc, ctx)
            alone[v] = _track(r, v, start)
            r.close()
        srv = asc.Server(ctx, 16 << 30, 4)
        sessions = {v: _open(asc, ctx, server=srv) for v in starts}
        shared, errors = {}, []

        def run(v):
            try:
                shared[v] = _track(sessions[v], v, starts[v])
            except Exception as e:  # reported below
                errors.append(e)

        threads = [threading.Thread(target=run, args=(v,)) for v in starts]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        st = srv.stats()
        for r in sessions.values():
            r.close()
        srv.close()
    finally:
        del os.environ["Autoscoper_RANDOM_SEED"]
    assert not errors, errors
    print("shared sessions:", {k: st[k] for k in ("batches", "requests", "evaluations")})
    for v in starts:
        assert shared[v] == alone[v], v
    assert st["requests"] > st["batches"]  # the two sessions' requests shared launches
