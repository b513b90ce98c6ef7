"""The display path (View::render over a camera's whole image plane) and
the remote session's saveFullDRR (message 12), against the oracle:

  * tiny phantom placed twice (two volume slots, no cutoff): the merged,
    filtered DRR channel within the bound that the 1e-4 DRR bar propagates
    to through the merge and the filters, the radiograph channel within
    1e-5 relative, and the middle channel equal to the Compositor's formula
    applied to the first and last, bit for bit;
  * the wrist sample trial, its four bones merged, with a cutoff: the same
    bound, where each bone's DRR bar includes the samples near the cutoff
    that a different sampler rounding may flip (the oracle's margin);
  * the PNG written from a composite decodes to its rounded, flipped rows;
  * saveFullDRR writes one PGM per camera, numbered across calls, equal to
    save_full_drr of the oracle's masked DRR but for a truncation step in
    a few pixels.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

from chain_bound import assert_within, chain_bound, drr_bar
from test_display import read_pgm, read_png
from test_gpu_sample_data import base_pose, build_scene, wrist  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


def full_view(cam, RW, RH):
    vp = np.float32([-cam[26] / 2, -cam[27] / 2, cam[26], cam[27]])
    w = int(np.float64(cam[26]) * RW / np.float64(cam[26]))
    h = int(np.float64(cam[27]) * RH / np.float64(cam[27]))
    return vp, w, h


def oracle_view(asc, orc, cam, preps, poses, params, frame, drr_recs, rad_recs, RW, RH):
    """-> (the oracle's composite, per-pixel bound on its DRR channel): each
    bone's DRR bar (1e-4 relative plus its cutoff margin) summed through the
    merge (min(a + b, 1) is 1-Lipschitz in each input), then carried
    through the DRR filter chain."""
    vp, w, h = full_view(cam, RW, RH)
    merged = np.zeros((h, w), np.float32)
    e = np.zeros((h, w))
    for prep, pose, (step, inten, cut) in zip(preps, poses, params):
        _, imv, _, _ = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], RW, RH)
        d, margin = orc.drr_cut_margin(prep["volume"], imv[:12], vp, step, inten, cut, w, h)
        merged = orc.merge(merged, d)
        e += drr_bar(d, margin)
    drr_f = orc.filter_chain(drr_recs, merged)
    rad_f = orc.filter_chain(rad_recs, orc.rad(frame, cam[24:28], vp, w, h))
    return orc.composite(drr_f, rad_f, np.ones_like(drr_f)), chain_bound(drr_recs, merged, e)


def check_compositor(g):
    """G = (R/2 + B/2) where the background mask is 1 (no background)."""
    want = g[..., 0] / np.float32(2) + g[..., 2] / np.float32(2)
    assert np.array_equal(g[..., 1], want)


def test_view_two_slots_matches_oracle(asc, ctx, orc, tiny, tmp_path):
    wl = tiny["wl"]
    RW, RH = tiny["render"]
    vol = asc.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3)
    sc = asc.Scene(ctx, RW, RH)
    params = [(0.37, 2.5, 0.0), (0.5, 1.5, 0.0)]
    for p in params:
        sc.add_volume(vol, *p)
    recs = []
    for cam, (_, rad) in zip(tiny["cams"], tiny["rads"]):
        c = sc.add_camera(asc.AscCamera.from_array(cam), asc.Image(ctx, rad))
        sc.set_filters(c, 0, wl.drr_filters)
        sc.set_filters(c, 1, wl.rad_filters)
        recs.append((asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)))
    poses = np.stack([wl.gt_poses[0], wl.gt_poses[0] + np.array([4.0, -3.0, 2.0, 10.0, -6.0, 4.0])])
    for c, cam in enumerate(tiny["cams"]):
        g = sc.render_view(c, poses)
        r, bound = oracle_view(asc, orc, cam, [tiny["prep"]] * 2, poses, params, tiny["rads"][c][1], *recs[c],
                               RW, RH)
        assert g.shape == r.shape
        peak = np.abs(r[..., 0]).max()
        err = np.abs(g[..., 0].astype(np.float64) - r[..., 0])
        print(f"camera {c}: DRR channel max error {err.max() / peak:.2e} of its peak {peak:.4f}")
        assert peak > 0
        assert_within(g[..., 0], r[..., 0], bound, f"camera {c} DRR channel")
        assert np.allclose(g[..., 2], r[..., 2], rtol=1e-5, atol=0)
        check_compositor(g)
        p = tmp_path / f"view{c}.png"
        asc.io.write_view_png(p, g)
        want = np.floor(np.clip(g, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)[::-1]
        assert np.array_equal(read_png(p), want)
    # one slot without a pose is refused
    with pytest.raises(asc.AscError):
        sc.render_view(0, poses[:1])
    sc.close()
    vol.close()


def test_wrist_view_matches_oracle(asc, ctx, orc, wrist):
    sc, vols, _, params = build_scene(asc, ctx, wrist)
    s = wrist["settings"]
    RW, RH = wrist["render"]
    poses = np.stack([base_pose(wrist, b) for b in range(len(vols))])
    recs = (asc.filters.to_records(s["drr_filters"]), asc.filters.to_records(s["rad_filters"]))
    for c, cam in enumerate(wrist["cams"]):
        g = sc.render_view(c, poses)
        r, bound = oracle_view(asc, orc, cam, wrist["preps"], poses, params, wrist["frames"][c], *recs, RW, RH)
        peak = np.abs(r[..., 0]).max()
        err = np.abs(g[..., 0].astype(np.float64) - r[..., 0])
        print(f"wrist camera {c}: max DRR-channel error {err.max():.3e}; "
              f"{(err <= 1e-4 * peak).mean() * 100:.3f} % of pixels within 1e-4 of the peak")
        assert peak > 0
        assert_within(g[..., 0], r[..., 0], bound, f"wrist camera {c} DRR channel")
        assert np.allclose(g[..., 2], r[..., 2], rtol=1e-5, atol=0)
        check_compositor(g)
    sc.close()
    for v in vols:
        v.close()


def msg(t, fmt="", *vals, tail=b""):
    return bytes([t]) + struct.pack("<" + fmt, *vals) + tail


def test_save_full_drr(asc, ctx, orc, wrist, tmp_path):
    r = asc.Remote(ctx)
    r.set_output_dir(tmp_path)
    assert r.handle(msg(1, tail=str(GOLD / "wrist.cfg").encode())) == b"\x01"
    vie = str(GOLD / "xParameters" / "control_settings.vie").encode()
    assert r.handle(msg(4, "i", -1, tail=vie)) == b"\x04"
    assert r.handle(msg(7, "ii6d", 0, 0, *base_pose(wrist, 0))) == b"\x07"
    pose = np.array(struct.unpack("<6d", r.handle(msg(6, "ii", 0, 0))[1:]))  # the curves' value
    assert r.handle(msg(12)) == b"\x0c"
    assert r.handle(msg(12)) == b"\x0c"
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == [f"image_cam{i:02d}.pgm" for i in range(4)]
    s = wrist["settings"]
    prep = wrist["preps"][0]
    v = wrist["vols"][0]
    vol = asc.Volume(ctx, v["data"], scale=tuple(v["voxel_mm"]), flip=v["flip"])
    cutoff = asc.io.drr_cutoff(s["cutoff_fraction"], vol.desc)
    vol.close()
    RW, RH = wrist["render"]
    for c, cam in enumerate(wrist["cams"]):
        vp, w, h = full_view(cam, RW, RH)
        _, imv, _, _ = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], RW, RH)
        d, _ = orc.drr(prep["volume"], imv[:12], vp, s["drr_step"], s["drr_intensity"], cutoff, w, h)
        d = orc.filter_chain(asc.filters.to_records(s["drr_filters"]), d)
        d = d * (d != 0).astype(np.float32)  # the DRR mask and multiplies leave the DRR as it is
        ref = tmp_path / f"ref{c}.pgm"
        asc.io.write_full_drr_pgm(ref, d)
        want = read_pgm(ref)
        got = read_pgm(tmp_path / f"image_cam{c:02d}.pgm")
        assert np.array_equal(got, read_pgm(tmp_path / f"image_cam{c + 2:02d}.pgm"))
        diff = got - want
        print(f"full DRR camera {c}: {np.count_nonzero(diff)} of {diff.size} pixels differ")
        assert (np.abs(diff) <= 1).all() and np.count_nonzero(diff) <= 0.002 * diff.size
    r.close()
