"""Several bones in one scene: each bone's trackFrame uses its own volume's
DRR against the shared radiographs (Tracker::trackFrame(volumeID),
Tracker.cpp:583-773, View::renderDrrSingle), and one fused batch mixes
poses of all bones (items of uneven bounding boxes and ray lengths).

Targets are the merged full-frame DRRs of all bones (View::renderDrr +
gpu::merge: min(a + b, 1)), rendered by the oracle. Each item is compared
with the oracle's trackFrame for that bone: 1 - NCC within 1e-5, render
sizes, in-view flags, mask counts and ray-sample counts bit-exact.
"""
import numpy as np
import pytest

from autoscoper_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def multi(asc, orc):
    wl = synth.multi_workload("tiny3", frames=1)
    cams = [orc.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    preps = [orc.volume_prep(v, (vox,) * 3) for v, vox in zip(wl.volumes, wl.voxel_mm)]
    RW, RH = wl.render
    rads = []
    for i, cam in enumerate(cams):
        fvp = synth.full_frame_viewport(cam[24:28], RW, RH)
        merged = np.zeros((RH, RW), np.float32)
        for b, prep in enumerate(preps):
            _, imv, _, _ = orc.view_geometry(cam, wl.gt_poses[0, b], prep["inv_scale"], prep["inv_trans"], RW, RH)
            img, _ = orc.drr(prep["volume"], imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, RW, RH)
            merged = orc.merge(merged, img).reshape(RH, RW)
        rads.append(synth.target_from_drr(merged, 90 + i))
    return dict(wl=wl, cams=cams, preps=preps, rads=rads, render=(RW, RH))


def build(asc, ctx, multi):
    wl = multi["wl"]
    RW, RH = multi["render"]
    sc = asc.Scene(ctx, RW, RH)
    vols, pivots = [], []
    for v, vox in zip(wl.volumes, wl.voxel_mm):
        vol = asc.Volume(ctx, v, scale=(vox,) * 3)
        pivot = asc.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
        sc.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, pivot)
        vols.append(vol)
        pivots.append(pivot)
    for i, cam in enumerate(multi["cams"]):
        c = sc.add_camera(asc.AscCamera.from_array(cam), asc.Image(ctx, multi["rads"][i][1]))
        sc.set_filters(c, 0, wl.drr_filters)
        sc.set_filters(c, 1, wl.rad_filters)
    return sc, vols, pivots


def oracle_for(orc, asc, multi, b):
    wl = multi["wl"]
    recs = [(asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)) for _ in multi["cams"]]
    return orc.Scene(multi["preps"][b], (synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), multi["cams"],
                     [r[1] for r in multi["rads"]], recs, multi["render"])


def test_mixed_bone_batch_matches_oracle(asc, ctx, orc, multi):
    sc, vols, _ = build(asc, ctx, multi)
    wl = multi["wl"]
    oracles = [oracle_for(orc, asc, multi, b) for b in range(3)]
    rng = np.random.default_rng(41)
    bones = np.array([0, 1, 2, 2, 0, 1, 1, 2, 0])
    poses = wl.gt_poses[0, bones] + np.concatenate([rng.uniform(-2, 2, (9, 3)), rng.uniform(-4, 4, (9, 3))], 1)
    costs, per = sc.eval_poses(poses, volumes=bones)
    info, samples = sc.last_info(len(poses) * 2)
    tot = 0
    for i, (b, p) in enumerate(zip(bones, poses)):
        r = oracles[b].track_frame(p)
        for c in range(2):
            assert info[2 * i + c, 0] == r["info"][c, 0]
            if r["info"][c, 0] == 1:
                assert tuple(info[2 * i + c, 1:4]) == tuple(r["info"][c, 1:4])
                assert abs(per[i, c] - r["corr"][c]) <= 1e-5
        tot += r["samples"]
    assert samples == tot
    # a bone alone and in the mix: identical
    for b in range(3):
        sel = np.nonzero(bones == b)[0]
        cb, pb = sc.eval_poses(poses[sel], volumes=bones[sel])
        assert np.array_equal(cb, costs[sel]) and np.array_equal(pb, per[sel])
    sc.close()
    for v in vols:
        v.close()


def test_minimization_func_per_bone(asc, ctx, orc, multi):
    """Tracker::minimizationFunc for different current volumes in one batch:
    each bone's stored pose and pivot (Trial::getVolumeMatrix)."""
    sc, vols, pivots = build(asc, ctx, multi)
    wl = multi["wl"]
    oracles = [oracle_for(orc, asc, multi, b) for b in range(3)]
    rng = np.random.default_rng(5)
    stores = np.stack([asc.geometry.pose_store_set(wl.gt_poses[0, b] + np.array([0.4, -0.3, 0.2, 1.0, -0.8, 0.5]))
                       for b in range(3)])
    bones = np.array([2, 0, 1, 0, 2, 1])
    manips = np.concatenate([rng.uniform(-1, 1, (6, 3)), rng.uniform(-2, 2, (6, 3))], 1)
    costs, _ = sc.eval_manips(stores[bones], manips, volumes=bones)
    for i, b in enumerate(bones):
        assert abs(costs[i] - oracles[b].min_func(stores[b], pivots[b], manips[i])) <= 2e-5
    sc.close()
    for v in vols:
        v.close()
