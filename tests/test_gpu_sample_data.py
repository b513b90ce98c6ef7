"""The reference's wrist sample trial end to end (tests/golden/sample_data):
the trial file, two MayaCam 2.0 calibrations, four segmented 12-bit bone
volumes (radius, MC2-MC3, MC3, ulna) at 0.39625 x 0.39625 x 0.625 mm, one
1760 x 1760 8-bit radiograph per camera, the sample filter settings
(control_settings.vie) and a 512 x 512 render, all read through the
product's readers, and compared with the oracle on the same inputs:

  * per bone, trackFrame of poses around an in-view pose: 1 - NCC within
    1e-5, render sizes, in-view flags, mask counts and ray-sample counts
    bit-exact;
  * the ray march of the radius against the oracle's DRR: pixels within
    1e-4 relative, zero pattern exact;
  * Tracker::optimize (the downhill simplex from a perturbed start) on the
    GPU scene against the oracle's: the same final pose.

The sample data carries no tracked poses, so the bones are placed where
both cameras see them: at the point nearest to both optical axes.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "sample_data"


@pytest.fixture(scope="module")
def wrist(asc, orc):
    trial = asc.Trial(GOLD / "wrist.cfg")
    settings = asc.io.read_filter_settings(GOLD / "xParameters" / "control_settings.vie")
    cams = [c.to_array() for c in trial.cameras]
    # the point nearest to both optical axes (each camera looks down its -z)
    M, b = np.zeros((3, 3)), np.zeros(3)
    for c in cams:
        d = -c[6:9] / np.linalg.norm(c[6:9])
        P = np.eye(3) - np.outer(d, d)
        M += P
        b += P @ c[9:12]
    iso = np.linalg.solve(M, b)
    vols = [trial.volume(i) for i in range(trial.n_volumes)]
    preps = [orc.volume_prep(v["data"], tuple(v["voxel_mm"]), v["flip"]) for v in vols]
    frames = [trial.frame(c, 0) for c in range(trial.n_cameras)]
    trial.close()
    return dict(settings=settings, cams=cams, iso=iso, vols=vols, preps=preps, frames=frames, render=trial.render)


def drr_params(asc, wl, desc):
    s = wl["settings"]
    return s["drr_step"], s["drr_intensity"], asc.io.drr_cutoff(s["cutoff_fraction"], desc)


def base_pose(wl, b):
    return np.concatenate([wl["iso"] + np.array(wl["preps"][b]["trans_center"]), [0.0, 0.0, 0.0]])


def build_scene(asc, ctx, wl):
    s = wl["settings"]
    RW, RH = wl["render"]
    sc = asc.Scene(ctx, RW, RH)
    vols, pivots, params = [], [], []
    for v in wl["vols"]:
        vol = asc.Volume(ctx, v["data"], scale=tuple(v["voxel_mm"]), flip=v["flip"])
        pivot = asc.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
        p = drr_params(asc, wl, vol.desc)
        sc.add_volume(vol, *p, pivot)
        vols.append(vol)
        pivots.append(pivot)
        params.append(p)
    for cam, fr in zip(wl["cams"], wl["frames"]):
        c = sc.add_camera(asc.AscCamera.from_array(cam), asc.Image(ctx, fr))
        sc.set_filters(c, 0, s["drr_filters"])
        sc.set_filters(c, 1, s["rad_filters"])
    return sc, vols, pivots, params


def oracle_scene(orc, asc, wl, b, params):
    s = wl["settings"]
    recs = [(asc.filters.to_records(s["drr_filters"]), asc.filters.to_records(s["rad_filters"]))
            for _ in wl["cams"]]
    return orc.Scene(wl["preps"][b], params, wl["cams"], wl["frames"], recs, wl["render"])


def test_trial_bones_match_oracle(asc, ctx, orc, wrist):
    sc, vols, _, params = build_scene(asc, ctx, wrist)
    rng = np.random.default_rng(21)
    B = len(vols)
    bones = np.repeat(np.arange(B), 5)
    poses = np.stack([base_pose(wrist, b) for b in bones])
    poses += np.concatenate([rng.uniform(-3, 3, (len(bones), 3)), rng.uniform(-8, 8, (len(bones), 3))], 1)
    costs, per = sc.eval_poses(poses, volumes=bones)
    info, samples = sc.last_info(len(poses) * 2)
    oracles = [oracle_scene(orc, asc, wrist, b, params[b]) for b in range(B)]
    total, seen = 0, 0
    for i, (b, p) in enumerate(zip(bones, poses)):
        r = oracles[b].track_frame(p)
        for c in range(2):
            assert info[2 * i + c, 0] == r["info"][c, 0]
            if r["info"][c, 0] == 1:
                seen += 1
                assert tuple(info[2 * i + c, 1:4]) == tuple(r["info"][c, 1:4])
                assert abs(per[i, c] - r["corr"][c]) <= 1e-5
        total += r["samples"]
    assert samples == total
    assert seen >= len(poses)  # most poses are seen by both cameras
    print(f"wrist: {len(poses)} poses of {B} bones, {seen} in-view camera images, {samples} ray samples; "
          f"costs {np.round(costs, 4)}")
    sc.close()
    for v in vols:
        v.close()


def test_radius_drr_matches_oracle(asc, ctx, orc, wrist):
    v = wrist["vols"][0]
    vol = asc.Volume(ctx, v["data"], scale=tuple(v["voxel_mm"]), flip=v["flip"])
    prep = wrist["preps"][0]
    step, intensity, cutoff = drr_params(asc, wrist, vol.desc)
    RW, RH = wrist["render"]
    for cam in wrist["cams"]:
        inview, imv, vp, (w, h) = orc.view_geometry(cam, base_pose(wrist, 0), prep["inv_scale"], prep["inv_trans"],
                                                    RW, RH)
        assert inview and w * h > 1000
        ref, _ = orc.drr(prep["volume"], imv[:12], vp, step, intensity, cutoff, w, h)
        out = ctx.alloc(4 * w * h)
        ctx.drr_render(vol, imv[:12], vp, step, intensity, cutoff, out, w, h)
        g = out.download((h, w))
        out.free()
        assert ((g == 0) == (ref == 0)).all()
        nz = ref != 0
        assert nz.mean() > 0.2
        assert (np.abs(g[nz].astype(np.float64) - ref[nz]) / ref[nz]).max() <= 1e-4
    vol.close()


def test_radius_simplex_matches_oracle(asc, ctx, orc, wrist):
    sc, vols, pivots, params = build_scene(asc, ctx, wrist)
    S = oracle_scene(orc, asc, wrist, 0, params[0])
    start = base_pose(wrist, 0) + np.array([1.0, -0.8, 0.6, 2.0, -1.5, 1.0])
    store0 = asc.geometry.pose_store_set(start)
    opts = dict(method=1, max_iter=100, max_stall=25, seed=5)
    s_gpu, st = sc.optimize_frame(store0, volume=0, **opts)
    s_orc, ost, n = S.optimize_frame(store0, pivots[0], nparticles=100, **opts)
    a, b = asc.geometry.pose_store_get(s_gpu), orc.pose_store_get(s_orc)
    print(f"wrist radius simplex: gpu {st.simplex_iterations} iterations, final {st.final_cost:.6f}; "
          f"oracle {ost[1]:.0f}, final {ost[3]:.6f}")
    assert np.allclose(a, b, atol=0.05)
    assert abs(st.final_cost - ost[3]) <= 1e-4
    sc.close()
    for v in vols:
        v.close()
