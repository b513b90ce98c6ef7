"""Requests of several jobs in one scene and one fused launch
(asc_scene_eval_requests): each job has its own volume, camera set, render
size and per-frame radiographs. Every request is compared with the oracle's
trackFrame for that job alone (1 - NCC within 1e-5; render sizes, in-view
flags, mask counts bit-exact); a job's costs are identical alone and in the
mix; removed slots are reused without changing results.
"""
import numpy as np
import pytest

from autoscoper_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs(asc, orc, tiny):
    """Job A: the tiny workload (2 cameras, 96x80 render, 2 frames).
    Job B: bone 1 of the three-bone workload seen by its first camera only,
    at a 72x64 render size, 2 frames."""
    wl3 = synth.multi_workload("tiny3", frames=1)
    camB = orc.camera_from_mayacam2(*wl3.cameras[0].mayacam2())
    prepB = orc.volume_prep(wl3.volumes[1], (wl3.voxel_mm[1],) * 3)
    RWb, RHb = 72, 64
    radsB = []
    for f in range(2):
        _, imv, _, _ = orc.view_geometry(camB, wl3.gt_poses[0, 1], prepB["inv_scale"], prepB["inv_trans"], 96, 80)
        fvp = synth.full_frame_viewport(camB[24:28], 96, 80)
        img, _ = orc.drr(prepB["volume"], imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, 96, 80)
        radsB.append(synth.target_from_drr(img, 300 + f)[1])
    # job A's second frame: its targets with other noise
    radsA2 = [synth.target_from_drr(np.flipud(tiny["rads"][c][0].astype(np.float64)), 400 + c)[1] for c in range(2)]
    return dict(wl3=wl3, camB=camB, prepB=prepB, renderB=(RWb, RHb), radsB=radsB, radsA2=radsA2)


def oracle_job(orc, asc, prep, cams, rads, render, wl):
    recs = [(asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)) for _ in cams]
    return orc.Scene(prep, (synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), cams, rads, recs, render)


def build(asc, ctx, tiny, jobs):
    wl, wl3 = tiny["wl"], jobs["wl3"]
    RW, RH = tiny["render"]
    sc = asc.Scene(ctx, RW, RH)
    keep = []
    volA = asc.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3)
    volB = asc.Volume(ctx, wl3.volumes[1], scale=(wl3.voxel_mm[1],) * 3)
    keep += [volA, volB]
    vA = sc.add_volume(volA, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0)
    vB = sc.add_volume(volB, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0)
    camsA = []
    for i, cam in enumerate(tiny["cams"]):
        c = sc.add_camera(asc.AscCamera.from_array(cam), None)
        sc.set_filters(c, 0, wl.drr_filters)
        sc.set_filters(c, 1, wl.rad_filters)
        camsA.append(c)
    cB = sc.add_camera(asc.AscCamera.from_array(jobs["camB"]), None)
    sc.set_filters(cB, 0, wl3.drr_filters)
    sc.set_filters(cB, 1, wl3.rad_filters)
    sA = sc.add_camera_set(camsA, RW, RH)
    sB = sc.add_camera_set([cB], *jobs["renderB"])
    fA = [sc.add_frame(sA, [asc.Image(ctx, tiny["rads"][c][1]) for c in range(2)]),
          sc.add_frame(sA, [asc.Image(ctx, jobs["radsA2"][c]) for c in range(2)])]
    fB = [sc.add_frame(sB, [asc.Image(ctx, jobs["radsB"][f])]) for f in range(2)]
    return sc, keep, (vA, vB), (sA, sB), (fA, fB), (camsA, cB)


def test_requests_of_two_jobs_match_oracle(asc, ctx, orc, tiny, jobs):
    sc, keep, (vA, vB), (sA, sB), (fA, fB), _ = build(asc, ctx, tiny, jobs)
    wl, wl3 = tiny["wl"], jobs["wl3"]
    oA = [oracle_job(orc, asc, tiny["prep"], tiny["cams"], [tiny["rads"][c][1] for c in range(2)], tiny["render"], wl),
          oracle_job(orc, asc, tiny["prep"], tiny["cams"], jobs["radsA2"], tiny["render"], wl)]
    oB = [oracle_job(orc, asc, jobs["prepB"], [jobs["camB"]], [jobs["radsB"][f]], jobs["renderB"], wl3)
          for f in range(2)]
    rng = np.random.default_rng(77)
    job = np.array([0, 1, 1, 0, 0, 1, 0, 1])
    frame = np.array([0, 0, 1, 1, 0, 1, 1, 0])
    base = np.where(job[:, None] == 0, wl.gt_poses[0], wl3.gt_poses[0, 1])
    poses = base + np.concatenate([rng.uniform(-2, 2, (8, 3)), rng.uniform(-4, 4, (8, 3))], 1)
    vols = np.where(job == 0, vA, vB)
    sets = np.where(job == 0, sA, sB)
    frames = np.array([fA[f] if j == 0 else fB[f] for j, f in zip(job, frame)])
    costs, per = sc.eval_requests(vols, sets, frames, poses)
    for i in range(8):
        r = (oA if job[i] == 0 else oB)[frame[i]].track_frame(poses[i])
        nc = 2 if job[i] == 0 else 1
        for c in range(nc):
            if r["info"][c, 0] == 1:
                assert abs(per[i, c] - r["corr"][c]) <= 1e-5
            else:
                assert per[i, c] == 9999.0
        ref = r["corr"].sum()
        assert abs(costs[i] - (ref if ref >= 0 else 9999)) <= 2e-5
    # each job alone equals its part of the mix
    for j in (0, 1):
        sel = np.nonzero(job == j)[0]
        cj, pj = sc.eval_requests(vols[sel], sets[sel], frames[sel], poses[sel])
        assert np.array_equal(cj, costs[sel]) and np.array_equal(pj, per[sel])
    sc.close()
    for v in keep:
        v.close()


def test_slots_are_reused(asc, ctx, orc, tiny, jobs):
    sc, keep, (vA, vB), (sA, sB), (fA, fB), (camsA, cB) = build(asc, ctx, tiny, jobs)
    wl3 = jobs["wl3"]
    pose = wl3.gt_poses[0, 1] + np.array([0.3, -0.2, 0.1, 1.0, 0.5, -0.5])
    c0, p0 = sc.eval_requests([vB], [sB], [fB[0]], pose[None])
    ca0, _ = sc.eval_requests([vA], [sA], [fA[0]], tiny["wl"].gt_poses[0][None])
    # drop job B entirely, then register it again: same costs
    sc.remove_camera_set(sB)
    sc.remove_camera(cB)
    sc.remove_volume(vB)
    cB2 = sc.add_camera(asc.AscCamera.from_array(jobs["camB"]), None)
    sc.set_filters(cB2, 0, wl3.drr_filters)
    sc.set_filters(cB2, 1, wl3.rad_filters)
    vB2 = sc.add_volume(keep[1], synth.DRR_STEP, synth.DRR_INTENSITY, 0.0)
    sB2 = sc.add_camera_set([cB2], *jobs["renderB"])
    fB2 = sc.add_frame(sB2, [asc.Image(ctx, jobs["radsB"][0])])
    assert (cB2, vB2) == (cB, vB)
    c1, p1 = sc.eval_requests([vB2], [sB2], [fB2], pose[None])
    assert np.array_equal(c0, c1) and np.array_equal(p0, p1)
    # job A is untouched by the churn
    ca1, _ = sc.eval_requests([vA], [sA], [fA[0]], tiny["wl"].gt_poses[0][None])
    assert np.array_equal(ca0, ca1)
    sc.close()
    for v in keep:
        v.close()


def test_submitted_batches_equal_blocking_batches(asc, ctx, tiny, jobs):
    """asc_scene_submit_requests / _finish_requests: two batches in flight
    on the scene's stream, finished in either order, give the same bits as
    asc_scene_eval_requests of each batch on another scene; the second batch
    is larger than the first, so that the fresh scene's buffers grow while
    the first is in flight; a third submit, graph replay and the last-batch
    queries are refused while batches are in flight."""
    ref, keep_r, *_ = build(asc, ctx, tiny, jobs)
    sc, keep, (vA, vB), (sA, sB), (fA, fB), _ = build(asc, ctx, tiny, jobs)
    wl, wl3 = tiny["wl"], jobs["wl3"]
    rng = np.random.default_rng(91)

    def batch(n):
        job = rng.integers(0, 2, n)
        frame = rng.integers(0, 2, n)
        base = np.where(job[:, None] == 0, wl.gt_poses[0], wl3.gt_poses[0, 1])
        poses = base + np.concatenate([rng.uniform(-2, 2, (n, 3)), rng.uniform(-4, 4, (n, 3))], 1)
        vols = np.where(job == 0, vA, vB)
        sets = np.where(job == 0, sA, sB)
        frames = np.array([fA[f] if j == 0 else fB[f] for j, f in zip(job, frame)])
        return vols, sets, frames, poses

    small, large = batch(6), batch(40)
    want = [ref.eval_requests(*small), ref.eval_requests(*large)]
    for order in ((0, 1), (1, 0)):
        t = [sc.submit_requests(*small), sc.submit_requests(*large)]
        assert t[0][0] != t[1][0]
        with pytest.raises(asc.AscError):
            sc.submit_requests(*small)  # both slots in flight
        with pytest.raises(asc.AscError):
            sc.last_moments(4)
        got = [None, None]
        for k in order:
            got[k] = sc.finish_requests(t[k])
        for k in (0, 1):
            assert np.array_equal(got[k][0], want[k][0]) and np.array_equal(got[k][1], want[k][1]), (order, k)
    # the last batch finished is the scene's last batch
    t = sc.submit_requests(*small)
    got = sc.finish_requests(t)
    assert np.array_equal(got[0], want[0][0])
    ref.eval_requests(*small)
    assert np.array_equal(sc.last_moments(6 * 2), ref.last_moments(6 * 2))
    with pytest.raises(asc.AscError):
        sc.finish_requests(t)  # already finished
    sc.use_graph(True)
    with pytest.raises(asc.AscError):
        sc.submit_requests(*small)
    sc.use_graph(False)
    for s in (sc, ref):
        s.close()
    for v in keep + keep_r:
        v.close()


def test_batches_with_no_item_in_view(asc, ctx, tiny, jobs):
    """Edge cases of the fused launch: a batch whose poses are all out of
    view (no tile to march or filter; each camera's value is the 9999
    sentinel of Tracker.cpp:614-618, summed over the set's two cameras as
    Tracker.cpp:814-826 sums them), blocking and submitted, then a batch of
    one pose, each followed by a normal batch whose costs equal those of a
    fresh scene (the work counters and slots are left clean)."""
    sc, keep, (vA, vB), (sA, sB), (fA, fB), _ = build(asc, ctx, tiny, jobs)
    ref, keep_r, *_ = build(asc, ctx, tiny, jobs)
    wl = tiny["wl"]
    rng = np.random.default_rng(5)
    far = wl.gt_poses[0] + np.array([900.0, 900.0, 900.0, 0, 0, 0])
    normal = wl.gt_poses[0] + np.concatenate([rng.uniform(-2, 2, (6, 3)), rng.uniform(-4, 4, (6, 3))], 1)
    vols, sets, frames = np.full(6, vA), np.full(6, sA), np.full(6, fA[0])
    want = ref.eval_requests(vols, sets, frames, normal)
    costs, per = sc.eval_requests(vols[:3], sets[:3], frames[:3], np.repeat(far[None], 3, 0))
    assert (costs == 2 * 9999.0).all() and (per[:, :2] == 9999.0).all()
    assert sc.last_info(6)[1] == 0  # no ray sample
    got = sc.eval_requests(vols, sets, frames, normal)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    t = sc.submit_requests(vols[:2], sets[:2], frames[:2], np.repeat(far[None], 2, 0))
    t2 = sc.submit_requests(vols, sets, frames, normal)
    c_far, _ = sc.finish_requests(t)
    got = sc.finish_requests(t2)
    assert (c_far == 2 * 9999.0).all()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    one = sc.eval_requests(vols[:1], sets[:1], frames[:1], normal[:1])
    assert one[0][0] == want[0][0] and np.array_equal(one[1][0], want[1][0])
    for s in (sc, ref):
        s.close()
    for v in keep + keep_r:
        v.close()


def test_debug_images_follow_the_last_batch(asc, ctx, tiny, jobs):
    """asc_scene_last_images after submitted batches: the images of the
    batch finished last when it was also submitted last (the same as from a
    blocking batch of its own), refused when an earlier batch finished last
    (the device then holds a later batch's images)."""
    sc, keep, (vA, vB), (sA, sB), (fA, fB), _ = build(asc, ctx, tiny, jobs)
    wl = tiny["wl"]
    rng = np.random.default_rng(8)
    pa = wl.gt_poses[0] + np.concatenate([rng.uniform(-1, 1, (2, 3)), rng.uniform(-2, 2, (2, 3))], 1)
    pb = wl.gt_poses[0] + np.concatenate([rng.uniform(-1, 1, (3, 3)), rng.uniform(-2, 2, (3, 3))], 1)
    sc.set_debug(True)
    sc.eval_requests([vA] * 3, [sA] * 3, [fA[0]] * 3, pb)
    info, _ = sc.last_info(6)
    w, h = int(info[0, 1]), int(info[0, 2])
    want = sc.last_images(0, 0, w, h)
    ta = sc.submit_requests([vA] * 2, [sA] * 2, [fA[1]] * 2, pa)
    tb = sc.submit_requests([vA] * 3, [sA] * 3, [fA[0]] * 3, pb)
    sc.finish_requests(ta)
    sc.finish_requests(tb)  # submitted last, finished last
    got = sc.last_images(0, 0, w, h)
    assert all(np.array_equal(g, x) for g, x in zip(got, want))
    ta = sc.submit_requests([vA] * 2, [sA] * 2, [fA[1]] * 2, pa)
    tb = sc.submit_requests([vA] * 3, [sA] * 3, [fA[0]] * 3, pb)
    sc.finish_requests(tb)
    sc.finish_requests(ta)  # finished last, but b's images are on the device
    with pytest.raises(asc.AscError):
        sc.last_images(0, 0, w, h)
    sc.set_debug(False)
    sc.close()
    for v in keep:
        v.close()
