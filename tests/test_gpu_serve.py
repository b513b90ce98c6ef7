"""Registration server (asc_server_*): tracking jobs of different shapes
served together with continuous batching.

* A job's tracked poses are identical (bit for bit) run alone and run in a
  mix with other jobs, with one fused launch in flight at a time or two: a
  request's cost does not depend on the launch it shares, and each bone's
  optimiser is deterministic given its costs.
* A job's frame loop (guess 1 = previous frame) matches the oracle's
  optimize_frame run frame by frame from the same start: poses within
  0.05 mm / 0.05 deg.
* A device-memory budget that fits one job at a time admits the second job
  only after the first finished, with the same results.
"""
import numpy as np
import pytest

from autoscoper_amd import synth

from test_gpu_optimize import START, pose_gap

pytestmark = pytest.mark.gpu

OPTS = dict(method=0, max_iter=4, max_stall=3, seed=7, particles=40, synchronous=True)


@pytest.fixture(scope="module")
def trial(orc, tiny):
    """The tiny workload as a 3-frame trial: targets rendered at the hidden
    poses of frames 0..2."""
    wl, prep = tiny["wl"], tiny["prep"]
    RW, RH = tiny["render"]
    frames = []
    for f in range(3):
        per_cam = []
        for i, cam in enumerate(tiny["cams"]):
            _, imv, _, _ = orc.view_geometry(cam, wl.gt_poses[f], prep["inv_scale"], prep["inv_trans"], RW, RH)
            fvp = synth.full_frame_viewport(cam[24:28], RW, RH)
            img, _ = orc.drr(prep["volume"], imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, RW, RH)
            per_cam.append(synth.target_from_drr(img, 70 + 10 * f + i)[1])
        frames.append(per_cam)
    return dict(frames=frames, start=wl.gt_poses[0] + START)


def submit_tiny(asc, server, tiny, trial, **over):
    wl = tiny["wl"]
    opts = dict(OPTS, **over)
    return server.submit([wl.volume], [wl.voxel_mm], trial["start"][None],
                         [asc.AscCamera.from_array(c) for c in tiny["cams"]], trial["frames"], tiny["render"],
                         drr_filters=wl.drr_filters, rad_filters=wl.rad_filters,
                         drr=(synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), guess=1, **opts)


def submit_bones(asc, server):
    """Three bones in one camera at its own render size, 2 frames."""
    wl3 = synth.multi_workload("tiny3", frames=2)
    import oracle as orc

    cam = orc.camera_from_mayacam2(*wl3.cameras[1].mayacam2())
    preps = [orc.volume_prep(v, (x,) * 3) for v, x in zip(wl3.volumes, wl3.voxel_mm)]
    frames = []
    for f in range(2):
        fvp = synth.full_frame_viewport(cam[24:28], 96, 80)
        merged = np.zeros((80, 96), np.float32)
        for b, prep in enumerate(preps):
            _, imv, _, _ = orc.view_geometry(cam, wl3.gt_poses[f, b], prep["inv_scale"], prep["inv_trans"], 96, 80)
            img, _ = orc.drr(prep["volume"], imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, 96, 80)
            merged = orc.merge(merged, img).reshape(80, 96)
        frames.append([synth.target_from_drr(merged, 600 + f)[1]])
    starts = wl3.gt_poses[0] + np.array([0.8, -0.6, 0.4, 1.5, -1.0, 1.0])
    return server.submit(wl3.volumes, wl3.voxel_mm, starts, [asc.AscCamera.from_array(cam)], frames, (80, 72),
                         drr_filters=wl3.drr_filters, rad_filters=wl3.rad_filters,
                         drr=(synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), guess=2, **dict(OPTS, seed=3))


@pytest.mark.parametrize("pipeline", [1, 2])
def test_job_alone_equals_job_in_the_mix(asc, ctx, tiny, trial, pipeline):
    """pipeline: fused launches the batcher keeps in flight (2: the next
    one is queued while one runs, the bones in two alternating cohorts)."""
    srv = asc.Server(ctx, 8 << 30, 8, pipeline=1)
    alone, _, _ = srv.wait(submit_tiny(asc, srv, tiny, trial))
    srv.close()
    srv = asc.Server(ctx, 8 << 30, 8, pipeline=pipeline)
    ids = [submit_bones(asc, srv), submit_tiny(asc, srv, tiny, trial), submit_tiny(asc, srv, tiny, trial, seed=11),
           submit_bones(asc, srv)]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    print("mix stats:", {k: st[k] for k in ("batches", "requests", "evaluations", "latency_p50_ms", "latency_p99_ms")})
    assert np.array_equal(res[1][0], alone)
    assert np.array_equal(res[0][0], res[3][0])  # the same job twice in one mix
    assert st["frames_done"] == 3 + 3 + 2 + 2
    assert st["requests"] > st["batches"]  # requests of different jobs shared launches


def test_pipeline_depth_changes_while_jobs_run(asc, ctx, tiny, trial):
    """asc_server_set_pipeline between 1 and 2 while jobs run, with two
    launches possibly in flight at the switch: every job finishes, with the
    same poses as alone."""
    import time

    srv = asc.Server(ctx, 8 << 30, 8, pipeline=1)
    alone, _, _ = srv.wait(submit_tiny(asc, srv, tiny, trial))
    srv.close()
    srv = asc.Server(ctx, 8 << 30, 8, pipeline=2)
    ids = [submit_tiny(asc, srv, tiny, trial), submit_bones(asc, srv), submit_tiny(asc, srv, tiny, trial)]
    for depth in (1, 2, 1, 2):
        time.sleep(0.005)
        srv.set_pipeline(depth)
    res = [srv.wait(i) for i in ids]
    with pytest.raises(asc.AscError):
        srv.set_pipeline(3)
    srv.close()
    assert np.array_equal(res[0][0], alone) and np.array_equal(res[2][0], alone)


def test_frame_loop_matches_oracle(asc, ctx, orc, tiny, trial):
    srv = asc.Server(ctx, 8 << 30, 4)
    poses, lat, st = srv.wait(submit_tiny(asc, srv, tiny, trial))
    srv.close()
    assert (lat >= 0).all() and st.frames == 3
    wl = tiny["wl"]
    pivot = orc.frame_from_xyzypr(list(tiny["prep"]["trans_center"]) + [0, 0, 0])
    recs = [(asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)) for _ in tiny["cams"]]
    pose = trial["start"]
    for f in range(3):
        S = orc.Scene(tiny["prep"], (synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), tiny["cams"], trial["frames"][f], recs,
                      tiny["render"])
        o = dict(OPTS)
        s_orc, _, _ = S.optimize_frame(orc.pose_store_set(pose), pivot, nparticles=o.pop("particles"), **o)
        pose = orc.pose_store_get(s_orc)
        dt, dr = pose_gap(orc, poses[f, 0], pose)
        print(f"frame {f}: gap to oracle {dt:.2e} mm {dr:.2e} deg")
        assert dt <= 0.05 and dr <= 0.05


def job_bytes(tiny, trial, particles):
    """The server's device-memory estimate of a tiny job: the linear volume
    copy and three 8-corner BRICK copies (one per ray axis), float frames,
    one PSO epoch of DRR scratch."""
    D, H, W = tiny["wl"].volume.shape
    RW, RH = tiny["render"]
    frames = sum(im.size for fr in trial["frames"] for im in fr) * 4
    vol = (W + 2) * (H + 2) * (D + 2) * 2 + (W + 8) * (H + 8) * (D + 8) * 2 * 8 * 3
    return vol + frames + particles * len(tiny["cams"]) * RW * RH * 4


@pytest.mark.parametrize("rule", ["active_jobs", "memory_budget"])
def test_admission_one_job_at_a_time(asc, ctx, tiny, trial, rule):
    srv = asc.Server(ctx, 8 << 30, 8)
    ref, _, _ = srv.wait(submit_tiny(asc, srv, tiny, trial))
    srv.close()
    if rule == "active_jobs":
        srv = asc.Server(ctx, 8 << 30, 1)
    else:
        one = job_bytes(tiny, trial, OPTS["particles"])
        srv = asc.Server(ctx, int(one * 1.5), 8)  # room for one job, not two
        with pytest.raises(asc.AscError):  # a job that can never fit is refused at once
            submit_tiny(asc, srv, tiny, trial, particles=400)
    a = submit_tiny(asc, srv, tiny, trial)
    b = submit_tiny(asc, srv, tiny, trial)
    pa, _, sa = srv.wait(a)
    pb, _, sb = srv.wait(b)
    st = srv.stats()
    srv.close()
    # the scene's batch scratch (kept after the jobs) is what the server counts
    assert 0 < st["device_bytes_in_use"]
    if rule == "memory_budget":
        assert st["device_bytes_in_use"] <= int(one * 1.5)
    print(f"{rule}: a queued {sa.queue_ms:.1f} ms, ran to {sa.total_ms:.1f} ms; b queued {sb.queue_ms:.1f} ms")
    # b was admitted only after a finished (b arrived a few ms after a)
    assert sb.queue_ms >= sa.total_ms - 20.0
    assert np.array_equal(pa, ref) and np.array_equal(pb, ref)


def test_front_ends_share_launches(asc, ctx, tiny, trial):
    """Two front ends on one server (the server's scene under
    asc_server_lock_scene; asc_server_eval_requests inside
    asc_server_stream_begin / _end), each posting a chain of single-pose
    requests from its own thread: their requests share launches, and every
    cost is the same bits as the chain run with nothing beside it."""
    import threading

    wl = tiny["wl"]
    RW, RH = tiny["render"]
    srv = asc.Server(ctx, 16 << 30, 4)
    vol = asc.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3)
    with srv.scene() as sc:
        v = sc.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0)
        cams = []
        for cam in tiny["cams"]:
            c = sc.add_camera(asc.AscCamera.from_array(cam), None)
            sc.set_filters(c, 0, wl.drr_filters)
            sc.set_filters(c, 1, wl.rad_filters)
            cams.append(c)
        cs = sc.add_camera_set(cams, RW, RH)
        frames = [sc.add_frame(cs, [asc.Image(ctx, r) for r in trial["frames"][f]]) for f in range(2)]
    rng = np.random.default_rng(17)
    chains = [wl.gt_poses[0] + np.concatenate([rng.uniform(-2, 2, (40, 3)), rng.uniform(-3, 3, (40, 3))], 1)
              for _ in range(2)]

    def run(k):
        return [srv.eval_requests([v], [cs], [frames[k]], p[None])[0][0] for p in chains[k]]

    alone = [run(k) for k in range(2)]
    before = srv.stats()
    shared = [None, None]
    errors = []

    def worker(k):
        try:
            with srv.stream():
                shared[k] = run(k)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    after = srv.stats()
    srv.close()
    vol.close()
    assert not errors, errors
    batches = after["batches"] - before["batches"]
    requests = after["requests"] - before["requests"]
    print(f"front ends: {requests} requests in {batches} launches")
    assert requests == 80
    assert requests > batches
    for k in range(2):
        assert np.array_equal(np.array(shared[k]), np.array(alone[k])), k
