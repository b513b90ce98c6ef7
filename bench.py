#!/usr/bin/env python3
"""Pose-evaluation throughput of the fused DRR -> filters -> NCC path on one
MI355X.

Workload (default "cfg2"): seeded long-bone phantom 256x256x320 u16 at
0.3 mm, two pinhole cameras 90 degrees apart with 1024x1024 detectors,
targets rendered at a hidden pose. One step = one PSO-sized batch of
`--particles` poses (uniform +/-3 mm / +/-3 deg around the hidden pose, the
reference's PSO start range) x 2 cameras evaluated by the fused path, inputs
resident in HBM. Prints one JSON line.

Beside it: the plain one-eval-at-a-time GPU forms (the scene's, and
xromm::Tracker::trackFrame through the C++ facade in the reference's launch
order, asynchronous and with a finish after every launch) and the CPU oracle
on the host cores (bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

STORAGES = {"u16": 0, "f16": 1, "bf16": 2, "f32": 3}
BYTES_PER_VOXEL = {"u16": 2, "f16": 2, "bf16": 2, "f32": 4}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _jsonable(o):
    """numpy scalars and arrays in the result line"""
    if hasattr(o, "tolist"):
        return o.tolist()
    raise TypeError(f"not JSON serializable: {type(o).__name__}")


LAYOUTS = {"linear": 0, "quad": 1, "octet": 2, "brick": 3}


def build_workload(A, synth, name: str, storage: str, ctx, layout: str = "auto"):
    t0 = time.time()
    wl = synth.workload(name, frames=2)
    log(f"[bench] phantom {wl.volume.shape} generated in {time.time() - t0:.1f}s")
    vol = A.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3, storage=STORAGES[storage])
    if layout != "auto":
        vol.set_layout(LAYOUTS[layout])
    RW, RH = wl.render
    cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    gt = wl.gt_poses[0]
    # targets: full-frame DRR at the hidden pose (product path), blur, noise, background
    rads = []
    full = ctx.alloc(4 * RW * RH)
    for i, cam in enumerate(cams):
        inview, imv, _, _ = A.geometry.view_geometry(cam, gt, vol.desc, RW, RH)
        assert inview
        fvp = synth.full_frame_viewport(list(cam.viewport), RW, RH)
        ctx.drr_render(vol, imv, fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, full, RW, RH)
        rads.append(synth.target_from_drr(full.download((RH, RW)), 500 + i))
    full.free()
    scene = A.Scene(ctx, RW, RH)
    pivot = A.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
    scene.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF, pivot)
    for i, cam in enumerate(cams):
        c = scene.add_camera(cam, A.Image(ctx, rads[i][1]))
        scene.set_filters(c, 0, wl.drr_filters)
        scene.set_filters(c, 1, wl.rad_filters)
    return wl, vol, cams, rads, scene, gt


MULTI = ("cfg3", "tiny3")


def build_multi(A, synth, name: str, storage: str, ctx, layout: str = "auto"):
    """Several bones in one scene; targets are the merged full-frame DRRs of
    all bones (View::renderDrr with gpu::merge)."""
    t0 = time.time()
    wl = synth.multi_workload(name, frames=2)
    log(f"[bench] {len(wl.volumes)} phantoms {[v.shape for v in wl.volumes]} generated in {time.time() - t0:.1f}s")
    vols = []
    for v, vox in zip(wl.volumes, wl.voxel_mm):
        vol = A.Volume(ctx, v, scale=(vox,) * 3, storage=STORAGES[storage])
        if layout != "auto":
            vol.set_layout(LAYOUTS[layout])
        vols.append(vol)
    RW, RH = wl.render
    cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    gts = wl.gt_poses[0]
    rads = []
    full, acc = ctx.alloc(4 * RW * RH), ctx.alloc(4 * RW * RH)
    for i, cam in enumerate(cams):
        fvp = synth.full_frame_viewport(list(cam.viewport), RW, RH)
        acc.upload(np.zeros((RH, RW), np.float32))
        for b, vol in enumerate(vols):
            inview, imv, _, _ = A.geometry.view_geometry(cam, gts[b], vol.desc, RW, RH)
            assert inview
            ctx.drr_render(vol, imv, fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, full, RW, RH)
            ctx.merge(acc, full, acc, RW, RH)
        rads.append(synth.target_from_drr(acc.download((RH, RW)), 500 + i))
    full.free()
    acc.free()
    scene = A.Scene(ctx, RW, RH)
    for vol in vols:
        pivot = A.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
        scene.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF, pivot)
    for i, cam in enumerate(cams):
        c = scene.add_camera(cam, A.Image(ctx, rads[i][1]))
        scene.set_filters(c, 0, wl.drr_filters)
        scene.set_filters(c, 1, wl.rad_filters)
    return wl, vols, cams, rads, scene, gts


def cpu_baseline(wl, cams, rads, gt, budget_s: float):
    """Oracle trackFrame (float-faithful restatement, OpenMP over rows) on the
    host cores: a bounded sample of the same workload's evaluations."""
    import oracle as O
    from autoscoper_amd import filters as F
    from autoscoper_amd import synth

    prep = O.volume_prep(wl.volume, (wl.voxel_mm,) * 3)
    recs = [(F.to_records(wl.drr_filters), F.to_records(wl.rad_filters)) for _ in cams]
    S = O.Scene(prep, (synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF), [c.to_array() for c in cams],
                [r[1] for r in rads], recs, wl.render)
    rng = np.random.default_rng(77)
    n, t0 = 0, time.perf_counter()
    while True:
        p = gt + np.concatenate([rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)])
        S.track_frame(p)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return n / el, n, el, threads


SERVE_LOAD = (0.5, 0.8, 1.0)  # arrival rates as fractions of the measured capacity
# trial templates of the serving stream: cameras, detector side, bones (indices into SERVE_POOL)
SERVE_TEMPLATES = [
    dict(ncam=1, det=512, bones=(0,)),
    dict(ncam=2, det=1024, bones=(2,)),
    dict(ncam=4, det=768, bones=(0, 1)),
    dict(ncam=3, det=512, bones=(1,)),
    dict(ncam=2, det=768, bones=(0, 1, 2)),
    dict(ncam=4, det=1024, bones=(2,)),
]
# bone volumes: 64^3 @0.5 mm up to 256x256x400 @0.3 mm (u16)
SERVE_POOL = [((64, 64, 64), 0.5, "ellipsoid", -60.0), ((160, 128, 128), 0.4, "longbone", 0.0),
              ((400, 256, 256), 0.3, "longbone", 70.0)]
SERVE_WALK = 24  # distinct target frames per template; a job walks them back and forth
SERVE_OPTIONS = dict(method=0, max_iter=6, max_stall=3, synchronous=True, speculative=True)


def serve_trials(A, synth, ctx):
    """Seeded trials of mixed shapes (1-4 cameras, 512^2-1024^2 detectors,
    1-3 bones from 64^3 to 256x256x400 u16), each with SERVE_WALK target
    frames rendered at the poses of a random walk (0.5 mm / 1 deg a frame).
    A job of F frames visits them back and forth (0, 1, .., 23, 22, .., 0,
    1, ..), so consecutive frames are always neighbouring walk poses."""
    t0 = time.time()
    vols = [synth.phantom(shape, vox, kind, seed=synth.SEED_PHANTOM + 20 * i, noise_seed=synth.SEED_NOISE + 20 * i)
            for i, (shape, vox, kind, _) in enumerate(SERVE_POOL)]
    dvols = [A.Volume(ctx, v, scale=(p[1],) * 3) for v, p in zip(vols, SERVE_POOL)]
    trials = []
    for ti, tp in enumerate(SERVE_TEMPLATES):
        det = tp["det"]
        specs = synth.biplane(sod_mm=800.0, size=(det, det), pixel_mm=307.2 / det, sid_mm=1100.0, n=tp["ncam"])
        cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in specs]
        gts = []
        for k, b in enumerate(tp["bones"]):
            shape, vox, _, z = SERVE_POOL[b]
            base = synth.volume_centre_pose(shape, vox) + np.array([0.0, 0.0, z, 0.0, 0.0, 0.0])
            gts.append(synth.random_walk(base, SERVE_WALK, 0.5, 1.0, 5000 + 10 * ti + k))
        gts = np.stack(gts, 1)  # frames x bones x 6
        full, acc = ctx.alloc(4 * det * det), ctx.alloc(4 * det * det)
        frames = []
        for f in range(SERVE_WALK):
            per = []
            for c, cam in enumerate(cams):
                fvp = synth.full_frame_viewport(list(cam.viewport), det, det)
                acc.upload(np.zeros((det, det), np.float32))
                for k, b in enumerate(tp["bones"]):
                    inview, imv, _, _ = A.geometry.view_geometry(cam, gts[f, k], dvols[b].desc, det, det)
                    if inview:
                        ctx.drr_render(dvols[b], imv, fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, full, det, det)
                        ctx.merge(acc, full, acc, det, det)
                per.append(synth.target_from_drr(acc.download((det, det)), 7000 + 100 * ti + 10 * f + c)[1])
            frames.append(per)
        full.free()
        acc.free()
        rng = np.random.default_rng(8000 + ti)
        start = gts[0] + np.concatenate([rng.uniform(-2, 2, (len(tp["bones"]), 3)),
                                         rng.uniform(-2, 2, (len(tp["bones"]), 3))], 1)
        trials.append(dict(vols=[vols[b] for b in tp["bones"]], vox=[SERVE_POOL[b][1] for b in tp["bones"]],
                           start=start, cams=cams, frames=frames, render=(det, det)))
    for v in dvols:
        v.close()
    log(f"[serve] {len(trials)} trial templates ready in {time.time() - t0:.1f}s")
    return trials


def walk_frames(frames, n):
    """n frames visiting `frames` back and forth."""
    T = len(frames)
    period = 2 * (T - 1)
    return [frames[f % period if f % period < T else period - f % period] for f in range(n)]


def serve_run(A, synth, ctx, trials, jobs, arrival=True, pipeline=2):
    """Serve `jobs` = [(arrival s, template, frames, particles, seed)] with one
    server (continuous batching, `pipeline` fused launches in flight) ->
    result dict."""
    import threading

    dflt, rflt = synth.default_filters()
    srv = A.Server(ctx, 64 << 30, 16, pipeline=pipeline)
    ids = [None] * len(jobs)
    t0 = time.perf_counter()

    def submitter():
        for k, (a, ti, nframes, particles, s) in enumerate(jobs):
            delay = t0 + a - time.perf_counter()
            if arrival and delay > 0:
                time.sleep(delay)
            tr = trials[ti]
            ids[k] = srv.submit(tr["vols"], tr["vox"], tr["start"], tr["cams"], walk_frames(tr["frames"], nframes),
                                tr["render"], drr_filters=dflt, rad_filters=rflt,
                                drr=(synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF), guess=1,
                                particles=particles, seed=s, **SERVE_OPTIONS)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    frames = 0
    for jid in ids:
        _, lat, st = srv.wait(jid)
        frames += st.frames
    el = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"jobs": len(jobs), "frames": frames, "seconds": el, "frames_per_s": frames / el,
            "latency_p50_ms": st["latency_p50_ms"], "latency_p99_ms": st["latency_p99_ms"], "launches": st["batches"],
            "requests_per_launch": st["requests"] / max(st["batches"], 1), "gpu_busy_frac": st["busy_ms"] / (1e3 * el)}


def serve_bench(A, synth, ctx, trials, loads, jobs_per_rate: int, max_frames: int, seed: int = 9000):
    """cfg5: first the capacity (frames/s with 8 jobs submitted at once),
    then Poisson job arrivals at each fraction of it, `jobs_per_rate` jobs of
    20..max_frames frames each; frames/s and per-frame latency per rate."""
    def draw(rng, n, fmax):
        return [(int(rng.integers(len(trials))), int(rng.integers(20, fmax + 1)), int(rng.choice([50, 100, 200])),
                 int(rng.integers(1 << 20))) for _ in range(n)]

    rng = np.random.default_rng(seed)
    sat_jobs = [(0.0, *j) for j in draw(rng, 8, 40)]
    sat = serve_run(A, synth, ctx, trials, sat_jobs, arrival=False)
    cap = sat["frames_per_s"]
    mean_frames = (20 + max_frames) / 2
    log(f"[serve] capacity {cap:.1f} frames/s (8 jobs at once, p99 {sat['latency_p99_ms']:.0f} ms)")
    one = serve_run(A, synth, ctx, trials, sat_jobs, arrival=False, pipeline=1)
    log(f"[serve] one launch in flight: {one['frames_per_s']:.1f} frames/s (p99 {one['latency_p99_ms']:.0f} ms)")
    out = {"capacity": sat, "capacity_one_launch_in_flight": one, "rates": []}
    for ri, load in enumerate(loads):
        rate = load * cap / mean_frames  # jobs per second
        rng = np.random.default_rng(seed + 1 + ri)
        t, jobs = 0.0, []
        for j in draw(rng, jobs_per_rate, max_frames):
            t += rng.exponential(1.0 / rate)
            jobs.append((t, *j))
        res = {"load": load, "rate_jobs_per_s": rate, **serve_run(A, synth, ctx, trials, jobs)}
        log(f"[serve] {res}")
        out["rates"].append(res)
    return out


def time_fused(ctx, scene, batches, bone_of, ncams, args):
    """Warm up, then time args.steps fused batches (pose upload, the three
    launches, result readback), nothing else in the timed loop. Returns
    (pose-evals/s, ms per step, summed per-kernel ms and summed ray samples
    over an eager pass of the same batches with the kernel timers on)."""
    P = len(bone_of)
    nb = args.warmup + args.steps
    scene.set_timing(False)
    scene.use_graph(args.graph)
    log(f"[bench] warmup {args.warmup} x {P} poses")
    for k in range(args.warmup):
        scene.eval_poses(batches[k], volumes=bone_of)
    ctx.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for k in range(args.warmup, nb):
        scene.eval_poses(batches[k], volumes=bone_of)
    ctx.synchronize()
    el = time.perf_counter() - t0
    # per-kernel times (HIP events around each launch) and ray samples
    kern_ms = np.zeros(3)
    samples = 0
    scene.use_graph(False)
    scene.set_timing(True)
    for k in range(args.warmup, nb):
        scene.eval_poses(batches[k], volumes=bone_of)
        kern_ms += scene.last_timing()
        samples += scene.last_info(P * ncams)[1]
    scene.set_timing(False)
    return P * args.steps / el, 1e3 * el / args.steps, kern_ms, samples


def time_two_in_flight(ctx, scene, batches, bone_of, args):
    """The same batches with two in flight (asc_scene_submit_requests /
    _finish_requests, as the registration server runs its cohorts): the host
    prepares batch k+1 while batch k runs. -> (pose-evals/s, ms per step)."""
    P = len(bone_of)
    nb = args.warmup + args.steps
    scene.use_graph(False)
    scene.set_timing(False)
    sets = np.full(P, -1, np.int32)
    ctx.synchronize()
    t0 = time.perf_counter()
    pending = None
    for k in range(args.warmup, nb):
        ticket = scene.submit_requests(bone_of, sets, None, batches[k])
        if pending is not None:
            scene.finish_requests(pending)
        pending = ticket
    scene.finish_requests(pending)
    el = time.perf_counter() - t0
    return P * args.steps / el, 1e3 * el / args.steps


def storage_variant(A, synth, ctx, args, batches, bone_of, storage):
    """The same workload and poses with another volume storage: pose-evals/s
    and the ray-march kernel's time (its error against the all-double oracle
    is bounded by tests/test_gpu_precision.py)."""
    wl, vol, cams, rads, scene, gt = build_workload(A, synth, args.workload, storage, ctx, args.layout)
    tuning = None
    if args.autotune and not (args.unroll or args.grid):
        tuning = scene.autotune(batches[0], volumes=bone_of, reps=5)
        log(f"[bench] {storage} autotune: {tuning}")
    if args.unroll:
        scene.set_tuning(scene.TUNE_UNROLL, args.unroll)
    if args.grid:
        scene.set_tuning(scene.TUNE_GRID, args.grid)
    evals_s, ms_step, kern_ms, _ = time_fused(ctx, scene, batches, bone_of, len(cams), args)
    log(f"[bench] {storage} volume: {evals_s:.1f} evals/s, kernels ms {kern_ms / args.steps}")
    g_s = None
    if not args.graph:
        g_s, _, _, _ = time_fused(ctx, scene, batches, bone_of, len(cams), argparse.Namespace(**{**vars(args), "graph": True}))
        log(f"[bench] {storage} volume, hipGraph replay: {g_s:.1f} evals/s")
    scene.close()
    vol.close()
    return {"value": evals_s, "unit": "pose-evals/s", "ms_per_step": ms_step, "launch_shape": tuning,
            "graph_replay_value": g_s,
            "kernel_ms": {"drr_tiles": kern_ms[0] / args.steps, "post_tiles": kern_ms[1] / args.steps}}


def plain_facade(synth, wl, cams, rads, poses, evals: int):
    """The reference's own form of the cost: xromm::Tracker::trackFrame of the
    C++ facade (tests/cpp/facade_check.bin), one pose at a time, one launch
    per operator in the reference's order (DRR, DRR filters, radiograph,
    radiograph filters, background, DRR mask, three multiplies, NCC with its
    blocking read-back). Timed as it runs by default and with every
    per-operator call finishing its kernel first (the OpenCL backend's
    clFinish after each launch), with the volume as the reference keeps it
    (u16, here in the densest layout) and as a plain FP32 copy in the linear
    layout; "reference_form" adds gpu::ncc as the reference computes it
    (six float tree sums of two or three launches each, each ended by a
    blocking read). Raises when the program is not built: the batching gain
    is quoted against this form."""
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(root, "tests", "cpp", "facade_check.bin")
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} is not built (python -m autoscoper_amd.build builds it); the plain "
                                "reference-form baseline cannot be timed without it")
    D, H, W = wl.volume.shape
    RW, RH = wl.render
    rh, rw = rads[0][1].shape
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(root, "autoscoper_amd") + os.pathsep + env.get("LD_LIBRARY_PATH", "")
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "facade_in.bin")
        with open(inp, "wb") as f:
            np.array([W, H, D, len(cams), RW, RH, len(poses), rw, rh], np.int32).tofile(f)
            np.array([wl.voxel_mm], np.float64).tofile(f)
            np.ascontiguousarray(wl.volume, np.uint16).tofile(f)
            np.asarray([c.to_array() for c in cams], np.float64).tofile(f)
            for r in rads:
                np.ascontiguousarray(r[1], np.uint16).tofile(f)
            np.asarray(poses, np.float64).tofile(f)
            np.array([synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF], np.float32).tofile(f)
        f32 = ["--storage", "f32", "--layout", "linear"]
        runs = {"u16": [], "u16_blocking": ["--blocking"], "f32_linear_blocking": f32 + ["--blocking"],
                "reference_form": f32 + ["--blocking", "--ncc-tree", "opencl"]}
        for key, extra in runs.items():
            out = subprocess.run([exe, inp, "--time", str(evals), *extra], capture_output=True, text=True,
                                 timeout=300, env=env)
            if out.returncode != 0:
                raise RuntimeError(f"facade_check failed: {out.stderr.strip()}")
            res[key] = json.loads(out.stdout.strip().splitlines()[-1])["evals_per_s"]
    return res


def pose_gap(A, a, b):
    """(translation distance mm, rotation angle deg) between two xyzypr poses."""
    fa, fb = A.geometry.frame_from_xyzypr(a), A.geometry.frame_from_xyzypr(b)
    Ra, Rb = fa[:9].reshape(3, 3), fb[:9].reshape(3, 3)
    c = (np.trace(Ra.T @ Rb) - 1.0) / 2.0
    return float(np.linalg.norm(fa[9:] - fb[9:])), float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))


def frame_targets(A, synth, ctx, wl, vol, cams, frames, seed0):
    """Targets of every frame: the full-frame DRR at the frame's hidden pose
    (product path), blurred, noised and given a background; -> per frame,
    per camera (float32 radiograph, u16 copy)."""
    RW, RH = wl.render
    full = ctx.alloc(4 * RW * RH)
    out = []
    for f in range(frames):
        per = []
        for i, cam in enumerate(cams):
            inview, imv, _, _ = A.geometry.view_geometry(cam, wl.gt_poses[f], vol.desc, RW, RH)
            assert inview
            fvp = synth.full_frame_viewport(list(cam.viewport), RW, RH)
            ctx.drr_render(vol, imv, fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, full, RW, RH)
            per.append(synth.target_from_drr(full.download((RH, RW)), seed0 + 10 * f + i))
        out.append(per)
    full.free()
    return out


def frame_loop(A, synth, ctx, name, storage, frames, opts, seed, restart=None, fp64_frames=0, fp64_budget=60.0,
               wl=None, targets=None, filters=None):
    """Tracker's frame loop on one GPU: for each frame of the hidden random
    walk, the scene's radiographs are the frame's targets, the start is the
    previous frame's result (guess 1; frame 0 starts 2 mm / 2 deg off its
    hidden pose), and asc_optimize_frame runs `opts` (PSO then simplex, or
    the simplex alone). With `restart`, every frame starts from its hidden
    pose plus a fresh perturbation of that size instead. Timed from the first
    frame's start to the last frame's end, radiograph switches included
    (targets are uploaded before). Returns frames/s, evaluations and the
    final pose's distance to the hidden pose per frame; for the first
    `fp64_frames` frames also the distance to the oracle's all-FP64 simplex
    polish started at the GPU's pose (the FP64 cost's own minimum nearby)."""
    wl = wl or synth.workload(name, frames=frames)
    drr_filters, rad_filters = filters or (wl.drr_filters, wl.rad_filters)
    vol = A.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3, storage=STORAGES[storage])
    RW, RH = wl.render
    cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    if targets is None:
        targets = frame_targets(A, synth, ctx, wl, vol, cams, frames, 900)
    imgs = [[A.Image(ctx, t[1]) for t in per] for per in targets]
    scene = A.Scene(ctx, RW, RH)
    pivot = A.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
    scene.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF, pivot)
    for i, cam in enumerate(cams):
        c = scene.add_camera(cam, imgs[0][i])
        scene.set_filters(c, 0, drr_filters)
        scene.set_filters(c, 1, rad_filters)
    rng = np.random.default_rng(synth.SEED_START)
    pert = lambda mm, deg: np.concatenate([rng.uniform(-mm, mm, 3), rng.uniform(-deg, deg, 3)])  # noqa: E731
    guess = wl.gt_poses[0] + pert(2.0, 2.0)
    starts = [wl.gt_poses[f] + pert(*restart) for f in range(frames)] if restart else None
    gaps, evals, finals, costs = [], 0, [], []
    t0 = time.perf_counter()
    for f in range(frames):
        for i in range(len(cams)):
            scene.set_rad(i, imgs[f][i])
        start = starts[f] if restart else guess
        s, st = scene.optimize_frame(A.geometry.pose_store_set(start), seed=seed, **opts)
        p = A.geometry.pose_store_get(s)
        evals += st.evaluations
        costs.append([st.initial_cost, st.pso_best_cost, st.final_cost, st.pso_epochs, st.simplex_iterations])
        gaps.append(pose_gap(A, p, wl.gt_poses[f]))
        finals.append((p, s))
        guess = p
    el = time.perf_counter() - t0
    res = {"frames": frames, "seconds": el, "frames_per_s": frames / el, "pose_evals": int(evals),
           "pose_evals_per_s": evals / el,
           "gap_to_hidden_mm": [g[0] for g in gaps], "gap_to_hidden_deg": [g[1] for g in gaps],
           "max_gap_to_hidden": [max(g[0] for g in gaps), max(g[1] for g in gaps)],
           "per_frame": {"fields": ["initial_cost", "pso_best_cost", "final_cost", "pso_epochs", "simplex_iterations"],
                         "values": costs},
           "final_poses": [f[0].tolist() for f in finals]}
    if fp64_frames:
        import oracle as O
        from autoscoper_amd import filters as F

        prep = O.volume_prep(wl.volume, (wl.voxel_mm,) * 3)
        recs = [(F.to_records(drr_filters), F.to_records(rad_filters)) for _ in cams]
        fp = []
        t1 = time.perf_counter()
        for f in range(min(fp64_frames, frames)):
            if time.perf_counter() - t1 > fp64_budget:
                break
            S = O.Scene(prep, (synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF), [c.to_array() for c in cams],
                        [t[1] for t in targets[f]], recs, wl.render)
            p_gpu, s_gpu = finals[f]
            s64, st64, n64 = S.optimize_frame(s_gpu, pivot, method=1, fp64=True)
            g = pose_gap(A, p_gpu, O.pose_store_get(s64))
            fp.append({"frame": f, "gap_mm": g[0], "gap_deg": g[1], "oracle_evals": n64,
                       "fp64_cost": float(st64[3])})
        res["fp64_polish"] = {"what": "oracle all-FP64 simplex (offsets 0.1) started at the GPU's final pose; the "
                                      "gap is how far the FP64 cost's own minimum lies from the GPU's pose",
                              "frames": fp, "seconds": time.perf_counter() - t1}
    scene.close()
    for per in imgs:
        for im in per:
            im.close()
    vol.close()
    return res


CFG2_TRACK = dict(method=0, max_iter=100, max_stall=101, synchronous=True, speculative=True, particles=100)


def track_filters(A, synth, name):
    """Filter chains of a frame loop: "vie" = the sample .vie settings
    (radiograph Sobel(3, 0.43) -> Contrast(2.5, 0.9), DRR Sobel(1.7, 0.1));
    "matched" = the DRR's Sobel(1.7, 0.1) -> Contrast(2.5, 0.9) on both
    images; "sobel" = Sobel(1.7, 0.1) on both."""
    F = A.filters
    if name == "vie":
        return synth.default_filters()
    if name == "matched":
        c = [F.sobel(1.7, 0.1), F.contrast(2.5, 0.9, 3)]
        return c, list(c)
    if name == "sobel":
        return [F.sobel(1.7, 0.1)], [F.sobel(1.7, 0.1)]
    raise ValueError(name)
CFG1_TRACK = dict(method=1, speculative=True)


def cfg1_leg(A, synth, ctx, args):
    """The small workload (cfg1) on the GPU and the CPU: the 128^3 u16 ellipsoid, two 512^2
    cameras; the fused batch rate (100 poses per launch), the simplex frame
    loop of 10 frames (each started 2 mm / 2 deg off its hidden pose), and
    the CPU oracle's evaluations and simplex frames on the host cores."""
    import oracle as O
    from autoscoper_amd import filters as F

    wl = synth.workload("small", frames=10)
    vol = A.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3)
    cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
    targets = frame_targets(A, synth, ctx, wl, vol, cams, 10, 300)
    RW, RH = wl.render
    scene = A.Scene(ctx, RW, RH)
    pivot = A.geometry.frame_from_xyzypr(list(vol.desc.trans_center) + [0, 0, 0])
    scene.add_volume(vol, synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF, pivot)
    imgs = [A.Image(ctx, t[1]) for t in targets[0]]
    for i, cam in enumerate(cams):
        c = scene.add_camera(cam, imgs[i])
        scene.set_filters(c, 0, wl.drr_filters)
        scene.set_filters(c, 1, wl.rad_filters)
    rng = np.random.default_rng(synth.SEED_START + 7)
    batches = [wl.gt_poses[0] + np.concatenate([rng.uniform(-3, 3, (100, 3)), rng.uniform(-3, 3, (100, 3))], 1)
               for _ in range(args.warmup + args.steps)]
    ev_s, ms, kern, _ = time_fused(ctx, scene, batches, np.zeros(100, np.int32), len(cams), args)
    scene.close()
    for im in imgs:
        im.close()
    vol.close()
    loop = frame_loop(A, synth, ctx, "small", "u16", 10, CFG1_TRACK, 1, restart=(2.0, 2.0), wl=wl, targets=targets)
    out = {"workload": "cfg1: 128x128x128 u16 ellipsoid @0.5mm, 2 cams 512x512",
           "fused": {"value": ev_s, "unit": "pose-evals/s", "ms_per_step": ms, "poses_per_step": 100,
                     "kernel_ms": {"drr_tiles": kern[0] / args.steps, "post_tiles": kern[1] / args.steps}},
           "simplex_frames": loop}
    if not args.no_cpu:
        prep = O.volume_prep(wl.volume, (wl.voxel_mm,) * 3)
        recs = [(F.to_records(wl.drr_filters), F.to_records(wl.rad_filters)) for _ in cams]
        S = O.Scene(prep, (synth.DRR_STEP, synth.DRR_INTENSITY, synth.DRR_CUTOFF), [c.to_array() for c in cams],
                    [t[1] for t in targets[0]], recs, wl.render)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds / 2:
            S.track_frame(batches[0][n % 100])
            n += 1
        ev_cpu = n / (time.perf_counter() - t0)
        t0 = time.perf_counter()
        rng = np.random.default_rng(synth.SEED_START)
        start = wl.gt_poses[0] + np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)])
        s_c, st_c, n_c = S.optimize_frame(A.geometry.pose_store_set(start), pivot, method=1)
        el = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": ev_cpu, "unit": "pose-evals/s", "cores": threads, "kind": "port",
                               "sample": f"{n} evaluations (float-faithful oracle trackFrame)",
                               "simplex_frame_s": el, "simplex_frame_evals": n_c,
                               "simplex_frame_gap_to_hidden": pose_gap(A, O.pose_store_get(s_c), wl.gt_poses[0])}
    return out


def gather_probe(ctx, scene, batches, bone_of, ncams, grids=(0,), reps=5):
    """The ray march's launch with its arithmetic replaced by the bare memory
    stream (ASC_TUNE_GATHER_PROBE): the same tiles, rays, sample counts and
    16-byte cell loads, at each persistent grid of `grids` (0 = one block
    per CU slot); the fastest is the roof. The scene is left at the last
    grid of `grids`. -> (ms per launch, ray samples per launch)."""
    P = len(bone_of)
    scene.use_graph(False)
    scene.set_timing(True)
    scene.set_tuning(scene.TUNE_GATHER_PROBE, 1)
    best = None
    try:
        for grid in grids:
            scene.set_tuning(scene.TUNE_GRID, grid)
            ms, samples = 0.0, 0
            for k in range(reps + 1):
                scene.eval_poses(batches[k % len(batches)], volumes=bone_of)
                if k:
                    ms += float(scene.last_timing()[0])
                    samples += scene.last_info(P * ncams)[1]
            if best is None or ms / reps < best[0]:
                best = (ms / reps, samples / reps)
    finally:
        scene.set_tuning(scene.TUNE_GATHER_PROBE, 0)
    return best


VALU_ISSUE_PER_S = 256 * 2.4e9  # VALU issue cycles per second over the chip: 256 CUs at the 2.4 GHz peak clock


def roofline(drr_ms, samples, bpv, probe, prof, vol_bytes, pixels):
    """Which roof binds drr_tiles, per launch:
      gather  16 B x samples (8 texels x 2 B) at BW_gather, the bandwidth the
              gather probe achieved on the same access stream;
      valu    the VALU issue cycles the kernel needs (rocprof
              SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2 per wave-sample, from
              profiles/, times the launch's wave-samples) at one issue cycle
              per CU per clock;
      hbm     compulsory bytes (the volume once, 4 B per bbox pixel) at 8 TB/s.
    frac = the largest roof time / the measured time."""
    S = samples
    Bg = 8 * bpv * S
    t_meas = drr_ms * 1e-3
    probe_ms, probe_samples = probe
    bw_gather = 8 * bpv * probe_samples / (probe_ms * 1e-3)
    t = {"gather": Bg / bw_gather, "hbm": (vol_bytes + 4 * pixels) / 8.0e12}
    valu = None
    if prof and "valu_issue_cycles_per_wave_sample" in prof:
        cyc = prof["valu_issue_cycles_per_wave_sample"] * S / 64
        t["valu"] = cyc / VALU_ISSUE_PER_S
        valu = {"issue_cycles_per_wave_sample": prof["valu_issue_cycles_per_wave_sample"],
                "instr_per_wave_sample": prof.get("valu_instr_per_wave_sample"),
                "achieved_G_issue_cycles_per_s": cyc / t_meas / 1e9, "peak_G_issue_cycles_per_s": VALU_ISSUE_PER_S / 1e9,
                "source": prof.get("source")}
    bound = max(t, key=t.get)
    out = {"bound": bound, "frac": t[bound] / t_meas, "t_roof_ms": {k: 1e3 * v for k, v in t.items()},
           "measured_ms": drr_ms, "kernel": "drr_tiles", "ray_samples_per_launch": S,
           "algorithmic_bytes_per_launch": Bg,
           "gather": {"bw_gather_GBps": bw_gather / 1e9, "probe_ms": probe_ms, "achieved_GBps": Bg / t_meas / 1e9,
                      "frac": Bg / t_meas / bw_gather},
           "valu": valu}
    if bound == "gather":
        out.update(achieved=Bg / t_meas / 1e9, peak=bw_gather / 1e9, unit="GB/s")
    elif bound == "valu":
        out.update(achieved=valu["achieved_G_issue_cycles_per_s"], peak=valu["peak_G_issue_cycles_per_s"],
                   unit="G VALU issue-cycles/s")
    else:
        out.update(achieved=(vol_bytes + 4 * pixels) / t_meas / 1e9, peak=8000.0, unit="GB/s")
    traffic = prof.get("traffic_bytes_per_launch") if prof else None
    out["traffic"] = traffic
    out["hbm"] = {"compulsory_bytes": vol_bytes + 4 * pixels, "measured_traffic_bytes": traffic,
                  "achieved_GBps": traffic / t_meas / 1e9 if traffic else None,
                  "frac_of_8TBps": traffic / t_meas / 8.0e12 if traffic else None}
    return out


def workload_name(args, wl, cams, P, multi) -> str:
    RW, RH = wl.render
    if multi:
        shapes = ", ".join(f"{v.shape[2]}x{v.shape[1]}x{v.shape[0]}@{vx}mm" for v, vx in zip(wl.volumes, wl.voxel_mm))
        return (f"{args.workload}: {len(wl.volumes)} bones ({shapes}) {args.storage}, {len(cams)} cams {RW}x{RH}, "
                f"{args.particles} poses per bone, {P} per batch")
    return (f"{args.workload}: {wl.volume.shape[2]}x{wl.volume.shape[1]}x{wl.volume.shape[0]} "
            f"{args.storage} volume @{wl.voxel_mm}mm, {len(cams)} cams {RW}x{RH}, {P} poses per batch")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed steps before the timed ones (the host side of a step settles over the first few dozen)")
    ap.add_argument("--particles", type=int, default=100)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--storage", default="u16", choices=sorted(STORAGES))
    ap.add_argument("--layout", default="auto", choices=["auto"] + sorted(LAYOUTS),
                    help="device volume layout (auto: the densest that fits, BRICK for the workloads here)")
    ap.add_argument("--graph", action="store_true", help="replay the fused launches from a hipGraph")
    ap.add_argument("--unroll", type=int, default=0, help="ray-march unroll (tuning; 0 = library default)")
    ap.add_argument("--grid", type=int, default=0, help="persistent grid in blocks (tuning; 0 = library default)")
    ap.add_argument("--plain-evals", type=int, default=10)
    ap.add_argument("--facade-evals", type=int, default=300,
                    help="pose evaluations timed through the C++ facade's trackFrame (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=3,
                    help="serving stream: jobs per arrival rate (0 = skip the stream)")
    ap.add_argument("--serve-frames", type=int, default=200, help="serving stream: frames per job, 20 to this")
    ap.add_argument("--serve-only", action="store_true", help="only the serving stream")
    ap.add_argument("--autotune", type=int, default=1,
                    help="choose the ray march's launch shape with asc_scene_autotune before timing (0 = defaults)")
    ap.add_argument("--variants", type=int, default=1,
                    help="also time the F16-volume variant of a u16 workload (0 = skip)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel over MIB MiB 5 times (counter calibration)")
    ap.add_argument("--frames", type=int, default=3,
                    help="cfg2 frame loop (PSO 100 x 100 + simplex per frame): frames per storage (0 = skip)")
    ap.add_argument("--fp64-frames", type=int, default=1,
                    help="frames of the u16 frame loop checked against the oracle's all-FP64 simplex polish")
    ap.add_argument("--cfg1", type=int, default=1, help="also run the small workload (cfg1: 128^3, 2 x 512^2) on GPU and CPU")
    ap.add_argument("--track-only", default="", metavar="STORAGES",
                    help="only the frame loop, for these comma-separated storages (with --seeds)")
    ap.add_argument("--seeds", default="1", help="PSO seeds of the --track-only frame loops, comma-separated")
    ap.add_argument("--track-filters", default="vie", help="filter chains of the frame loops: vie, matched or sobel "
                                                          "(comma-separated for --track-only)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(jobs and particle batches would shard across GPUs without collectives)",
              file=sys.stderr)
        return 2

    import autoscoper_amd as A
    from autoscoper_amd import synth

    ctx = A.Context(0)
    if args.calib_copy:
        nbytes = args.calib_copy << 20
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        for _ in range(5):
            ctx.copy(dst, src, nbytes)
        ctx.synchronize()
        log(f"[bench] copied {nbytes} bytes x 5")
        return 0
    if args.track_only:
        wl = synth.workload(args.workload, frames=args.frames)
        vol = A.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3)
        cams = [A.geometry.camera_from_mayacam2(*c.mayacam2()) for c in wl.cameras]
        targets = frame_targets(A, synth, ctx, wl, vol, cams, args.frames, 900)
        vol.close()
        res = {}
        for fl in args.track_filters.split(","):
            for st in args.track_only.split(","):
                for seed in [int(x) for x in args.seeds.split(",")]:
                    log(f"[track] {fl} {st} seed {seed}: {args.frames} frames")
                    r = frame_loop(A, synth, ctx, args.workload, st, args.frames, CFG2_TRACK, seed,
                                   fp64_frames=args.fp64_frames, wl=wl, targets=targets,
                                   filters=track_filters(A, synth, fl))
                    log(f"[track] {fl} {st} seed {seed}: {r['frames_per_s']:.2f} frames/s, "
                        f"max gap {r['max_gap_to_hidden']}")
                    res[f"{fl}/{st}/seed{seed}"] = r
        print(json.dumps({"frame_loops": res, "workload": args.workload, "options": CFG2_TRACK}, default=_jsonable),
              flush=True)
        ctx.close()
        return 0
    if args.serve_only:
        trials = serve_trials(A, synth, ctx)
        print(json.dumps({"serving": serve_bench(A, synth, ctx, trials, SERVE_LOAD, args.serve_jobs,
                                                 args.serve_frames)}, default=_jsonable), flush=True)
        ctx.close()
        return 0
    multi = args.workload in MULTI
    if multi:
        wl, vols, cams, rads, scene, gts = build_multi(A, synth, args.workload, args.storage, ctx, args.layout)
    else:
        wl, vol, cams, rads, scene, gt = build_workload(A, synth, args.workload, args.storage, ctx, args.layout)
        vols, gts = [vol], gt[None]
    RW, RH = wl.render
    B = len(vols)
    P = args.particles * B  # poses per fused launch: `particles` per bone
    bone_of = np.repeat(np.arange(B, dtype=np.int32), args.particles)
    rng = np.random.default_rng(synth.SEED_START)
    nb = args.warmup + args.steps
    batches = [gts[bone_of] + np.concatenate([rng.uniform(-3, 3, (P, 3)), rng.uniform(-3, 3, (P, 3))], 1)
               for _ in range(nb)]

    tuning = None
    if args.autotune and not (args.unroll or args.grid):
        # launch shape chosen by the autotuner on the first warm-up batch
        # (outside the timed region; no variant changes a result)
        tuning = scene.autotune(batches[0], volumes=bone_of, reps=5)
        log(f"[bench] autotune: {tuning}")
    if args.unroll:
        scene.set_tuning(scene.TUNE_UNROLL, args.unroll)
    if args.grid:
        scene.set_tuning(scene.TUNE_GRID, args.grid)
    evals_s, ms_step, kern_ms, samples = time_fused(ctx, scene, batches, bone_of, len(cams), args)
    graph_replay = two_in_flight = None
    if not args.graph:
        # the same steps replayed from a hipGraph per batch
        g_args = argparse.Namespace(**{**vars(args), "graph": True})
        g_s, g_ms, _, _ = time_fused(ctx, scene, batches, bone_of, len(cams), g_args)
        scene.use_graph(False)
        scene.set_timing(True)
        graph_replay = {"value": g_s, "unit": "pose-evals/s", "ms_per_step": g_ms}
        log(f"[bench] hipGraph replay {g_s:.1f} evals/s")
        p_s, p_ms = time_two_in_flight(ctx, scene, batches, bone_of, args)
        two_in_flight = {"value": p_s, "unit": "pose-evals/s", "ms_per_step": p_ms}
        log(f"[bench] two batches in flight {p_s:.1f} evals/s")
        scene.set_timing(True)
    info, last_samples = scene.last_info(P * len(cams))
    drr_ms = kern_ms[0] / args.steps
    bpv = BYTES_PER_VOXEL[args.storage]
    log(f"[bench] {evals_s:.1f} evals/s, {ms_step:.2f} ms/step, kernels ms {kern_ms / args.steps}")
    # the probe at the default grid and at the march's tuned one (the scene
    # keeps the tuned grid afterwards)
    tuned_grid = args.grid or (tuning or {}).get("grid", 0)
    probe = gather_probe(ctx, scene, batches, bone_of, len(cams), (0, tuned_grid)) if bpv == 2 else None
    if probe:
        log(f"[bench] gather probe: {probe[0]:.3f} ms per launch, {16 * probe[1] / probe[0] / 1e9:.1f} TB/s")

    # plain one-eval-at-a-time form
    t0 = time.perf_counter()
    for k in range(args.plain_evals):
        scene.track_frame_plain(batches[args.warmup][k % P], volume=int(bone_of[k % P]))
    plain_s = args.plain_evals / (time.perf_counter() - t0)
    log(f"[bench] plain form {plain_s:.1f} evals/s")
    facade = None
    if args.facade_evals > 0 and not multi:
        facade = plain_facade(synth, wl, cams, rads, batches[args.warmup], args.facade_evals)
        log(f"[bench] facade trackFrame {facade} evals/s")

    cpu = None
    if not args.no_cpu and not multi:
        log(f"[bench] CPU oracle baseline (~{args.cpu_seconds:.0f}s)")
        v, n, el_c, threads = cpu_baseline(wl, cams, rads, gt, args.cpu_seconds)
        cpu = {"value": v, "unit": "pose-evals/s", "cores": threads, "kind": "port",
               "sample": f"{n} pose evaluations of the same workload (float-faithful oracle trackFrame, "
                         f"OpenMP over rows, {el_c:.1f}s)"}

    # per-launch counters of the ray-march kernel from the committed rocprofv3
    # runs: calibrated HBM traffic (FETCH_SIZE / WRITE_SIZE) and VALU issue
    # cycles per wave-sample (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2)
    pj = None
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"drr_tiles_{args.workload}_{args.storage}_counters.json")
    if P == 100 and os.path.exists(prof):
        with open(prof) as fh:
            pj = json.load(fh)
        pj["source"] = os.path.relpath(prof, os.path.dirname(os.path.abspath(__file__)))
    tiles = info[:, 4].sum()
    pixels = int((info[:, 1] * info[:, 2]).sum())
    vol_bytes = sum(int(np.prod(v.desc.dims)) * bpv for v in vols)
    roof = None
    if probe:
        roof = roofline(drr_ms, samples / args.steps, bpv, probe, pj, vol_bytes, pixels)
        log(f"[bench] roofline: bound {roof['bound']}, frac {roof['frac']:.3f}, roofs ms {roof['t_roof_ms']}")
    out = {
        "metric": "pose evals/sec (DRR+filters+NCC, 2 cams 1024\u00b2); served frames/s + p99 latency",
        "value": evals_s,
        "unit": "pose-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom, rendered targets)",
        "config": {
            "workload": workload_name(args, wl, cams, P, multi),
            "volume_storage": args.storage,
            "volume_layout": args.layout,
            "particles_per_step": P,
            "cameras": len(cams),
            "render": [RW, RH],
            "graph": bool(args.graph),
            "launch_shape": tuning,
        },
        "roofline": roof,
        "kernel_ms": {"drr_tiles": drr_ms, "post_tiles": kern_ms[1] / args.steps,
                      "finalize_items": kern_ms[2] / args.steps},
        "cpu_baseline": cpu,
        "graph_replay": graph_replay,
        "two_in_flight": two_in_flight,
        "plain_per_eval": {"value": plain_s, "unit": "pose-evals/s",
                           "speedup_of_fused": evals_s / plain_s if plain_s > 0 else None},
        "plain_facade": None if facade is None else {
            "form": "xromm::Tracker::trackFrame through the C++ facade: one pose, one launch per operator in the "
                    "reference's order, blocking NCC read-back per camera; 'blocking' also waits after every "
                    "launch (the OpenCL backend's clFinish); f32_linear: FP32 volume copy, linear layout; "
                    "reference_form (value): f32_linear, blocking, and gpu::ncc as six float tree sums with a "
                    "blocking read each (Ncc.cpp:180-220), about 28 launches and 6 reads per camera",
            "value": facade["reference_form"], "unit": "pose-evals/s", "evals": args.facade_evals,
            "rates": facade,
            "speedup_of_fused": evals_s / facade["reference_form"]},
        "per_eval": {"bbox_pixels": int((info[:, 1] * info[:, 2]).sum() / P),
                     "ray_samples": float(last_samples) / P, "tiles": int(tiles / P)},
    }
    if args.variants and not multi and args.storage == "u16":
        out["storage_variants"] = {st: storage_variant(A, synth, ctx, args, batches, bone_of, st) for st in ("f16",)}
        if facade is not None:
            f16 = out["storage_variants"]["f16"]
            out["plain_facade"]["speedup_of_fused_f16"] = f16["value"] / facade["reference_form"]
            if f16["graph_replay_value"]:
                out["plain_facade"]["speedup_of_fused_f16_graph"] = f16["graph_replay_value"] / facade["reference_form"]
        if probe:
            # an F16 volume's march reads the same 16-byte cells as the u16
            # one's, so the u16 gather probe is its gather roof too
            f16 = out["storage_variants"]["f16"]
            f16["roofline"] = {"bound": "gather", "frac": probe[0] / f16["kernel_ms"]["drr_tiles"],
                               "probe_ms": probe[0], "measured_ms": f16["kernel_ms"]["drr_tiles"]}
    if args.frames > 0 and not multi:
        # Tracker's frame loop at this workload: PSO 100 particles x 100
        # epochs (every epoch runs: stall limit above the epoch count), then
        # the simplex, each frame started from the previous frame's result
        wl_f = synth.workload(args.workload, frames=args.frames)
        targets = frame_targets(A, synth, ctx, wl_f, vol, cams, args.frames, 900)
        out["frame_loop"] = {"options": CFG2_TRACK, "guess": "previous frame's result (frame 0: 2 mm / 2 deg off)"}
        for st in ("u16", "f16"):
            log(f"[bench] frame loop, {st} volume, {args.frames} frames")
            out["frame_loop"][st] = frame_loop(A, synth, ctx, args.workload, st, args.frames, CFG2_TRACK, 1,
                                               fp64_frames=args.fp64_frames if st == "u16" else 0,
                                               wl=wl_f, targets=targets)
            log(f"[bench] {st}: {out['frame_loop'][st]['frames_per_s']:.2f} frames/s, "
                f"max gap to hidden {out['frame_loop'][st]['max_gap_to_hidden']}")
    if args.cfg1 and not multi:
        log("[bench] cfg1 leg")
        out["cfg1"] = cfg1_leg(A, synth, ctx, args)
    if args.serve_jobs > 0:
        scene.close()
        for v in vols:
            v.close()
        vols = []
        trials = serve_trials(A, synth, ctx)
        out["serving"] = {
            "workload": f"cfg5 stream: Poisson job arrivals at 0.5/0.8/1.0 x measured capacity, {args.serve_jobs} "
                        f"jobs per rate; jobs of 1-4 cameras (512^2-1024^2), 1-3 bones (64^3 .. 256x256x400 u16), "
                        f"20-{args.serve_frames} frames, PSO 50-200 particles (synchronous, <= 7 epochs) + simplex "
                        "per frame; targets rendered from hidden random-walk poses",
            "per_frame_latency": "frame start (job arrival or previous frame done) to frame done",
            **serve_bench(A, synth, ctx, trials, SERVE_LOAD, args.serve_jobs, args.serve_frames),
        }
    print(json.dumps(out, default=_jsonable), flush=True)
    scene.close()
    for v in vols:
        v.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
