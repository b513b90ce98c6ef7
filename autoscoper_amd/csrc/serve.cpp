// Registration server: continuous batching of many tracking jobs on one GPU.
//
// A job is a trial: one or more bones (volumes), a camera set with its
// render size and filter chains, a stack of radiograph frames, optimiser
// options and a guess heuristic. Each bone of an admitted job runs the
// reference's frame loop (Tracker::optimize per frame, Tracker.cpp:337-580,
// with guess 1 = previous frame and guess 2 = linear extrapolation,
// :380-427) in its own host thread; every cost evaluation the optimiser
// asks for becomes a request to the batcher. The batcher thread is the only
// one that evaluates: it registers staged jobs in one shared scene and
// evaluates the pending requests of all running bones in one fused launch
// (asc_scene_eval_requests) per tick. A stager thread, with a HIP stream of
// its own, admits queued jobs when their device memory fits the budget,
// makes their device volumes and radiographs, and frees the device copies
// of finished jobs, so that neither the host-side volume packing nor the
// copies and frees stall the evaluations. A request's cost does not depend
// on what else is in the launch, so every job's trajectory is the same
// alone and in the mix.
//
// Front ends that keep their own trial state (a socket session of the
// reference's protocol) share the same scene and launches: they register
// their volumes, cameras and camera sets in the server's scene under its
// lock, and their evaluations are requests like the bones' (external
// requests), so several clients' optimisations ride in one fused launch.
//
// Per-frame latency of a job: from the moment the frame could start (job
// arrival for the first frame, completion of the previous frame by all
// bones for the next ones) to its completion by all bones.
#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "geometry.hpp"

using Clock = std::chrono::steady_clock;

namespace {

double ms_since(Clock::time_point t0, Clock::time_point t) { return std::chrono::duration<double, std::milli>(t - t0).count(); }

struct Job;

// One cost evaluation batch: of a bone (job set), n manipulator offsets
// relative to the bone's current float pose store; of a front end (job
// null), n (volume, camera set, frame, pose) requests of the shared scene.
struct Request {
  Job* job;
  int bone;
  int frame;
  const double* manips;
  int n;
  const float* store;
  double* costs;
  // external requests
  const int* ext_vols = nullptr;
  const int* ext_sets = nullptr;
  const int* ext_frames = nullptr;
  const double* ext_poses = nullptr;
  double* per_camera = nullptr;  // n x 8, optional
  bool done = false;
  int status = ASC_OK;
};

struct Job {
  int id = 0;
  // host copy of the description
  int n_bones = 0;
  std::vector<std::vector<uint16_t>> vol_data;
  std::vector<std::array<int, 3>> vol_dims;
  std::vector<std::array<double, 3>> vol_scale;
  int storage = ASC_STORAGE_U16;
  std::vector<std::array<double, 6>> init_pose;
  float step = 0.5f, intensity = 1.0f, cutoff = 0.0f;
  std::vector<asc_camera> cams;
  int RW = 0, RH = 0;
  std::vector<asc_filter> drr_chain, rad_chain;
  int n_frames = 0;
  std::vector<std::vector<unsigned char>> frames;  // [f * ncam + c]
  std::vector<std::array<int, 2>> frame_dims;      // per camera
  int bps = 16;
  asc_optimize_options opt{};
  int guess = 1;
  size_t device_bytes = 0;
  // timing
  Clock::time_point arrival;
  double admit_ms = -1.0;
  // device state (batcher thread only)
  std::vector<asc_volume> dvols;
  std::vector<int> svols, scams;
  int set = -1;
  std::vector<asc_image> images;
  std::vector<int> sframes;
  std::vector<std::array<double, 12>> pivots;
  // results
  std::vector<double> poses;                     // [f][b][6]
  std::vector<std::vector<double>> bone_done_ms;  // [b][f], since arrival
  std::vector<double> frame_latency_ms;          // [f]
  std::vector<std::thread> threads;
  int bones_left = 0;
  int64_t evaluations = 0;
  int status = ASC_OK;
  size_t data_bytes = 0;     // volumes and frames
  size_t scratch_share = 0;  // one PSO epoch of batch scratch
  // admitted: taken from the queue by the stager; registered: in the scene
  // (or failed) with bones_left set; released: out of the scene, queued for
  // the stager to free; finished: device copies freed, results final
  bool admitted = false, registered = false, finished = false, released = false;
};

}  // namespace

struct asc_server_s {
  asc_context ctx;
  asc_context up_ctx = nullptr;  // the stager's context (its own stream)
  asc_scene scene = nullptr;
  // held while the scene is launched or edited (the batcher, or a front
  // end between asc_server_lock_scene and asc_server_unlock_scene)
  std::mutex scene_mu;
  std::thread::id scene_owner;
  // admission: device bytes of the active jobs' volumes and frames, and
  // their shares of the batch scratch; the scene's scratch only grows, so
  // what it holds counts when it exceeds the shares
  size_t budget = 0, used_data = 0, used_share = 0;
  int max_active = 1, active = 0;
  std::mutex mu;
  std::condition_variable cv_batcher, cv_req, cv_done, cv_stager;
  std::deque<Job*> waiting;                  // submitted, not admitted
  std::deque<Job*> staged;                   // device copies made, not registered
  std::deque<Job*> to_free;                  // out of the scene, device copies to free
  std::map<int, std::unique_ptr<Job>> jobs;  // all not yet collected by asc_server_wait
  std::vector<Request*> pending;
  // submitted launches, oldest first (at most two: the scene's slots); the
  // completer thread waits for each and hands its results out
  struct Launch {
    int ticket = 0;
    std::vector<Request*> reqs;
    size_t n = 0;
    int bone_reqs = 0, ext_reqs = 0;
    Clock::time_point t;
  };
  std::deque<Launch> launches;
  int depth = 2;  // launches in flight (asc_server_set_pipeline)
  bool stop_completer = false;
  std::condition_variable cv_completer;
  // the current run of launches in flight started at busy_from; the last
  // one ended at busy_to
  Clock::time_point busy_from, busy_to;
  int running_bones = 0;
  int streams = 0;  // front ends inside asc_server_stream_begin / _end
  int next_id = 1;
  bool stop = false, stop_stager = false;
  std::thread batcher, stager, completer;
  Clock::time_point t0;
  // statistics
  size_t scene_scratch = 0;  // the scene's device scratch after the last launch
  int64_t batches = 0, requests = 0, evaluations = 0, frames_done = 0, jobs_done = 0;
  double busy_ms = 0.0;
  std::vector<double> latencies_ms;
};

namespace {

int bone_cost(const double* manips, int n, double* costs, void* user);

struct BoneCtx {
  asc_server s;
  Job* job;
  int bone;
  int frame;
  float* store;
};

int bone_cost(const double* manips, int n, double* costs, void* user)
{
  BoneCtx* b = (BoneCtx*)user;
  asc_server s = b->s;
  Request r;
  r.job = b->job;
  r.bone = b->bone;
  r.frame = b->frame;
  r.manips = manips;
  r.n = n;
  r.store = b->store;
  r.costs = costs;
  std::unique_lock<std::mutex> lk(s->mu);
  s->pending.push_back(&r);
  s->cv_batcher.notify_one();
  s->cv_req.wait(lk, [&] { return r.done; });
  return r.status;
}

// Tracker::optimize's initial guess for frame f (Tracker.cpp:380-427).
void initial_guess(const Job& j, int b, int f, double* out)
{
  const int B = j.n_bones;
  auto pose = [&](int fr) { return &j.poses[((size_t)fr * B + b) * 6]; };
  if (f == 0) {
    std::memcpy(out, j.init_pose[b].data(), 6 * sizeof(double));
  } else if (j.guess == 2 && f >= 2) {
    // float curves hold the stored poses; extrapolate frames f-2, f-1
    float s1[7], s2[7];
    double p1[6], p2[6];
    asc_pose_store_set(s1, pose(f - 2));
    asc_pose_store_get(s1, p1);
    asc_pose_store_set(s2, pose(f - 1));
    asc_pose_store_get(s2, p2);
    const asc::Frame x = asc::frame_linear_extrap(asc::frame_from_xyzypr(p1), asc::frame_from_xyzypr(p2));
    asc::frame_to_xyzypr(x, out);
  } else {
    std::memcpy(out, pose(f - 1), 6 * sizeof(double));
  }
}

void bone_main(asc_server s, Job* j, int b)
{
  const int F = j->n_frames, B = j->n_bones;
  int status = ASC_OK;
  for (int f = 0; f < F && status == ASC_OK; f++) {
    double guess[6];
    initial_guess(*j, b, f, guess);
    float store[7];
    asc_pose_store_set(store, guess);
    BoneCtx ctx{s, j, b, f, store};
    asc_optimize_options opt = j->opt;
    asc_optimize_stats st;
    status = asc_optimize_frame_fn(bone_cost, &ctx, store, j->pivots[b].data(), &opt, &st);
    double result[6];
    asc_pose_store_get(store, result);
    std::lock_guard<std::mutex> lk(s->mu);
    std::memcpy(&j->poses[((size_t)f * B + b) * 6], result, sizeof result);
    j->bone_done_ms[b][f] = ms_since(j->arrival, Clock::now());
    j->evaluations += st.evaluations;
    if (status != ASC_OK) j->status = status;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->running_bones--;
  j->bones_left--;
  s->cv_batcher.notify_one();
}

// Device volumes and radiographs of an admitted job (stager thread, on the
// server's upload context). Each create call has finished its copies and
// layout kernels when it returns, so the batcher's stream may read them.
int upload_job(asc_context up, Job& j)
{
  int rc;
  for (int b = 0; b < j.n_bones; b++) {
    asc_volume v = nullptr;
    asc_volume_desc d;
    const int flip[3] = {0, 0, 0};
    const std::array<int, 3>& dm = j.vol_dims[b];
    if ((rc = asc_volume_create(up, j.vol_data[b].data(), dm[0], dm[1], dm[2], 16, flip, j.vol_scale[b].data(),
                                j.storage, &v, &d)))
      return rc;
    j.dvols.push_back(v);
    // Tracker.cpp:281-290: the pivot is the volume centre
    const double ypr0[6] = {d.trans_center[0], d.trans_center[1], d.trans_center[2], 0, 0, 0};
    std::array<double, 12> pv;
    asc_frame_from_xyzypr(ypr0, pv.data());
    j.pivots.push_back(pv);
  }
  const int C = (int)j.cams.size();
  for (int f = 0; f < j.n_frames; f++)
    for (int c = 0; c < C; c++) {
      asc_image im = nullptr;
      if ((rc = asc_image_create(up, j.frames[(size_t)f * C + c].data(), j.frame_dims[c][0], j.frame_dims[c][1],
                                 j.bps, &im)))
        return rc;
      j.images.push_back(im);
    }
  return ASC_OK;
}

// Scene entries of a staged job (batcher thread).
int register_job(asc_server s, Job& j)
{
  int rc;
  for (int b = 0; b < j.n_bones; b++) {
    int idx;
    if ((rc = asc_scene_add_volume(s->scene, j.dvols[b], j.step, j.intensity, j.cutoff, j.pivots[b].data(), &idx)))
      return rc;
    j.svols.push_back(idx);
  }
  const int C = (int)j.cams.size();
  for (int c = 0; c < C; c++) {
    int idx;
    if ((rc = asc_scene_add_camera(s->scene, &j.cams[c], nullptr, &idx))) return rc;
    j.scams.push_back(idx);
    if ((rc = asc_scene_set_filters(s->scene, idx, 0, j.drr_chain.data(), (int)j.drr_chain.size()))) return rc;
    if ((rc = asc_scene_set_filters(s->scene, idx, 1, j.rad_chain.data(), (int)j.rad_chain.size()))) return rc;
  }
  if ((rc = asc_scene_add_camera_set(s->scene, j.scams.data(), C, j.RW, j.RH, &j.set))) return rc;
  for (int f = 0; f < j.n_frames; f++) {
    int fid;
    if ((rc = asc_scene_add_frame(s->scene, j.set, &j.images[(size_t)f * C], &fid))) return rc;
    j.sframes.push_back(fid);
  }
  return ASC_OK;
}

// Scene entries out (batcher thread). No later launch names the job's
// device copies, so the stager may free them while the batcher evaluates.
void unregister_job(asc_server s, Job& j)
{
  for (int f : j.sframes) asc_scene_remove_frame(s->scene, f);
  if (j.set >= 0) asc_scene_remove_camera_set(s->scene, j.set);
  for (int c : j.scams) asc_scene_remove_camera(s->scene, c);
  for (int v : j.svols) asc_scene_remove_volume(s->scene, v);
  j.sframes.clear();
  j.scams.clear();
  j.svols.clear();
  j.set = -1;
}

// Device copies out (stager thread).
void free_job(Job& j)
{
  for (asc_image im : j.images) asc_image_destroy(im);
  for (asc_volume v : j.dvols) asc_volume_destroy(v);
  j.images.clear();
  j.dvols.clear();
}

// Stager thread: frees finished jobs first (their memory may admit the
// next one), then admits the queue's head when it fits the budget.
void stager_main(asc_server s)
{
  std::unique_lock<std::mutex> lk(s->mu);
  auto fits = [&] {
    if (s->waiting.empty() || s->active >= s->max_active) return false;
    const Job* j = s->waiting.front();
    return s->used_data + j->data_bytes + std::max(s->scene_scratch, s->used_share + j->scratch_share) <= s->budget;
  };
  for (;;) {
    s->cv_stager.wait(lk, [&] { return s->stop_stager || !s->to_free.empty() || fits(); });
    if (!s->to_free.empty()) {
      Job* j = s->to_free.front();
      s->to_free.pop_front();
      lk.unlock();
      free_job(*j);
      lk.lock();
      s->active--;
      s->used_data -= j->data_bytes;
      s->used_share -= j->scratch_share;
      j->finished = true;
      s->cv_done.notify_all();
    } else if (fits()) {
      Job* j = s->waiting.front();
      s->waiting.pop_front();
      s->active++;
      s->used_data += j->data_bytes;
      s->used_share += j->scratch_share;
      j->admitted = true;
      j->admit_ms = ms_since(j->arrival, Clock::now());
      lk.unlock();
      const int rc = upload_job(s->up_ctx, *j);
      lk.lock();
      if (rc != ASC_OK) j->status = rc;
      s->staged.push_back(j);
      s->cv_batcher.notify_one();
    } else {
      break;  // stopping, nothing left to free
    }
  }
}

// Called with the lock held: scene entries and bone threads for the staged
// jobs.
void register_staged(asc_server s, std::unique_lock<std::mutex>& lk)
{
  while (!s->staged.empty()) {
    Job* j = s->staged.front();
    s->staged.pop_front();
    int rc = j->status;
    if (rc == ASC_OK) {
      lk.unlock();
      {
        std::lock_guard<std::mutex> sl(s->scene_mu);
        rc = register_job(s, *j);
      }
      lk.lock();
    }
    j->registered = true;
    if (rc != ASC_OK) {
      j->status = rc;
      j->bones_left = 0;
      continue;
    }
    j->bones_left = j->n_bones;
    s->running_bones += j->n_bones;
    for (int b = 0; b < j->n_bones; b++) j->threads.emplace_back(bone_main, s, j, b);
  }
}

// Called with the lock held: take jobs whose bones are all done out of the
// scene and hand their device copies to the stager.
void reap(asc_server s, std::unique_lock<std::mutex>& lk)
{
  for (auto& kv : s->jobs) {
    Job& j = *kv.second;
    if (!j.registered || j.released || j.bones_left > 0) continue;
    std::vector<std::thread> th;
    th.swap(j.threads);
    lk.unlock();
    for (std::thread& t : th) t.join();
    {
      std::lock_guard<std::mutex> sl(s->scene_mu);
      unregister_job(s, j);
    }
    lk.lock();
    j.released = true;
    // frame latency: completion by all bones, from when the frame could start
    double prev = 0.0;
    for (int f = 0; f < j.n_frames; f++) {
      double done = 0.0;
      for (int b = 0; b < j.n_bones; b++) done = std::max(done, j.bone_done_ms[b][f]);
      j.frame_latency_ms[f] = done - prev;
      prev = done;
      if (j.status == ASC_OK) s->latencies_ms.push_back(j.frame_latency_ms[f]);
    }
    s->frames_done += j.status == ASC_OK ? j.n_frames : 0;
    s->jobs_done++;
    s->to_free.push_back(&j);
    s->cv_stager.notify_one();
  }
}

// Hand a launch's results to its requesters (lock held).
void deliver(asc_server s, const std::vector<Request*>& reqs, const std::vector<double>& costs,
             const std::vector<double>& per, int rc)
{
  size_t k = 0;
  for (Request* r : reqs) {
    for (int i = 0; i < r->n; i++, k++) {
      if (rc != ASC_OK) continue;
      r->costs[i] = costs[k];
      if (r->per_camera) std::memcpy(r->per_camera + 8 * i, &per[k * 8], 8 * sizeof(double));
    }
    r->status = rc;
    r->done = true;
  }
  s->cv_req.notify_all();
}

// Completer thread: waits for the oldest submitted launch (outside the
// scene lock, so that the batcher can submit the next one meanwhile), takes
// its costs and wakes its requesters.
void completer_main(asc_server s)
{
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_completer.wait(lk, [&] { return s->stop_completer || !s->launches.empty(); });
    if (s->launches.empty()) break;  // stopping
    const int ticket = s->launches.front().ticket;
    lk.unlock();
    int rc = asc_scene_wait_requests(s->scene, ticket);
    lk.lock();
    asc_server_s::Launch b = std::move(s->launches.front());
    lk.unlock();
    std::vector<double> costs(b.n), per(b.n * 8, 0.0);
    size_t scratch = 0;
    {
      std::lock_guard<std::mutex> sl(s->scene_mu);
      // finish even after a failed wait: it frees the slot
      const int frc = asc_scene_finish_requests(s->scene, ticket, costs.data(), per.data());
      if (rc == ASC_OK) rc = frc;
      asc_scene_scratch_bytes(s->scene, &scratch);
    }
    const Clock::time_point t = Clock::now();
    lk.lock();
    s->launches.pop_front();
    s->batches++;
    s->requests += (int64_t)b.reqs.size();
    s->evaluations += (int64_t)b.n;
    // busy: the union of the launches' submit-to-finish intervals
    if (s->launches.empty()) {
      s->busy_ms += ms_since(s->busy_from, t);
      s->busy_to = t;
    }
    s->scene_scratch = scratch;
    deliver(s, b.reqs, costs, per, rc);
    s->cv_batcher.notify_one();  // a slot is free
    s->cv_stager.notify_one();   // the scratch may have grown: re-check admission
  }
}

// Poses a launch must hold before the batcher splits the requesters into
// two cohorts (half of them start a launch).
constexpr int kCohortPoses = 32;

void batcher_main(asc_server s)
{
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    reap(s, lk);
    register_staged(s, lk);
    // a tick: wait until every running bone and every announced stream of
    // external requests has a request pending or in flight (between two
    // optimiser steps they do only host work), or briefly for stragglers;
    // any other external request waits the same way. With two launches in
    // flight and nothing in flight, half of them start a launch once it
    // holds kCohortPoses poses, so that the requesters settle into two
    // cohorts whose launches alternate: one runs while the other's
    // requesters do their host work. (Requests of a few poses keep sharing
    // one launch: there a launch's fixed cost outweighs the host work.)
    auto ready = [&] {
      if (s->stop) return true;
      if ((int)s->launches.size() >= s->depth) return false;
      int bone_reqs = 0, poses = 0;
      for (const Request* r : s->pending) {
        bone_reqs += r->job != nullptr;
        poses += r->n;
      }
      const int ext_reqs = (int)s->pending.size() - bone_reqs;
      int fly_bone = 0, fly_ext = 0;
      for (const auto& b : s->launches) {
        fly_bone += b.bone_reqs;
        fly_ext += b.ext_reqs;
      }
      int need_bone = s->running_bones - fly_bone, need_ext = s->streams - fly_ext;
      if (s->depth > 1 && s->launches.empty() && poses >= kCohortPoses) {
        need_bone = (need_bone + 1) / 2;
        need_ext = (need_ext + 1) / 2;
      }
      return (!s->pending.empty() && bone_reqs >= need_bone && ext_reqs >= need_ext) || !s->staged.empty() ||
             std::any_of(s->jobs.begin(), s->jobs.end(), [](const auto& kv) {
               return kv.second->registered && !kv.second->released && kv.second->bones_left == 0;
             });
    };
    if (s->pending.empty())
      s->cv_batcher.wait(lk, ready);
    else
      s->cv_batcher.wait_for(lk, std::chrono::microseconds(500), ready);
    if (s->stop && s->pending.empty() && s->running_bones == 0) break;
    if (s->pending.empty()) continue;
    if ((int)s->launches.size() >= s->depth) {
      // both slots in flight: wait for the completer (stragglers included)
      s->cv_batcher.wait(lk, [&] { return (int)s->launches.size() < s->depth; });
      continue;
    }
    asc_server_s::Launch b;
    b.reqs.swap(s->pending);
    lk.unlock();
    // one fused launch for every pending request
    size_t n = 0;
    for (Request* r : b.reqs) {
      n += r->n;
      (r->job ? b.bone_reqs : b.ext_reqs)++;
    }
    b.n = n;
    std::vector<int> vols(n), sets(n), frames(n);
    std::vector<double> poses(n * 6);
    size_t k = 0;
    for (Request* r : b.reqs) {
      if (!r->job) {
        for (int i = 0; i < r->n; i++, k++) {
          vols[k] = r->ext_vols[i];
          sets[k] = r->ext_sets[i];
          frames[k] = r->ext_frames ? r->ext_frames[i] : -1;
          std::memcpy(&poses[k * 6], r->ext_poses + 6 * i, 6 * sizeof(double));
        }
        continue;
      }
      Job& j = *r->job;
      for (int i = 0; i < r->n; i++, k++) {
        vols[k] = j.svols[r->bone];
        sets[k] = j.set;
        frames[k] = j.sframes[r->frame];
        asc_compose_pose(r->store, j.pivots[r->bone].data(), r->manips + 6 * i, &poses[k * 6]);
      }
    }
    b.t = Clock::now();
    int rc;
    {
      std::lock_guard<std::mutex> sl(s->scene_mu);
      rc = asc_scene_submit_requests(s->scene, (int)n, vols.data(), sets.data(), frames.data(), poses.data(),
                                     &b.ticket);
    }
    lk.lock();
    if (rc != ASC_OK) {
      deliver(s, b.reqs, {}, {}, rc);
      continue;
    }
    if (s->launches.empty()) s->busy_from = std::max(b.t, s->busy_to);
    s->launches.push_back(std::move(b));
    s->cv_completer.notify_one();
  }
}

template <class T>
std::vector<unsigned char> bytes_of(const T* p, size_t n)
{
  return std::vector<unsigned char>((const unsigned char*)p, (const unsigned char*)p + n * sizeof(T));
}

}  // namespace

extern "C" {

ASC_API int asc_server_create(asc_context ctx, size_t hbm_budget, int max_active_jobs, asc_server* out)
{
  if (!ctx || !out || max_active_jobs < 1) return asc::fail(ASC_ERR_INVALID, "bad argument");
  asc_server s = new asc_server_s();
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->max_active = max_active_jobs;
  s->t0 = Clock::now();
  int device = 0;
  int rc = asc_context_device(ctx, &device);
  if (rc == ASC_OK) rc = asc_context_create(device, &s->up_ctx);
  if (rc == ASC_OK) rc = asc_scene_create(ctx, 1, 1, &s->scene);
  if (rc) {
    if (s->up_ctx) asc_context_destroy(s->up_ctx);
    delete s;
    return rc;
  }
  s->batcher = std::thread(batcher_main, s);
  s->completer = std::thread(completer_main, s);
  s->stager = std::thread(stager_main, s);
  *out = s;
  return ASC_OK;
}

ASC_API int asc_server_submit(asc_server s, const asc_job_desc* d, int* job_id)
{
  if (!s || !d || !job_id) return asc::fail(ASC_ERR_INVALID, "bad argument");
  if (d->n_volumes < 1 || d->n_cameras < 1 || d->n_cameras > 8 || d->n_frames < 1)
    return asc::fail(ASC_ERR_INVALID, "a job needs 1+ volumes, 1 to 8 cameras and 1+ frames");
  if (!d->volumes || !d->volume_dims || !d->voxel_mm || !d->initial_poses || !d->cameras || !d->frames ||
      !d->frame_dims)
    return asc::fail(ASC_ERR_INVALID, "null job field");
  if (d->frame_bps != 8 && d->frame_bps != 16 && d->frame_bps != 32)
    return asc::fail(ASC_ERR_UNSUPPORTED, "frame bit depth must be 8, 16 or 32");
  if (d->render_width < 1 || d->render_height < 1) return asc::fail(ASC_ERR_INVALID, "bad render size");
  if (d->options.particles < 1) return asc::fail(ASC_ERR_INVALID, "particles must be >= 1");
  std::unique_ptr<Job> j(new Job());
  j->n_bones = d->n_volumes;
  size_t dev = 0;
  for (int b = 0; b < d->n_volumes; b++) {
    const int* dm = d->volume_dims + 3 * b;
    if (dm[0] < 1 || dm[1] < 1 || dm[2] < 1) return asc::fail(ASC_ERR_INVALID, "bad volume shape");
    const size_t nv = (size_t)dm[0] * dm[1] * dm[2];
    j->vol_data.emplace_back((const uint16_t*)d->volumes[b], (const uint16_t*)d->volumes[b] + nv);
    j->vol_dims.push_back({dm[0], dm[1], dm[2]});
    j->vol_scale.push_back({d->voxel_mm[3 * b], d->voxel_mm[3 * b + 1], d->voxel_mm[3 * b + 2]});
    std::array<double, 6> p;
    std::memcpy(p.data(), d->initial_poses + 6 * b, sizeof(double) * 6);
    j->init_pose.push_back(p);
    // the padded linear copy plus three 8-corner BRICK copies, one per ray
    // axis, each covering at most (n + 8) cells per axis
    const size_t tex = d->storage == ASC_STORAGE_F32 ? 4 : 2;
    dev += (size_t)(dm[0] + 2) * (dm[1] + 2) * (dm[2] + 2) * tex;
    dev += (size_t)(dm[0] + 8) * (dm[1] + 8) * (dm[2] + 8) * tex * 8 * 3;
  }
  j->storage = d->storage;
  j->step = d->drr_step;
  j->intensity = d->drr_intensity;
  j->cutoff = d->drr_cutoff;
  j->cams.assign(d->cameras, d->cameras + d->n_cameras);
  j->RW = d->render_width;
  j->RH = d->render_height;
  if (d->n_drr_filters > 0) j->drr_chain.assign(d->drr_filters, d->drr_filters + d->n_drr_filters);
  if (d->n_rad_filters > 0) j->rad_chain.assign(d->rad_filters, d->rad_filters + d->n_rad_filters);
  j->n_frames = d->n_frames;
  j->bps = d->frame_bps;
  for (int c = 0; c < d->n_cameras; c++) j->frame_dims.push_back({d->frame_dims[2 * c], d->frame_dims[2 * c + 1]});
  for (int f = 0; f < d->n_frames; f++)
    for (int c = 0; c < d->n_cameras; c++) {
      const size_t n = (size_t)j->frame_dims[c][0] * j->frame_dims[c][1];
      const void* src = d->frames[(size_t)f * d->n_cameras + c];
      if (!src) return asc::fail(ASC_ERR_INVALID, "null frame");
      j->frames.push_back(d->frame_bps == 8    ? bytes_of((const uint8_t*)src, n)
                          : d->frame_bps == 16 ? bytes_of((const uint16_t*)src, n)
                                               : bytes_of((const float*)src, n));
      dev += n * 4;
    }
  // the job's share of the batch scratch, one PSO epoch of items: DRR
  // images, per-tile moment partials (9 doubles per 32 x 32 tile), results
  // and descriptors
  const size_t items = (size_t)d->options.particles * d->n_volumes * d->n_cameras;
  const size_t tiles = (size_t)((d->render_width + 31) / 32) * ((d->render_height + 31) / 32);
  j->data_bytes = dev;
  j->scratch_share = items * ((size_t)d->render_width * d->render_height * 4 + tiles * 72 + 80 + 256);
  j->device_bytes = j->data_bytes + j->scratch_share;
  dev = j->device_bytes;
  j->opt = d->options;
  j->guess = d->guess;
  j->poses.assign((size_t)d->n_frames * d->n_volumes * 6, 0.0);
  j->bone_done_ms.assign(d->n_volumes, std::vector<double>(d->n_frames, 0.0));
  j->frame_latency_ms.assign(d->n_frames, 0.0);
  std::lock_guard<std::mutex> lk(s->mu);
  if (dev > s->budget) return asc::fail(ASC_ERR_NOMEM, "job does not fit the server's device-memory budget");
  j->id = s->next_id++;
  j->arrival = Clock::now();
  *job_id = j->id;
  s->waiting.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->cv_stager.notify_one();
  return ASC_OK;
}

ASC_API int asc_server_wait(asc_server s, int job_id, double* poses, double* frame_latency_ms, asc_job_stats* stats)
{
  if (!s) return asc::fail(ASC_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return asc::fail(ASC_ERR_INVALID, "unknown job");
  Job* j = it->second.get();
  s->cv_done.wait(lk, [&] { return j->finished; });
  const int status = j->status;
  if (poses) std::memcpy(poses, j->poses.data(), j->poses.size() * sizeof(double));
  if (frame_latency_ms) std::memcpy(frame_latency_ms, j->frame_latency_ms.data(), j->n_frames * sizeof(double));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->queue_ms = j->admit_ms;
    double last = 0.0;
    for (const auto& bd : j->bone_done_ms) last = std::max(last, bd.back());
    stats->total_ms = last;
    stats->evaluations = j->evaluations;
    stats->frames = j->n_frames;
  }
  s->jobs.erase(it);
  if (status != ASC_OK) return asc::fail(status, "job failed");
  return ASC_OK;
}

ASC_API int asc_server_set_pipeline(asc_server s, int depth)
{
  if (!s || depth < 1 || depth > 2) return asc::fail(ASC_ERR_INVALID, "pipeline depth is 1 or 2");
  std::lock_guard<std::mutex> lk(s->mu);
  s->depth = depth;
  s->cv_batcher.notify_one();
  return ASC_OK;
}

ASC_API int asc_server_lock_scene(asc_server s, asc_scene* scene)
{
  if (!s || !scene) return asc::fail(ASC_ERR_INVALID, "bad argument");
  s->scene_mu.lock();
  s->scene_owner = std::this_thread::get_id();
  *scene = s->scene;
  return ASC_OK;
}

ASC_API int asc_server_unlock_scene(asc_server s)
{
  if (!s) return asc::fail(ASC_ERR_INVALID, "null server");
  if (s->scene_owner != std::this_thread::get_id())
    return asc::fail(ASC_ERR_INVALID, "the scene is not locked by this thread");
  s->scene_owner = std::thread::id();
  s->scene_mu.unlock();
  return ASC_OK;
}

ASC_API int asc_server_stream_begin(asc_server s)
{
  if (!s) return asc::fail(ASC_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->streams++;
  return ASC_OK;
}

ASC_API int asc_server_stream_end(asc_server s)
{
  if (!s) return asc::fail(ASC_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->streams <= 0) return asc::fail(ASC_ERR_INVALID, "asc_server_stream_end without asc_server_stream_begin");
  s->streams--;
  // the batcher may be waiting for this stream's next request
  s->cv_batcher.notify_one();
  return ASC_OK;
}

ASC_API int asc_server_eval_requests(asc_server s, int n, const int* volumes, const int* sets, const int* frames,
                                     const double* xyzypr, double* costs, double* per_camera)
{
  if (!s || n < 1 || !volumes || !sets || !xyzypr || !costs) return asc::fail(ASC_ERR_INVALID, "bad argument");
  if (s->scene_owner == std::this_thread::get_id())
    return asc::fail(ASC_ERR_INVALID, "asc_server_eval_requests while this thread holds the scene lock");
  Request r;
  r.job = nullptr;
  r.bone = -1;
  r.frame = -1;
  r.manips = nullptr;
  r.n = n;
  r.store = nullptr;
  r.costs = costs;
  r.ext_vols = volumes;
  r.ext_sets = sets;
  r.ext_frames = frames;
  r.ext_poses = xyzypr;
  r.per_camera = per_camera;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->stop) return asc::fail(ASC_ERR_INVALID, "server is stopping");
  s->pending.push_back(&r);
  s->cv_batcher.notify_one();
  s->cv_req.wait(lk, [&] { return r.done; });
  return r.status;
}

ASC_API int asc_server_stats(asc_server s, asc_serve_stats* out)
{
  if (!s || !out) return asc::fail(ASC_ERR_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->jobs_done = s->jobs_done;
  out->frames_done = s->frames_done;
  out->batches = s->batches;
  out->requests = s->requests;
  out->evaluations = s->evaluations;
  out->busy_ms = s->busy_ms;
  out->elapsed_ms = ms_since(s->t0, Clock::now());
  std::vector<double> l = s->latencies_ms;
  std::sort(l.begin(), l.end());
  auto q = [&](double p) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
  out->latency_p50_ms = q(0.50);
  out->latency_p99_ms = q(0.99);
  out->latency_max_ms = l.empty() ? 0.0 : l.back();
  out->active_jobs = s->active;
  out->waiting_jobs = (int)s->waiting.size();
  out->device_bytes_in_use = s->used_data + std::max(s->scene_scratch, s->used_share);
  return ASC_OK;
}

ASC_API int asc_server_destroy(asc_server s)
{
  if (!s) return ASC_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // jobs still queued are dropped; running ones finish first
    for (Job* j : s->waiting) {
      j->admitted = true;
      j->finished = true;
      j->status = ASC_ERR_INVALID;
    }
    s->waiting.clear();
    s->cv_done.wait(lk, [&] {
      for (auto& kv : s->jobs)
        if (kv.second->admitted && !kv.second->finished) return false;
      return true;
    });
    s->stop = true;
    s->stop_stager = true;
    s->cv_batcher.notify_one();
    s->cv_stager.notify_one();
  }
  s->batcher.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_completer = true;
    s->cv_completer.notify_one();
  }
  s->completer.join();
  s->stager.join();
  asc_scene_destroy(s->scene);
  asc_context_destroy(s->up_ctx);
  delete s;
  return ASC_OK;
}

}  // extern "C"
