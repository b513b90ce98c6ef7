// Remote control in the reference's socket protocol, version 2
// (autoscoper/src/net/Socket.cpp:53-444; its client is PyAutoscoper,
// scripts/python/PyAutoscoper/connect.py): a session that holds what the
// GUI holds for a trial (cameras, volumes, radiograph frames, filter
// settings, background threshold and the per-volume pose curves of
// KeyCurve.hpp) and answers the 18 messages on 127.0.0.1. Cost evaluation
// and optimisation run on this library's fused scene.
//
// A message is one read of what the client sent: a type byte, then the
// little-endian fields the type defines; every answer starts with the type
// byte (0 on failure).
#include <algorithm>
#include <arpa/inet.h>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <errno.h>
#include <limits>
#include <map>
#include <mutex>
#include <netinet/in.h>
#include <poll.h>
#include <string>
#include <sys/socket.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

#include "geometry.hpp"

namespace asc {
namespace {

// ---------------------------------------------------------------------------
// Pose curves (KeyCurve.hpp): x, y, z are float curves through their keys
// with SMOOTH tangents (cubic Hermite segments); the orientation is a Quatf
// curve evaluated by slerp. Without keys a float curve is 0 and the
// orientation the identity; before the first key and after the last the
// nearest key holds.
// ---------------------------------------------------------------------------
struct FloatKey {
  float value = 0, in_tangent = 0, out_tangent = 0;
  float a = 0, b = 0, c = 0, d = 0;
};

struct FloatCurve {
  std::map<int, FloatKey> keys;

  // KeyCurve::update_curve for the key at `it` (SMOOTH, unlocked tangents)
  void update(std::map<int, FloatKey>::iterator it)
  {
    auto prev = it, next = it;
    if (prev != keys.begin()) --prev;
    if (next != std::prev(keys.end())) ++next;
    const float time = (float)it->first;
    const float next_val = next->second.value, prev_val = prev->second.value;
    const float tan = (next_val - prev_val) / (float)(next->first - prev->first);
    it->second.in_tangent = tan;
    it->second.out_tangent = tan;
    if (next != it) {
      FloatKey& k = it->second;
      const float D1 = k.out_tangent * (next->first - time);
      const float D2 = next->second.in_tangent * (next->first - time);
      k.a = k.value;
      k.b = D1;
      k.c = 3.0f * (next_val - k.value) - 2.0f * D1 - D2;
      k.d = 2.0f * (k.value - next_val) + D1 + D2;
    }
  }

  // KeyCurve::insert + key_changed: the next key, this one, and the two
  // before it are brought up to date, the next one first
  void insert(int time, float value)
  {
    auto it = keys.insert_or_assign(time, FloatKey{}).first;
    it->second.value = value;
    auto next = it;
    if (next != std::prev(keys.end())) ++next;
    auto prev = it;
    if (prev != keys.begin()) --prev;
    auto prev2 = prev;
    if (prev2 != keys.begin()) --prev2;
    if (next != it) update(next);
    update(it);
    if (prev != it) update(prev);
    if (prev2 != prev) update(prev2);
  }

  bool has(int time) const { return keys.count(time) != 0; }

  float operator()(float time) const
  {
    if (keys.empty()) return 0.0f;
    // the map's key type is int: the float time converts (truncates) to it
    auto p1 = keys.upper_bound((int)time);
    if (p1 == keys.begin()) return p1->second.value;
    auto p0 = std::prev(p1);
    if (p1 == keys.end()) return p0->second.value;
    const float t = (time - p0->first) / (float)(p1->first - p0->first);
    const FloatKey& k = p0->second;
    return k.a + t * k.b + t * t * k.c + t * t * t * k.d;
  }
};

struct Quatf {
  float w = 1, x = 0, y = 0, z = 0;
};

float qdot(const Quatf& p, const Quatf& q) { return p.w * q.w + p.x * q.x + p.y * q.y + p.z * q.z; }

// the reference's float helpers call the double math functions
// (Quaternion.hpp includes <cmath> and calls sqrt/acos/sin unqualified)
Quatf qunit(const Quatf& q)
{
  const float l = (float)std::sqrt((double)qdot(q, q));
  return Quatf{q.w / l, q.x / l, q.y / l, q.z / l};
}

// slerp(p, q, t) of Quaternion.hpp:380-403
Quatf qslerp(const Quatf& p, const Quatf& q, float t)
{
  const float eps = std::numeric_limits<float>::epsilon();
  if (t < eps) return p;
  if (t > 1.0f - eps) return q;
  const float cos_rad = qdot(qunit(p), qunit(q));
  float rad = (float)std::acos((double)cos_rad);
  if (rad < eps) {  // nlerp
    if (qdot(p, q) > 0.0f)
      return qunit(Quatf{p.w + t * (q.w - p.w), p.x + t * (q.x - p.x), p.y + t * (q.y - p.y), p.z + t * (q.z - p.z)});
    return qunit(Quatf{p.w - t * (q.w + p.w), p.x - t * (q.x + p.x), p.y - t * (q.y + p.y), p.z - t * (q.z + p.z)});
  }
  float s0, s1;
  if (cos_rad > 0.0f) {
    s0 = (float)std::sin((double)(rad * (1.0f - t)));
    s1 = (float)std::sin((double)(rad * t));
  } else {
    rad = (float)(M_PI - (double)rad);  // M_PI is a double
    s0 = (float)std::sin((double)(rad * (1.0f - t)));
    s1 = -(float)std::sin((double)(rad * t));
  }
  return qunit(Quatf{s0 * p.w + s1 * q.w, s0 * p.x + s1 * q.x, s0 * p.y + s1 * q.y, s0 * p.z + s1 * q.z});
}

struct QuatCurve {
  std::map<int, Quatf> keys;
  void insert(int time, const Quatf& q) { keys[time] = q; }
  bool has(int time) const { return keys.count(time) != 0; }
  Quatf operator()(float time) const
  {
    if (keys.empty()) return Quatf{};
    auto p1 = keys.upper_bound((int)time);
    if (p1 == keys.begin()) return p1->second;
    auto p0 = std::prev(p1);
    if (p1 == keys.end()) return p0->second;
    const float t = (time - p0->first) / (float)(p1->first - p0->first);
    return qslerp(p0->second, p1->second, t);
  }
};

struct PoseCurves {
  FloatCurve x, y, z;
  QuatCurve q;

  void clear()
  {
    x.keys.clear();
    y.keys.clear();
    z.keys.clear();
    q.keys.clear();
  }
  bool keyed(int f) const { return x.has(f) || y.has(f) || z.has(f) || q.has(f); }
  // the curves at frame f as a pose store {x, y, z, qw, qx, qy, qz}
  void store_at(int f, float* s) const
  {
    const Quatf r = q((float)f);
    s[0] = x((float)f);
    s[1] = y((float)f);
    s[2] = z((float)f);
    s[3] = r.w;
    s[4] = r.x;
    s[5] = r.y;
    s[6] = r.z;
  }
  void insert_store(int f, const float* s)
  {
    x.insert(f, s[0]);
    y.insert(f, s[1]);
    z.insert(f, s[2]);
    q.insert(f, Quatf{s[3], s[4], s[5], s[6]});
  }
  // xyzypr: the float x, y, z and Quatf(yaw, pitch, roll), as the GUI's
  // setPose and load_tracking_results insert them
  void insert_pose(int f, const double* p)
  {
    float s[7];
    asc_pose_store_set(s, p);
    insert_store(f, s);
  }
  void pose_at(int f, double* p) const
  {
    float s[7];
    store_at(f, s);
    asc_pose_store_get(s, p);
  }
};

// ---------------------------------------------------------------------------
// The session
// ---------------------------------------------------------------------------
struct Bone {
  std::vector<uint8_t> data;
  int dims[3] = {0, 0, 0}, bps = 16, flip[3] = {0, 0, 0};
  double voxel[3] = {1, 1, 1};
  PoseCurves curves;
  bool visible = true;
  asc_volume dev = nullptr;
  asc_volume_desc desc;
  double pivot[12];
};

struct CamState {
  asc_filter_settings vie;  // filters and DRR renderer values applied to this camera
  asc_image rad = nullptr;  // the current frame
  asc_image bg = nullptr;
  int rad_frame = -1;
};

struct Session {
  asc_context ctx = nullptr;
  asc_trial trial = nullptr;
  asc_trial_info info;
  std::vector<asc_camera> cams;
  std::vector<Bone> bones;
  std::vector<CamState> cam_state;
  int frame = 0;
  int guess = 0;
  double bg_threshold = -1.0;
  asc_scene scene = nullptr;
  bool scene_dirty = true;
  // shared mode: the scene is the registration server's, and evaluations are
  // requests to its batcher (several sessions share launches). The session
  // owns only its entries in that scene.
  asc_server srv = nullptr;
  std::vector<int> vol_slot, cam_slot;  // session volume / camera -> scene slot
  int set = -1;                         // the session's camera set (shared mode)
  int current_volume = 0;       // Trial::current_volume, set by optimizeFrame
  std::string full_drr_dir = ".";  // where message 12 writes its images
  int full_drr_count = 0;        // save_full_drr's running file number

  ~Session() { close_trial(); }

  void close_trial()
  {
    if (srv && set >= 0) {
      // out of the shared scene before the device copies go (no launch
      // in flight reads them: the session's own requests are all done)
      asc_scene sc = nullptr;
      asc_server_lock_scene(srv, &sc);
      asc_scene_remove_camera_set(sc, set);
      for (int c : cam_slot) asc_scene_remove_camera(sc, c);
      for (int v : vol_slot)
        if (v >= 0) asc_scene_remove_volume(sc, v);
      asc_server_unlock_scene(srv);
    } else if (scene && !srv) {
      asc_scene_destroy(scene);
    }
    scene = nullptr;
    set = -1;
    vol_slot.clear();
    cam_slot.clear();
    for (auto& c : cam_state) {
      if (c.rad) asc_image_destroy(c.rad);
      if (c.bg) asc_image_destroy(c.bg);
    }
    cam_state.clear();
    for (auto& b : bones)
      if (b.dev) asc_volume_destroy(b.dev);
    bones.clear();
    cams.clear();
    if (trial) asc_trial_close(trial);
    trial = nullptr;
    frame = 0;
    bg_threshold = -1.0;
  }
};

void default_settings(asc_filter_settings& s)
{
  std::memset(&s, 0, sizeof s);
  // the GUI's DRR renderer spin boxes (ModelViewTreeWidgetItem.cpp:88-90)
  s.sample_distance = 0.62;
  s.ray_intensity = 0.49;
  s.cutoff_fraction = 0.0;
  s.drr_step = (float)std::exp(7 * 0.62 - 5);
  s.drr_intensity = (float)std::exp(15 * 0.49 - 5);
}

int open_trial(Session& S, const std::string& path)
{
  asc_trial t = nullptr;
  int rc = asc_trial_open(path.c_str(), &t);
  if (rc) return rc;
  S.close_trial();
  S.trial = t;
  asc_trial_get_info(t, &S.info);
  S.cams.resize(S.info.n_cameras);
  S.cam_state.resize(S.info.n_cameras);
  for (int c = 0; c < S.info.n_cameras; c++) {
    asc_trial_camera(t, c, &S.cams[c]);
    default_settings(S.cam_state[c].vie);
  }
  S.bones.resize(S.info.n_volumes);
  for (int v = 0; v < S.info.n_volumes; v++) {
    Bone& b = S.bones[v];
    if ((rc = asc_trial_volume_info(t, v, b.dims, &b.bps, b.voxel, b.flip))) return rc;
    b.data.resize((size_t)b.dims[0] * b.dims[1] * b.dims[2] * (b.bps / 8));
    if ((rc = asc_trial_volume_read(t, v, b.data.data(), b.data.size()))) return rc;
  }
  S.scene_dirty = true;
  return ASC_OK;
}

// The session's scene for the length of one operation: its own, or the
// registration server's under the server's scene lock (shared mode).
struct SceneUse {
  Session& S;
  explicit SceneUse(Session& s) : S(s)
  {
    if (S.srv) asc_server_lock_scene(S.srv, &S.scene);
  }
  ~SceneUse()
  {
    if (S.srv) asc_server_unlock_scene(S.srv);
  }
};

// Scene entries and radiographs, with the scene in use.
int update_scene(Session& S)
{
  int rc;
  if (S.cam_slot.empty()) {
    for (size_t c = 0; c < S.cams.size(); c++) {
      int idx;
      if ((rc = asc_scene_add_camera(S.scene, &S.cams[c], nullptr, &idx))) return rc;
      S.cam_slot.push_back(idx);
    }
    if (S.srv &&
        (rc = asc_scene_add_camera_set(S.scene, S.cam_slot.data(), (int)S.cam_slot.size(), S.info.render_width,
                                       S.info.render_height, &S.set)))
      return rc;
    S.vol_slot.assign(S.bones.size(), -1);
    S.scene_dirty = true;
  }
  if (S.scene_dirty) {
    // volumes with the DRR renderer values of camera 0 (the scene keeps one
    // set per volume), filters per camera
    const asc_filter_settings& v0 = S.cam_state[0].vie;
    for (size_t v = 0; v < S.bones.size(); v++) {
      Bone& b = S.bones[v];
      float cutoff;
      asc_drr_cutoff(v0.cutoff_fraction, &b.desc, &cutoff);
      if (S.vol_slot[v] >= 0) asc_scene_remove_volume(S.scene, S.vol_slot[v]);
      if ((rc = asc_scene_add_volume(S.scene, b.dev, v0.drr_step, v0.drr_intensity, cutoff, b.pivot, &S.vol_slot[v])))
        return rc;
    }
    for (size_t c = 0; c < S.cams.size(); c++) {
      const asc_filter_settings& s = S.cam_state[c].vie;
      if ((rc = asc_scene_set_filters(S.scene, S.cam_slot[c], 0, s.drr_filters, s.n_drr_filters))) return rc;
      if ((rc = asc_scene_set_filters(S.scene, S.cam_slot[c], 1, s.rad_filters, s.n_rad_filters))) return rc;
    }
    S.scene_dirty = false;
  }
  // radiographs of the current frame (Tracker.cpp:376-381)
  for (size_t c = 0; c < S.cams.size(); c++) {
    CamState& cs = S.cam_state[c];
    if (cs.rad_frame == S.frame && cs.rad) continue;
    int dims[2], bps;
    if ((rc = asc_trial_frame_info(S.trial, (int)c, S.frame, dims, &bps))) return rc;
    std::vector<uint8_t> px((size_t)dims[0] * dims[1] * (bps / 8));
    if ((rc = asc_trial_frame_read(S.trial, (int)c, S.frame, px.data(), px.size()))) return rc;
    asc_image img = nullptr;
    if ((rc = asc_image_create(S.ctx, px.data(), dims[0], dims[1], bps, &img))) return rc;
    if ((rc = asc_scene_set_rad(S.scene, S.cam_slot[c], img))) {
      asc_image_destroy(img);
      return rc;
    }
    if (cs.rad) asc_image_destroy(cs.rad);
    cs.rad = img;
    cs.rad_frame = S.frame;
  }
  return ASC_OK;
}

// Device side of the session, created on the first request that needs it.
int ensure_scene(Session& S)
{
  if (!S.trial) return fail(ASC_ERR_INVALID, "no trial loaded");
  if (!S.ctx) return fail(ASC_ERR_NO_DEVICE, "the remote session has no device context");
  int rc;
  if (S.cam_slot.empty()) {
    for (Bone& b : S.bones) {
      if (b.dev) continue;
      if (b.bps != 8 && b.bps != 16) return fail(ASC_ERR_UNSUPPORTED, "volumes must be 8 or 16 bit");
      if ((rc = asc_volume_create(S.ctx, b.data.data(), b.dims[0], b.dims[1], b.dims[2], b.bps, b.flip, b.voxel,
                                  ASC_STORAGE_U16, &b.dev, &b.desc)))
        return rc;
      // Tracker.cpp:281-290: the pivot is the volume centre
      const double c[6] = {b.desc.trans_center[0], b.desc.trans_center[1], b.desc.trans_center[2], 0, 0, 0};
      asc_frame_from_xyzypr(c, b.pivot);
    }
    if (!S.srv && !S.scene &&
        (rc = asc_scene_create(S.ctx, S.info.render_width, S.info.render_height, &S.scene)))
      return rc;
  }
  SceneUse use(S);
  return update_scene(S);
}

// n poses of one volume at the current frame -> costs and per-camera
// values (n x 8): through the server's batcher in shared mode, else the
// session's own scene.
int eval_poses(Session& S, int volume, const double* poses, int n, double* costs, double* per8)
{
  std::vector<int> vols(n, S.vol_slot[volume]);
  if (S.srv) {
    std::vector<int> sets(n, S.set);
    return asc_server_eval_requests(S.srv, n, vols.data(), sets.data(), nullptr, poses, costs, per8);
  }
  return asc_scene_eval_poses(S.scene, vols.data(), poses, n, costs, per8);
}

// AutoscoperMainWindow::setBackground (:876-894) with Video::create_background_image.
int set_background(Session& S, double threshold)
{
  int rc;
  if ((rc = ensure_scene(S))) return rc;
  for (size_t c = 0; c < S.cams.size(); c++) {
    CamState& cs = S.cam_state[c];
    if (!cs.bg) {
      int dims[2], bps;
      if ((rc = asc_trial_frame_info(S.trial, (int)c, 0, dims, &bps))) return rc;
      std::vector<float> bg((size_t)dims[0] * dims[1]);
      if ((rc = asc_trial_background(S.trial, (int)c, bg.data(), bg.size()))) return rc;
      if ((rc = asc_image_create(S.ctx, bg.data(), dims[0], dims[1], 32, &cs.bg))) return rc;
    }
    SceneUse use(S);
    if ((rc = asc_scene_set_background(S.scene, S.cam_slot[c], cs.bg, (float)threshold))) return rc;
  }
  S.bg_threshold = threshold;
  return ASC_OK;
}

unsigned seed_from_env()
{
  // initializeRandom (PSO.cpp:131-149): Autoscoper_RANDOM_SEED, else the clock
  if (const char* s = std::getenv("Autoscoper_RANDOM_SEED")) return (unsigned)std::atoi(s);
  return (unsigned)std::time(nullptr);
}

// The optimiser's cost in shared mode: manipulator offsets composed with
// the current store and the volume's pivot (Tracker::minimizationFunc,
// Tracker.cpp:775-827, as the scene's own path composes them), evaluated
// by the server's batcher.
struct SharedCost {
  Session* S;
  int volume;
  const float* store;
};

int shared_cost(const double* manips, int n, double* costs, void* user)
{
  const SharedCost* c = (const SharedCost*)user;
  Session& S = *c->S;
  std::vector<double> poses((size_t)n * 6), per((size_t)n * 8);
  for (int i = 0; i < n; i++)
    asc_compose_pose(c->store, S.bones[c->volume].pivot, manips + 6 * i, &poses[(size_t)i * 6]);
  return eval_poses(S, c->volume, poses.data(), n, costs, per.data());
}

// Tracker::optimize (Tracker.cpp:337-580) for one volume and frame.
int optimize(Session& S, int volume, int frame, int dframe, int repeats, int method, unsigned max_iter,
             double min_limit, double max_limit, int cf_model, unsigned stall_iter)
{
  if (!S.trial || volume < 0 || volume >= (int)S.bones.size())
    return fail(ASC_ERR_INVALID, "optimize: bad volume");
  if (frame < 0 || frame >= S.info.n_frames) return fail(ASC_ERR_INVALID, "Tracker::optimize(): Invalid frame.");
  S.current_volume = volume;  // AutoscoperMainWindow::optimizeFrame sets it first
  if (cf_model != 0) return fail(ASC_ERR_UNSUPPORTED, "only the NCC cost (cf_model 0) is built");
  S.frame = frame;
  int rc;
  if ((rc = ensure_scene(S))) return rc;
  PoseCurves& cv = S.bones[volume].curves;
  const int behind = dframe > 0 ? frame : S.info.n_frames - frame - 1;
  if (S.guess == 2 && behind > 1) {
    double p1[6], p2[6], a[12], b[12], x[12], out[6];
    cv.pose_at(frame - 2 * dframe, p1);
    cv.pose_at(frame - dframe, p2);
    asc_frame_from_xyzypr(p1, a);
    asc_frame_from_xyzypr(p2, b);
    asc_frame_linear_extrap(a, b, x);
    asc_frame_to_xyzypr(x, out);
    cv.insert_pose(frame, out);
  } else if (S.guess == 1 && behind > 0) {
    double p[6];
    cv.pose_at(frame - dframe, p);
    cv.insert_pose(frame, p);
  }
  asc_optimize_options o;
  std::memset(&o, 0, sizeof o);
  o.method = method;
  o.max_iter = max_iter;
  o.max_stall = stall_iter;
  o.min_limit = min_limit;
  o.max_limit = max_limit;
  for (int k = 0; k < 6; k++) o.offsets[k] = S.info.offsets[k];
  o.seed = seed_from_env();
  o.particles = 100;
  o.volume = S.vol_slot[volume];
  for (int r = 0; r < repeats; r++) {
    float store[7];
    cv.store_at(frame, store);
    asc_optimize_stats st;
    if (S.srv) {
      // announced, so that the optimisations of other sessions share its launches
      SharedCost c{&S, volume, store};
      if ((rc = asc_server_stream_begin(S.srv))) return rc;
      rc = asc_optimize_frame_fn(shared_cost, &c, store, S.bones[volume].pivot, &o, &st);
      const int rc_end = asc_server_stream_end(S.srv);
      if (!rc) rc = rc_end;
    } else {
      rc = asc_optimize_frame(S.scene, store, &o, &st);
    }
    if (rc) return rc;
    cv.insert_store(frame, store);
  }
  return ASC_OK;
}

// Tracker::getImageData (Tracker.cpp:865-921): the whole camera view at
// the trial's render size, as 8-bit radiograph, DRR and mask images.
// The masked DRR, radiograph and mask of one volume over the camera's whole
// image plane, as Tracker::getImageData (Tracker.cpp:829-863) and
// Tracker::getFullDRR (Tracker.cpp:1035-1086) render them: host planes
// {rad, drr, mask}, each w x h.
int render_full_view(Session& S, int volume, int camera, const double* pose, unsigned& w, unsigned& h,
                     std::vector<float>& host)
{
  int rc;
  if ((rc = ensure_scene(S))) return rc;
  if (volume < 0 || volume >= (int)S.bones.size() || camera < 0 || camera >= (int)S.cams.size())
    return fail(ASC_ERR_INVALID, "getImageData: bad volume or camera");
  const asc_camera& cam = S.cams[camera];
  Bone& b = S.bones[volume];
  const CamState& cs = S.cam_state[camera];
  const asc_filter_settings& v0 = S.cam_state[0].vie;
  // modelview = camera^-1 * pose; the DRR renderer gets its inverse
  double cf[12], pf[12], ci[12], mv[12], mvi[12];
  std::memcpy(cf, cam.rotation, 9 * sizeof(double));
  std::memcpy(cf + 9, cam.translation, 3 * sizeof(double));
  asc_frame_from_xyzypr(pose, pf);
  asc_frame_inverse(cf, ci);
  asc_frame_mul(ci, pf, mv);
  asc_frame_inverse(mv, mvi);
  double m16[16];
  for (int r = 0; r < 3; r++) {
    for (int c = 0; c < 3; c++) m16[4 * r + c] = mvi[3 * c + r];
    m16[4 * r + 3] = mvi[9 + r];
  }
  m16[12] = m16[13] = m16[14] = 0.0;
  m16[15] = 1.0;
  float imv16[16];
  if ((rc = asc_inv_model_view(m16, b.desc.inv_scale, b.desc.inv_trans, imv16))) return rc;
  const double vpd[4] = {-cam.viewport[2] / 2, -cam.viewport[3] / 2, cam.viewport[2], cam.viewport[3]};
  const float vp[4] = {(float)vpd[0], (float)vpd[1], (float)vpd[2], (float)vpd[3]};
  const float plane[4] = {(float)cam.viewport[0], (float)cam.viewport[1], (float)cam.viewport[2],
                          (float)cam.viewport[3]};
  w = (unsigned)(vpd[2] * S.info.render_width / cam.viewport[2]);
  h = (unsigned)(vpd[3] * S.info.render_height / cam.viewport[3]);
  const size_t n = (size_t)w * h;
  float* buf = nullptr;
  if ((rc = asc_malloc(S.ctx, 6 * n * sizeof(float), (void**)&buf))) return rc;
  float *drr = buf, *rad = buf + n, *tmp = buf + 2 * n, *scratch = buf + 3 * n, *bgm = buf + 4 * n,
        *mask = buf + 5 * n;
  float cutoff;
  asc_drr_cutoff(v0.cutoff_fraction, &b.desc, &cutoff);
  rc = asc_drr_render(S.ctx, b.dev, imv16, vp, v0.drr_step, v0.drr_intensity, cutoff, tmp, w, h);
  if (!rc) rc = asc_filter_chain(S.ctx, cs.vie.drr_filters, cs.vie.n_drr_filters, tmp, drr, scratch, (int)w, (int)h);
  if (!rc) rc = asc_rad_render(S.ctx, cs.rad, plane, vp, tmp, w, h);
  if (!rc) rc = asc_filter_chain(S.ctx, cs.vie.rad_filters, cs.vie.n_rad_filters, tmp, rad, scratch, (int)w, (int)h);
  const bool use_bg = S.bg_threshold >= 0 && cs.bg;
  if (!rc && use_bg)
    rc = asc_background_render(S.ctx, cs.bg, plane, vp, (float)S.bg_threshold, bgm, w, h);
  if (!rc) rc = asc_mask_multiply(S.ctx, drr, rad, use_bg ? bgm : nullptr, mask, w, h);
  host.assign(3 * n, 0.0f);
  if (!rc) rc = asc_synchronize(S.ctx);
  if (!rc) rc = asc_memcpy_d2h(S.ctx, host.data(), rad, n * sizeof(float));
  if (!rc) rc = asc_memcpy_d2h(S.ctx, host.data() + n, drr, n * sizeof(float));
  if (!rc) rc = asc_memcpy_d2h(S.ctx, host.data() + 2 * n, mask, n * sizeof(float));
  asc_free(S.ctx, buf);
  return rc;
}

int image_data(Session& S, int volume, int camera, const double* pose, unsigned& w, unsigned& h,
               std::vector<uint8_t>& out)
{
  std::vector<float> host;
  const int rc = render_full_view(S, volume, camera, pose, w, h, host);
  if (rc) return rc;
  // get_image (Tracker.cpp:844-863): (unsigned char)(255 * value)
  out.resize(host.size());
  for (size_t i = 0; i < host.size(); i++) out[i] = (uint8_t)(int)(255 * host[i]);
  return ASC_OK;
}

// Tracker::getFullDRR (Tracker.cpp:1035-1086) of the current volume at its
// pose curves' value at the current frame: per camera, the masked DRR over
// the whole image plane, written by save_full_drr as image_camNN.pgm, NN
// counting up over the session.
int save_full_drr(Session& S)
{
  if (!S.trial) return fail(ASC_ERR_INVALID, "saveFullDRR: no trial");
  const int v = S.current_volume;
  if (v < 0 || v >= (int)S.bones.size()) return fail(ASC_ERR_INVALID, "saveFullDRR: bad volume");
  double pose[6];
  S.bones[v].curves.pose_at(S.frame, pose);
  for (int c = 0; c < (int)S.cams.size(); c++) {
    unsigned w, h;
    std::vector<float> host;
    int rc = render_full_view(S, v, c, pose, w, h, host);
    if (rc) return rc;
    char name[32];
    std::snprintf(name, sizeof name, "image_cam%02d.pgm", S.full_drr_count++);
    const std::string path = S.full_drr_dir + "/" + name;
    if ((rc = asc_write_full_drr_pgm(path.c_str(), host.data() + (size_t)w * h, (int)w, (int)h))) return rc;
  }
  return ASC_OK;
}

// ---------------------------------------------------------------------------
// Messages
// ---------------------------------------------------------------------------
struct Reader {
  const uint8_t* p;
  size_t n, at = 1;
  bool ok = true;
  template <class T>
  T get()
  {
    T v{};
    if (at + sizeof(T) > n) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  std::string rest() { return at <= n ? std::string((const char*)p + at, n - at) : std::string(); }
};

template <class T>
void put(std::vector<uint8_t>& out, const T& v)
{
  const uint8_t* b = (const uint8_t*)&v;
  out.insert(out.end(), b, b + sizeof(T));
}

int tra_flags(int matrix, int rows, int commas, int cm, int rad)
{
  return (matrix ? ASC_TRA_MATRIX : 0) | (rows ? ASC_TRA_ROWS : 0) | (commas ? ASC_TRA_COMMAS : 0) |
         (cm ? ASC_TRA_CM : 0) | (rad ? ASC_TRA_RADIANS : 0);
}

bool file_exists(const std::string& p)
{
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// Answers one message; returns false when the connection is to be closed.
bool handle_message(Session& S, const uint8_t* msg, size_t len, std::vector<uint8_t>& out)
{
  out.clear();
  if (len == 0) return true;
  const uint8_t type = msg[0];
  Reader r{msg, len};
  auto reply = [&](bool ok) { out.push_back(ok ? type : 0); };
  switch (type) {
    case 0: out.push_back(0); break;  // connection test
    case 1: {                         // load trial
      const std::string path = r.rest();
      reply(file_exists(path) && open_trial(S, path) == ASC_OK);
      break;
    }
    case 2:    // load tracking data
    case 3: {  // save tracking data
      const int volume = r.get<int32_t>(), matrix = r.get<int32_t>(), rows = r.get<int32_t>(),
                commas = r.get<int32_t>(), cm = r.get<int32_t>(), rad = r.get<int32_t>(),
                interpolate = r.get<int32_t>();
      const std::string path = r.rest();
      if (!r.ok || !S.trial) {
        reply(false);
        break;
      }
      const int B = (int)S.bones.size(), F = S.info.n_frames;
      const int v0 = volume == -1 ? 0 : volume, v1 = volume == -1 ? B : volume + 1;
      if (v0 < 0 || v1 > B) {
        reply(false);
        break;
      }
      const int flags = tra_flags(matrix, rows, commas, cm, rad);
      const int nv = v1 - v0;
      std::vector<double> poses((size_t)F * nv * 6);
      if (type == 2) {
        int got = 0;
        if (!file_exists(path) || asc_read_tra(path.c_str(), F, nv, flags, poses.data(), &got)) {
          reply(false);
          break;
        }
        for (int v = v0; v < v1; v++) S.bones[v].curves.clear();
        for (int f = 0; f < got; f++)
          for (int v = v0; v < v1; v++) {
            const double* p = &poses[((size_t)f * nv + (v - v0)) * 6];
            if (!std::isnan(p[0])) S.bones[v].curves.insert_pose(f, p);
          }
        reply(true);
      } else {
        for (int f = 0; f < F; f++)
          for (int v = v0; v < v1; v++) {
            double* p = &poses[((size_t)f * nv + (v - v0)) * 6];
            if (!interpolate && !S.bones[v].curves.keyed(f))
              std::fill(p, p + 6, std::numeric_limits<double>::quiet_NaN());
            else
              S.bones[v].curves.pose_at(f, p);
          }
        // the parent directory must exist (it is created) and be writable
        const size_t slash = path.find_last_of('/');
        if (slash != std::string::npos && slash > 0) {
          std::string dir = path.substr(0, slash), part;
          for (size_t i = 1; i <= dir.size(); i++)
            if (i == dir.size() || dir[i] == '/') ::mkdir(dir.substr(0, i).c_str(), 0777);
        }
        reply(asc_write_tra(path.c_str(), poses.data(), F, nv, flags) == ASC_OK);
      }
      break;
    }
    case 4: {  // load filter settings for a camera (-1: all)
      const int camera = r.get<int32_t>();
      const std::string path = r.rest();
      asc_filter_settings s;
      if (!r.ok || !S.trial || !file_exists(path) || asc_read_filter_settings(path.c_str(), &s)) {
        reply(false);
        break;
      }
      const int c0 = camera == -1 ? 0 : camera, c1 = camera == -1 ? (int)S.cams.size() : camera + 1;
      if (c0 < 0 || c1 > (int)S.cams.size()) {
        reply(false);
        break;
      }
      for (int c = c0; c < c1; c++) {
        asc_filter_settings& cur = S.cam_state[c].vie;
        // sections absent from the file leave the camera's settings as they are
        if (s.has_drr_renderer) {
          cur.sample_distance = s.sample_distance;
          cur.ray_intensity = s.ray_intensity;
          cur.cutoff_fraction = s.cutoff_fraction;
          cur.drr_step = s.drr_step;
          cur.drr_intensity = s.drr_intensity;
        }
        if (s.has_drr_filters) {
          cur.n_drr_filters = s.n_drr_filters;
          std::copy(s.drr_filters, s.drr_filters + s.n_drr_filters, cur.drr_filters);
        }
        if (s.has_rad_filters) {
          cur.n_rad_filters = s.n_rad_filters;
          std::copy(s.rad_filters, s.rad_filters + s.n_rad_filters, cur.rad_filters);
        }
      }
      S.scene_dirty = true;
      reply(true);
      break;
    }
    case 5: {  // set the current frame
      const int f = r.get<int32_t>();
      if (!r.ok || !S.trial) {
        reply(false);
        break;
      }
      S.frame = f;
      reply(true);
      break;
    }
    case 6: {  // get a pose
      const int volume = r.get<int32_t>();
      int frame = r.get<int32_t>();
      if (!r.ok || volume < 0 || volume >= (int)S.bones.size()) {
        reply(false);
        break;
      }
      if (frame == -1) frame = S.frame;
      double p[6];
      S.bones[volume].curves.pose_at(frame, p);
      reply(true);
      for (double v : p) put(out, v);
      break;
    }
    case 7: {  // set a pose
      const int volume = r.get<int32_t>();
      int frame = r.get<int32_t>();
      double p[6];
      for (double& v : p) v = r.get<double>();
      if (!r.ok || volume < 0 || volume >= (int)S.bones.size()) {
        reply(false);
        break;
      }
      if (frame == -1) frame = S.frame;
      S.bones[volume].curves.insert_pose(frame, p);
      reply(true);
      break;
    }
    case 8: {  // trackFrame of a pose at the current frame
      const int volume = r.get<int32_t>();
      double p[6];
      for (double& v : p) v = r.get<double>();
      double per[8];
      double cost;
      if (!r.ok || volume < 0 || volume >= (int)S.bones.size() || ensure_scene(S) ||
          eval_poses(S, volume, p, 1, &cost, per)) {
        reply(false);
        break;
      }
      reply(true);
      out.push_back((uint8_t)S.cams.size());
      for (size_t c = 0; c < S.cams.size(); c++) put(out, per[c]);
      break;
    }
    case 9: {  // background threshold
      const double t = r.get<double>();
      // AutoscoperMainWindow::setBackground keeps the old threshold when a
      // background image cannot be made (fewer than two frames); the
      // answer is 9 either way (Socket.cpp:300-311)
      if (r.ok) set_background(S, t);
      reply(r.ok);
      break;
    }
    case 10: {  // full-view images of a pose
      const int volume = r.get<int32_t>(), camera = r.get<int32_t>();
      double p[6];
      for (double& v : p) v = r.get<double>();
      unsigned w = 0, h = 0;
      std::vector<uint8_t> img;
      if (!r.ok || image_data(S, volume, camera, p, w, h, img)) {
        reply(false);
        break;
      }
      reply(true);
      put(out, (int32_t)w);
      put(out, (int32_t)h);
      out.insert(out.end(), img.begin(), img.end());
      break;
    }
    case 11: {  // optimise
      const int volume = r.get<int32_t>(), frame = r.get<int32_t>(), repeats = r.get<int32_t>(),
                max_iter = r.get<int32_t>();
      const double lo = r.get<double>(), hi = r.get<double>();
      const int stall = r.get<int32_t>(), dframe = r.get<int32_t>(), method = r.get<int32_t>(),
                cf = r.get<int32_t>(), heuristic = r.get<int32_t>();
      if (!r.ok) {
        reply(false);
        break;
      }
      S.guess = heuristic;
      reply(optimize(S, volume, frame, dframe, repeats, method, (unsigned)max_iter, lo, hi, cf, (unsigned)stall) ==
            ASC_OK);
      break;
    }
    case 12:  // save the full DRR images; the reference answers 12 whatever happened
      save_full_drr(S);
      reply(true);
      break;
    case 13: return false;         // close the connection
    case 14: {
      reply(true);
      put(out, (int32_t)S.bones.size());
      break;
    }
    case 15: {
      reply(true);
      put(out, (int32_t)(S.trial ? S.info.n_frames : 0));
      break;
    }
    case 16: {
      reply(true);
      put(out, (int32_t)2);  // AUTOSCOPER_SOCKET_VERSION
      break;
    }
    case 17: {  // volume visibility (display only)
      const int volume = r.get<int32_t>(), visible = r.get<int32_t>();
      if (!r.ok || volume < 0 || volume >= (int)S.bones.size()) {
        reply(false);
        break;
      }
      S.bones[volume].visible = visible != 0;
      reply(true);
      break;
    }
    default: out.push_back(0); break;
  }
  return true;
}

// One message, whatever happens inside: a failed allocation or a library
// exception answers the failure byte and keeps the connection.
bool handle(Session& S, const uint8_t* msg, size_t len, std::vector<uint8_t>& out)
{
  try {
    return handle_message(S, msg, len, out);
  } catch (const std::exception& e) {
    fail(ASC_ERR_IO, std::string("message failed: ") + e.what());
    out.assign(1, 0);
    return true;
  }
}

}  // namespace
}  // namespace asc

using namespace asc;

struct asc_remote_s {
  Session session;
  std::mutex mu;  // one message at a time
  int listen_fd = -1;
  int port = 0;
  std::atomic<bool> stop{false};
  std::thread io;
};

static void serve_loop(asc_remote r)
{
  std::vector<pollfd> fds;
  std::vector<uint8_t> buf(1 << 20), out;
  fds.push_back(pollfd{r->listen_fd, POLLIN, 0});
  while (!r->stop.load()) {
    if (::poll(fds.data(), fds.size(), 100) <= 0) continue;
    if (fds[0].revents & POLLIN) {
      const int c = ::accept(r->listen_fd, nullptr, nullptr);
      if (c >= 0) fds.push_back(pollfd{c, POLLIN, 0});
    }
    for (size_t i = 1; i < fds.size();) {
      bool keep = true;
      if (fds[i].revents & (POLLIN | POLLHUP | POLLERR)) {
        const ssize_t n = ::recv(fds[i].fd, buf.data(), buf.size(), 0);
        if (n <= 0) {
          keep = false;
        } else {
          std::lock_guard<std::mutex> lk(r->mu);
          keep = handle(r->session, buf.data(), (size_t)n, out);
          size_t sent = 0;
          while (keep && sent < out.size()) {
            const ssize_t k = ::send(fds[i].fd, out.data() + sent, out.size() - sent, MSG_NOSIGNAL);
            if (k <= 0) {
              keep = false;
              break;
            }
            sent += (size_t)k;
          }
        }
      }
      if (!keep) {
        ::close(fds[i].fd);
        fds.erase(fds.begin() + (long)i);
      } else {
        fds[i].revents = 0;
        i++;
      }
    }
  }
  for (size_t i = 1; i < fds.size(); i++) ::close(fds[i].fd);
}

extern "C" {

ASC_API int asc_remote_create(asc_context ctx, asc_remote* out)
{
  if (!out) return fail(ASC_ERR_INVALID, "null pointer");
  asc_remote r = new asc_remote_s();
  r->session.ctx = ctx;
  *out = r;
  return ASC_OK;
}

ASC_API int asc_remote_create_shared(asc_context ctx, asc_server server, asc_remote* out)
{
  if (!out || !server) return fail(ASC_ERR_INVALID, "null pointer");
  asc_remote r = new asc_remote_s();
  r->session.ctx = ctx;
  r->session.srv = server;
  *out = r;
  return ASC_OK;
}

ASC_API int asc_remote_listen(asc_remote r, int port)
{
  if (!r || port < 0 || port > 65535) return fail(ASC_ERR_INVALID, "bad argument");
  if (r->listen_fd >= 0) return fail(ASC_ERR_INVALID, "already listening");
  const int fd = ::socket(AF_INET, SOCK_STREAM, 0);
  if (fd < 0) return fail(ASC_ERR_IO, std::string("socket: ") + std::strerror(errno));
  const int one = 1;
  ::setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
  sockaddr_in a;
  std::memset(&a, 0, sizeof a);
  a.sin_family = AF_INET;
  a.sin_port = htons((uint16_t)port);
  a.sin_addr.s_addr = htonl(INADDR_LOOPBACK);  // QHostAddress::LocalHost
  if (::bind(fd, (sockaddr*)&a, sizeof a) != 0 || ::listen(fd, 8) != 0) {
    const std::string e = std::strerror(errno);
    ::close(fd);
    return fail(ASC_ERR_IO, "bind/listen on 127.0.0.1:" + std::to_string(port) + ": " + e);
  }
  socklen_t len = sizeof a;
  ::getsockname(fd, (sockaddr*)&a, &len);
  r->port = ntohs(a.sin_port);
  r->listen_fd = fd;
  r->stop = false;
  r->io = std::thread(serve_loop, r);
  return ASC_OK;
}

ASC_API int asc_remote_port(asc_remote r, int* port)
{
  if (!r || !port) return fail(ASC_ERR_INVALID, "null pointer");
  *port = r->port;
  return ASC_OK;
}

ASC_API int asc_remote_handle(asc_remote r, const void* message, size_t length, void* reply, size_t capacity,
                              size_t* reply_length)
{
  if (!r || (!message && length) || !reply_length) return fail(ASC_ERR_INVALID, "null pointer");
  std::vector<uint8_t> out;
  bool keep;
  {
    std::lock_guard<std::mutex> lk(r->mu);
    keep = handle(r->session, (const uint8_t*)message, length, out);
  }
  *reply_length = out.size();
  if (reply && capacity && !out.empty()) std::memcpy(reply, out.data(), std::min(capacity, out.size()));
  return keep ? ASC_OK : 1;
}

ASC_API int asc_remote_set_output_dir(asc_remote r, const char* dir)
{
  if (!r || !dir || !*dir) return fail(ASC_ERR_INVALID, "bad argument");
  std::lock_guard<std::mutex> lk(r->mu);
  r->session.full_drr_dir = dir;
  return ASC_OK;
}

ASC_API int asc_remote_destroy(asc_remote r)
{
  if (!r) return ASC_OK;
  r->stop = true;
  if (r->io.joinable()) r->io.join();
  if (r->listen_fd >= 0) ::close(r->listen_fd);
  delete r;
  return ASC_OK;
}

}  // extern "C"
