// Host-side geometry and status helpers (plain C++, no HIP headers).
#pragma once

#include <string>
#include <vector>

#include "../../include/autoscoper_hip.h"

namespace asc {

// ---------------------------------------------------------------------------
// Error plumbing: a thread-local message and an int status.
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// ---------------------------------------------------------------------------
// Host geometry (geometry.cpp). Frames: {r[9], t[3]} as CoordFrame.
// ---------------------------------------------------------------------------
struct Frame {
  double r[9];
  double t[3];
};
Frame frame_identity();
void frame_rotate(Frame& f, const double* axis, double angle_deg);
Frame frame_from_xyzypr(const double* p);
Frame frame_from_xyz_axis_angle(const double* p);
void frame_to_xyzypr(const Frame& f, double* p);
Frame frame_inverse(const Frame& f);
Frame frame_mul(const Frame& a, const Frame& b);
void frame_point(const Frame& f, const double* p, double* q);
void frame_row_order(const Frame& f, double* m16);
Frame frame_linear_extrap(const Frame& self, const Frame& x);
Frame frame_load(const double* f12);
void frame_store(const Frame& f, double* f12);

void inv_model_view(const double* m16, const float* inv_scale, const float* inv_trans, float* out16);
bool calc_viewport(const Frame& modelview, const asc_camera& cam, const float* inv_trans, const float* inv_scale,
                   double* vp);
// Full per-camera geometry of one pose: returns 1 in view, 0 not in view.
int view_geometry(const asc_camera& cam, const double* xyzypr, const float* inv_scale, const float* inv_trans,
                  int render_w, int render_h, float* imv12, double* vp4, unsigned* size2);
// The camera's inverse frame and the radiograph's film-plane bounds
// {left, bottom, right, top}, which view_geometry derives per camera.
struct CameraView {
  Frame inv;
  double r[4];
};
CameraView camera_view(const asc_camera& cam);
// The 8 corners of a volume's box in volume coordinates, which
// view_geometry derives per volume from inv_trans and inv_scale.
struct BoxCorners {
  double p[24];
};
BoxCorners box_corners(const float* inv_trans, const float* inv_scale);
// view_geometry from camera_view(cam), frame_from_xyzypr(xyzypr) and
// box_corners(inv_trans, inv_scale), for callers that reuse them across
// cameras, poses and volumes (same results).
int view_geometry(const asc_camera& cam, const CameraView& cv, const Frame& pose, const BoxCorners& box,
                  const float* inv_scale, const float* inv_trans, int render_w, int render_h, float* imv12,
                  double* vp4, unsigned* size2);

struct PoseStore {
  float x, y, z;
  float qw, qx, qy, qz;
};
void pose_store_set(PoseStore& s, const double* xyzypr);
void pose_store_get(const PoseStore& s, double* xyzypr);
void compose_pose(const PoseStore& s, const Frame& pivot, const double* manip, double* out_xyzypr);

// Camera::calculateViewport and calculateImagePlane (Camera.cpp:424-484) of
// a camera whose rotation, translation and size are set.
int camera_finish(asc_camera& c, double cx, double cy, double fx, double fy, double z);

// Gaussian/sharpen weights (GaussianFilter.cpp:75-127): returns N.
int gauss_weights(float radius, std::vector<float>& w);

}  // namespace asc
