// Host-side pose and projection geometry in double precision.
//
// These computations decide integer results (the projected bounding box and
// the render size), so every expression keeps the reference's operation
// order and this file is compiled without FMA contraction. The transcendental
// calls are the C library's, as in the reference.
//
// References (libautoscoper/src/): CoordFrame.cpp, Camera.cpp:229-484,
// RayCaster.cpp:92-142, Tracker.cpp:583-629, 775-796, 923-1032,
// math/Quaternion.hpp:80-94, 220-243, GaussianFilter.cpp:75-127.
#include <cmath>
#include <cstring>
#include <utility>

#include "geometry.hpp"

namespace asc {

Frame frame_identity()
{
  Frame f{};
  f.r[0] = f.r[4] = f.r[8] = 1.0;
  return f;
}

Frame frame_load(const double* f12)
{
  Frame f;
  std::memcpy(f.r, f12, sizeof f.r);
  std::memcpy(f.t, f12 + 9, sizeof f.t);
  return f;
}

void frame_store(const Frame& f, double* f12)
{
  std::memcpy(f12, f.r, sizeof f.r);
  std::memcpy(f12 + 9, f.t, sizeof f.t);
}

// Right-multiply the rotation by an axis-angle rotation (degrees).
// CoordFrame::rotate, CoordFrame.cpp:262-333.
void frame_rotate(Frame& f, const double* axis_in, double angle_deg)
{
  double ax[3] = {axis_in[0], axis_in[1], axis_in[2]};
  const double c = cos(M_PI * angle_deg / 180.0);
  const double s = sin(M_PI * angle_deg / 180.0);
  const double mag = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  if (mag == 0) return;
  for (double& a : ax) a /= mag;
  const double oc = 1.0 - c;
  // rotation matrix R (row-major as the reference lays it out)
  const double R[9] = {(oc * (ax[0] * ax[0])) + c,         (oc * (ax[0] * ax[1])) + s * ax[2],
                       (oc * (ax[0] * ax[2])) - s * ax[1], (oc * (ax[0] * ax[1])) - s * ax[2],
                       (oc * (ax[1] * ax[1])) + c,         (oc * (ax[1] * ax[2])) + s * ax[0],
                       (oc * (ax[0] * ax[2])) + s * ax[1], (oc * (ax[1] * ax[2])) - s * ax[0],
                       (oc * (ax[2] * ax[2])) + c};
  double out[9];
  for (int row = 0; row < 3; row++)
    for (int col = 0; col < 3; col++)
      out[3 * row + col] =
        f.r[3 * row + 0] * R[col] + f.r[3 * row + 1] * R[3 + col] + f.r[3 * row + 2] * R[6 + col];
  std::memcpy(f.r, out, sizeof out);
}

// CoordFrame::from_xyzypr, CoordFrame.cpp:100-108: yaw about z, then pitch
// about the new y, then roll about the new x; then translate.
Frame frame_from_xyzypr(const double* p)
{
  Frame f = frame_identity();
  frame_rotate(f, f.r + 6, p[3]);
  frame_rotate(f, f.r + 3, p[4]);
  frame_rotate(f, f.r + 0, p[5]);
  for (int i = 0; i < 3; i++) f.t[i] += p[i];
  return f;
}

// CoordFrame::from_xyzAxis_angle, CoordFrame.cpp:163-175: the rotation
// vector's length is the angle in degrees.
Frame frame_from_xyz_axis_angle(const double* p)
{
  Frame f = frame_identity();
  const double angle = sqrt(p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
  if (angle != 0) frame_rotate(f, p + 3, angle);
  for (int i = 0; i < 3; i++) f.t[i] += p[i];
  return f;
}

// CoordFrame::to_xyzypr, CoordFrame.cpp:177-189.
void frame_to_xyzypr(const Frame& f, double* p)
{
  for (int i = 0; i < 3; i++) p[i] = f.t[i];
  p[3] = 180.0 * atan2(f.r[1], f.r[0]) / M_PI;
  p[4] = 180.0 * atan2(-f.r[2], sqrt(f.r[5] * f.r[5] + f.r[8] * f.r[8])) / M_PI;
  p[5] = 180 * atan2(f.r[5], f.r[8]) / M_PI;
}

// CoordFrame::inverse, CoordFrame.cpp:335-353.
Frame frame_inverse(const Frame& f)
{
  Frame o;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o.r[3 * i + j] = f.r[3 * j + i];
  for (int i = 0; i < 3; i++)
    o.t[i] = -(o.r[i] * f.t[0] + o.r[3 + i] * f.t[1] + o.r[6 + i] * f.t[2]);
  return o;
}

// CoordFrame::operator*, CoordFrame.cpp:478-527.
Frame frame_mul(const Frame& a, const Frame& b)
{
  Frame o;
  for (int col = 0; col < 3; col++)
    for (int i = 0; i < 3; i++)
      o.r[3 * col + i] = a.r[i] * b.r[3 * col + 0] + a.r[3 + i] * b.r[3 * col + 1] + a.r[6 + i] * b.r[3 * col + 2];
  for (int i = 0; i < 3; i++) o.t[i] = a.r[i] * b.t[0] + a.r[3 + i] * b.t[1] + a.r[6 + i] * b.t[2] + a.t[i];
  return o;
}

// CoordFrame::point_to_world_space, CoordFrame.cpp:355-370.
void frame_point(const Frame& f, const double* p, double* q)
{
  for (int i = 0; i < 3; i++) q[i] = f.r[i] * p[0] + f.r[3 + i] * p[1] + f.r[6 + i] * p[2] + f.t[i];
}

// CoordFrame::to_matrix_row_order, CoordFrame.cpp:223-241.
void frame_row_order(const Frame& f, double* m)
{
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) m[4 * row + col] = f.r[3 * col + row];
    m[4 * row + 3] = f.t[row];
  }
  m[12] = m[13] = m[14] = 0.0;
  m[15] = 1.0;
}

// CoordFrame::linear_extrap, CoordFrame.cpp:419-476 (used by guess 2).
Frame frame_linear_extrap(const Frame& s, const Frame& x)
{
  const double t = 2.0;
  double tr[3];
  for (int i = 0; i < 3; i++) tr[i] = t * (x.t[i] - s.t[i]);
  double A[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[3 * i + j] = x.r[j] * s.r[i] + x.r[3 + j] * s.r[3 + i] + x.r[6 + j] * s.r[6 + i];
  double axis[3] = {1.0, 0.0, 0.0};
  const double ang = acos((A[0] + A[4] + A[8] - 1.0) / 2.0);
  if (ang > 0.0 && ang < M_PI) {
    axis[0] = A[5] - A[7];
    axis[1] = A[6] - A[2];
    axis[2] = A[1] - A[3];
  }
  Frame o = s;
  frame_rotate(o, axis, t * 180.0 * ang / M_PI);
  for (int i = 0; i < 3; i++) o.t[i] += tr[i];
  return o;
}

// RayCaster::setInvModelView, RayCaster.cpp:92-142: fold the volume's
// normalising scale/offset into rows 0-2 (double arithmetic, stored float).
void inv_model_view(const double* m, const float* inv_scale, const float* inv_trans, float* out)
{
  for (int row = 0; row < 3; row++)
    for (int col = 0; col < 4; col++)
      out[4 * row + col] = (float)(m[4 * row + col] * inv_scale[row] + m[12 + col] * inv_trans[row]);
  for (int col = 0; col < 4; col++) out[12 + col] = (float)m[12 + col];
}

// The camera's side of Tracker::calculate_viewport (Tracker.cpp:923-1032):
// its inverse frame and the radiograph's film-plane (z = -2) bounds, the
// projections of image-plane corners 1 and 3. They depend on the camera
// alone, so a batch computes them once per camera.
CameraView camera_view(const asc_camera& cam)
{
  Frame cf;
  std::memcpy(cf.r, cam.rotation, sizeof cf.r);
  std::memcpy(cf.t, cam.translation, sizeof cf.t);
  CameraView cv;
  cv.inv = frame_inverse(cf);
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  frame_point(cv.inv, cam.image_plane + 3, a);
  frame_point(cv.inv, cam.image_plane + 9, b);
  cv.r[0] = -2 * a[0] / a[2];
  cv.r[1] = -2 * a[1] / a[2];
  cv.r[2] = -2 * b[0] / b[2];
  cv.r[3] = -2 * b[1] / b[2];
  return cv;
}

// The volume's side of Tracker::calculate_viewport: the 8 corners of its
// box in volume coordinates (the unit box undone by inv_trans, inv_scale).
BoxCorners box_corners(const float* inv_trans, const float* inv_scale)
{
  static const double unit_corners[24] = {0, 0, -1, 0, 0, 0, 0, 1, -1, 0, 1, 0,
                                          1, 0, -1, 1, 0, 0, 1, 1, -1, 1, 1, 0};
  BoxCorners b;
  for (int j = 0; j < 8; j++)
    for (int a = 0; a < 3; a++) b.p[3 * j + a] = (unit_corners[3 * j + a] - inv_trans[a]) / inv_scale[a];
  return b;
}

// Tracker::calculate_viewport, Tracker.cpp:923-1032: film-plane (z = -2)
// bounding box of the projected unit box, clipped to the radiograph.
// Note the box starts at {1, 1, -1, -1}, as in the reference.
static bool clip_viewport(const Frame& mv, const CameraView& cv, const BoxCorners& box, double* vp)
{
  double lo_x = 1.0, lo_y = 1.0, hi_x = -1.0, hi_y = -1.0;
  for (int j = 0; j < 8; j++) {
    double q[3];
    frame_point(mv, box.p + 3 * j, q);
    const double fx = -2 * q[0] / q[2];
    const double fy = -2 * q[1] / q[2];
    if (lo_x > fx) lo_x = fx;
    if (lo_y > fy) lo_y = fy;
    if (hi_x < fx) hi_x = fx;
    if (hi_y < fy) hi_y = fy;
  }
  const double r0 = cv.r[0], r1 = cv.r[1], r2 = cv.r[2], r3 = cv.r[3];
  if (lo_x > r2 && hi_x > r2) return false;
  if (lo_y > r3 && hi_y > r3) return false;
  if (lo_x < r0 && hi_x < r0) return false;
  if (lo_y < r1 && hi_y < r1) return false;
  if (lo_x == 1.0 && lo_y == 1.0 && hi_x == -1.0 && hi_y == -1.0) return false;
  if (lo_x < r0) lo_x = r0;
  if (lo_y < r1) lo_y = r1;
  if (hi_x > r2) hi_x = r2;
  if (hi_y > r3) hi_y = r3;
  vp[0] = lo_x;
  vp[1] = lo_y;
  vp[2] = hi_x - lo_x;
  vp[3] = hi_y - lo_y;
  return true;
}

bool calc_viewport(const Frame& mv, const asc_camera& cam, const float* inv_trans, const float* inv_scale,
                   double* vp)
{
  return clip_viewport(mv, camera_view(cam), box_corners(inv_trans, inv_scale), vp);
}

// Tracker.cpp:604-629 for one camera, from the camera's view and the
// pose's frame (frame_from_xyzypr).
int view_geometry(const asc_camera& cam, const CameraView& cv, const Frame& pose, const BoxCorners& box,
                  const float* inv_scale, const float* inv_trans, int render_w, int render_h, float* imv12,
                  double* vp4, unsigned* size2)
{
  const Frame mv = frame_mul(cv.inv, pose);
  double m16[16];
  frame_row_order(frame_inverse(mv), m16);
  float imv16[16];
  inv_model_view(m16, inv_scale, inv_trans, imv16);
  std::memcpy(imv12, imv16, 12 * sizeof(float));
  if (!clip_viewport(mv, cv, box, vp4)) return 0;
  size2[0] = (unsigned)(vp4[2] * render_w / cam.viewport[2]);
  size2[1] = (unsigned)(vp4[3] * render_h / cam.viewport[3]);
  return 1;
}

int view_geometry(const asc_camera& cam, const double* xyzypr, const float* inv_scale, const float* inv_trans,
                  int render_w, int render_h, float* imv12, double* vp4, unsigned* size2)
{
  return view_geometry(cam, camera_view(cam), frame_from_xyzypr(xyzypr), box_corners(inv_trans, inv_scale),
                       inv_scale, inv_trans, render_w, render_h, imv12, vp4, size2);
}

// Quat<float>(yaw, pitch, roll), Quaternion.hpp:80-94.
static void quatf_from_ypr(float yaw, float pitch, float roll, float* q)
{
  const float cr = cosf(roll * float(M_PI) / float(360)), sr = sinf(roll * float(M_PI) / float(360));
  const float cp = cosf(pitch * float(M_PI) / float(360)), sp = sinf(pitch * float(M_PI) / float(360));
  const float cy = cosf(yaw * float(M_PI) / float(360)), sy = sinf(yaw * float(M_PI) / float(360));
  q[0] = cr * cp * cy + sr * sp * sy;
  q[1] = sr * cp * cy - cr * sp * sy;
  q[2] = cr * sp * cy + sr * cp * sy;
  q[3] = cr * cp * sy - sr * sp * cy;
}

// Quat<float>::toEuler, Quaternion.hpp:220-243 -> {roll, pitch, yaw} degrees.
static void quatf_to_euler(const float* q, float* rpy)
{
  const float w = q[0], x = q[1], y = q[2], z = q[3];
  const float roll = atan2f(float(2) * (w * x + y * z), float(1) - float(2) * (x * x + y * y));
  const float sinp = sqrtf(float(1) + float(2) * (w * y - x * z));
  const float cosp = sqrtf(float(1) - float(2) * (w * y - x * z));
  const float pitch = float(2) * atan2f(sinp, cosp) - float(M_PI) / float(2);
  const float yaw = atan2f(float(2) * (w * z + x * y), float(1) - float(2) * (y * y + z * z));
  rpy[0] = roll * (float(180) / float(M_PI));
  rpy[1] = pitch * (float(180) / float(M_PI));
  rpy[2] = yaw * (float(180) / float(M_PI));
}

// KeyCurve<float>::insert of x, y, z and KeyCurve<Quatf>::insert of
// Quatf(yaw, pitch, roll) (Tracker.cpp:513-517).
void pose_store_set(PoseStore& s, const double* p)
{
  s.x = (float)p[0];
  s.y = (float)p[1];
  s.z = (float)p[2];
  float q[4];
  quatf_from_ypr((float)p[3], (float)p[4], (float)p[5], q);
  s.qw = q[0];
  s.qx = q[1];
  s.qy = q[2];
  s.qz = q[3];
}

// Tracker.cpp:777-784: x, y, z from float curves; yaw, pitch, roll from
// the quaternion curve's Euler angles.
void pose_store_get(const PoseStore& s, double* p)
{
  const float q[4] = {s.qw, s.qx, s.qy, s.qz};
  float rpy[3];
  quatf_to_euler(q, rpy);
  p[0] = s.x;
  p[1] = s.y;
  p[2] = s.z;
  p[3] = rpy[2];
  p[4] = rpy[1];
  p[5] = rpy[0];
}

// Tracker.cpp:785-796: xcframe = base * V^-1 * manip * V.
void compose_pose(const PoseStore& s, const Frame& pivot, const double* manip, double* out)
{
  double base[6];
  pose_store_get(s, base);
  Frame xc = frame_from_xyzypr(base);
  xc = frame_mul(xc, frame_inverse(pivot));
  xc = frame_mul(xc, frame_from_xyz_axis_angle(manip));
  xc = frame_mul(xc, pivot);
  frame_to_xyzypr(xc, out);
}

// GaussianFilter::set_radius / SharpenFilter::makeFilter weights.
int gauss_weights(float radius, std::vector<float>& w)
{
  if (radius < 0) radius = 0;
  const int fr = (int)(3 * radius);
  const int n = 2 * fr + 1;
  w.assign((size_t)n * n, 0.0f);
  if (n == 1) return 1;
  float sum = 0.0f;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const int d2 = (i - fr) * (i - fr) + (j - fr) * (j - fr);
      w[(size_t)i * n + j] = (float)exp(d2 / (-2.0 * radius));
      sum = sum + w[(size_t)i * n + j];
    }
  for (float& v : w) v = v / sum;
  return n;
}

// Camera::calculateViewport and calculateImagePlane (Camera.cpp:424-484)
// for principal point (cx, cy), focal lengths (fx, fy) and film distance z.
int camera_finish(asc_camera& c, double cx, double cy, double fx, double fy, double z)
{
  if (fx == 0 || fy == 0) return fail(ASC_ERR_INVALID, "Invalid camera parameters (fx or fy is zero)");
  c.viewport[0] = -(2.0f * cx - c.size[0]) / fx;
  c.viewport[1] = -(2.0f * cy - c.size[1]) / fy;
  c.viewport[2] = 2.0f * c.size[0] / fx;
  c.viewport[3] = 2.0f * c.size[1] / fy;
  const Frame cf = frame_load(c.rotation);  // rotation[9] then translation[3]
  const double* tn = c.translation;
  const double distance = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
  const double scale = -1.5 * distance / z;
  const double ipt[3] = {scale * (c.size[0] / 2.0 - cx), scale * (c.size[1] / 2.0 - cy), scale * z};
  double ctr[3];
  frame_point(cf, ipt, ctr);
  const double hw = scale * c.size[0] / 2.0, hh = scale * c.size[1] / 2.0;
  const double* right = cf.r;
  const double* up = cf.r + 3;
  for (int i = 0; i < 3; i++) {
    c.image_plane[0 + i] = ctr[i] - hw * right[i] + hh * up[i];
    c.image_plane[3 + i] = ctr[i] - hw * right[i] - hh * up[i];
    c.image_plane[6 + i] = ctr[i] + hw * right[i] - hh * up[i];
    c.image_plane[9 + i] = ctr[i] + hw * right[i] + hh * up[i];
  }
  return ASC_OK;
}

}  // namespace asc
