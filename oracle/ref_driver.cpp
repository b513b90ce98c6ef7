// TEST INFRASTRUCTURE ONLY. Glue that drives the reference's own host C++
// (libautoscoper/src/CoordFrame.cpp, PSO.cpp, DownhillSimplex.cpp and the
// header-only math/Quaternion.hpp), compiled from the sources where they lie
// by oracle/Makefile's `ref` target into oracle/_ref/libref_host.so, so the
// tests can check the oracle's restatement of that code against the code
// itself. Nothing here is linked into the product library.
//
// The reference optimisers call the global callbacks FUNC (simplex, 1-indexed
// vertex rows, Tracker.cpp:88-91) and PSO_FUNC (PSO manipulator,
// Tracker.cpp:93-96); here both forward to a cost function chosen per call.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

#include "CoordFrame.hpp"
#include "DownhillSimplex.hpp"
#include "PSO.hpp"
#include "Quaternion.hpp"

typedef double (*ref_cost_fn)(const double* x, void* user);

static ref_cost_fn g_fn = nullptr;
static void* g_user = nullptr;
static long g_calls = 0;

double FUNC(double* P)
{
  g_calls++;
  return g_fn(P + 1, g_user);
}

double PSO_FUNC(double* P)
{
  g_calls++;
  return g_fn(P, g_user);
}

using xromm::CoordFrame;

static void frame_store(const CoordFrame& f, double* out12)
{
  // rotation (column-major, columns = images of the axes) then translation
  std::memcpy(out12, f.rotation(), 9 * sizeof(double));
  std::memcpy(out12 + 9, f.translation(), 3 * sizeof(double));
}

static CoordFrame frame_load(const double* f12) { return CoordFrame(f12, f12 + 9); }

extern "C" {

void ref_frame_from_xyzypr(const double* p, double* out) { frame_store(CoordFrame::from_xyzypr(p), out); }

void ref_frame_from_xyz_axis_angle(const double* p, double* out)
{
  frame_store(CoordFrame::from_xyzAxis_angle(p), out);
}

void ref_frame_to_xyzypr(const double* f, double* p) { frame_load(f).to_xyzypr(p); }

void ref_frame_mul(const double* a, const double* b, double* out) { frame_store(frame_load(a) * frame_load(b), out); }

void ref_frame_inverse(const double* a, double* out) { frame_store(frame_load(a).inverse(), out); }

void ref_frame_linear_extrap(const double* a, const double* b, double* out)
{
  frame_store(frame_load(a).linear_extrap(frame_load(b)), out);
}

void ref_frame_matrix_row_order(const double* a, double* m16) { frame_load(a).to_matrix_row_order(m16); }

// A pose written into single-key float curves and read back the way
// Tracker::minimizationFunc reads it (Tracker.cpp:413,426,778-784): the
// translation as float, the rotation as Quatf(yaw, pitch, roll) and
// toEuler().
void ref_pose_store_roundtrip(const double* xyzypr, double* out)
{
  const float x = (float)xyzypr[0], y = (float)xyzypr[1], z = (float)xyzypr[2];
  const Quatf q(xyzypr[3], xyzypr[4], xyzypr[5]);
  const Vec3f e = q.toEuler();
  const double r[6] = { x, y, z, e.z, e.y, e.x };
  std::memcpy(out, r, sizeof r);
}

// Tracker::minimizationFunc's pose composition (Tracker.cpp:785-791) on a
// base pose already read back from the curves.
void ref_compose_pose(const double* base_xyzypr, const double* pivot12, const double* manip, double* out)
{
  CoordFrame xc = CoordFrame::from_xyzypr(base_xyzypr);
  const CoordFrame pivot = frame_load(pivot12);
  const CoordFrame m = CoordFrame::from_xyzAxis_angle(manip);
  xc = xc * pivot.inverse() * m * pivot;
  xc.to_xyzypr(out);
}

// PSO.cpp:198-295 with srand(seed) (what initializeRandom does when
// Autoscoper_RANDOM_SEED is set). Returns the number of cost calls.
long ref_pso(ref_cost_fn fn, void* user, unsigned seed, float lo, float hi, unsigned max_epochs, unsigned max_stall,
             float* best_pos, float* best_ncc)
{
  g_fn = fn;
  g_user = user;
  g_calls = 0;
  srand(seed);
  std::streambuf* old = std::cout.rdbuf();
  std::ostringstream sink;
  std::cout.rdbuf(sink.rdbuf());  // the reference logs every epoch
  const Particle best = pso(lo, hi, max_epochs, max_stall);
  std::cout.rdbuf(old);
  for (int d = 0; d < NUM_OF_DIMENSIONS; d++) best_pos[d] = best.Position[d];
  *best_ncc = best.NCC;
  return g_calls;
}

// DownhillSimplex.cpp:39-172. P is (ndim+1) x ndim row-major and Y has
// ndim+1 entries, both 0-indexed; they are moved into the reference's
// 1-indexed MAT and back. Returns ITER.
int ref_amoeba(ref_cost_fn fn, void* user, double* P, double* Y, int ndim, double ftol)
{
  g_fn = fn;
  g_user = user;
  g_calls = 0;
  MAT M;
  double YY[MP];
  std::memset(M, 0, sizeof M);
  std::memset(YY, 0, sizeof YY);
  for (int i = 0; i <= ndim; i++) {
    for (int j = 0; j < ndim; j++) M[i + 1][j + 1] = P[i * ndim + j];
    YY[i + 1] = Y[i];
  }
  int iter = 0;
  AMOEBA(M, YY, ndim, ftol, &iter);
  for (int i = 0; i <= ndim; i++) {
    for (int j = 0; j < ndim; j++) P[i * ndim + j] = M[i + 1][j + 1];
    Y[i] = YY[i + 1];
  }
  return iter;
}

long ref_calls() { return g_calls; }

}  // extern "C"

// The reference's pose curves (KeyCurve.hpp, header only): keys at `times`
// inserted in the given order, then the curve evaluated at `at`.
#include "KeyCurve.hpp"

extern "C" {

void ref_float_curve(int n, const int* times, const float* values, int m, const float* at, float* out)
{
  KeyCurve<float> c;
  for (int i = 0; i < n; i++) c.insert(times[i], values[i]);
  for (int i = 0; i < m; i++) out[i] = c(at[i]);
}

void ref_quat_curve(int n, const int* times, const float* wxyz, int m, const float* at, float* out)
{
  KeyCurve<Quatf> c;
  for (int i = 0; i < n; i++) c.insert(times[i], Quatf(wxyz[4 * i], wxyz[4 * i + 1], wxyz[4 * i + 2], wxyz[4 * i + 3]));
  for (int i = 0; i < m; i++) {
    const Quatf q = c(at[i]);
    out[4 * i] = q.w;
    out[4 * i + 1] = q.x;
    out[4 * i + 2] = q.y;
    out[4 * i + 3] = q.z;
  }
}

}  // extern "C"
