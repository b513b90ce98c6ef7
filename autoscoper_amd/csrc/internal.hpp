// Internal declarations shared by the HIP sources of libautoscoper_hip.
//
// Floating-point policy: every translation unit is compiled with
// -ffp-contract=off and correctly rounded fp32 division/sqrt, so code written
// as `a * b + c` evaluates exactly like the reference kernels' IEEE float
// operations. Hot loops that may contract use explicit fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/autoscoper_hip.h"
#include "geometry.hpp"

namespace asc {

int hip_fail(hipError_t e, const char* what);

#define ASC_HIP(call)                                \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return asc::hip_fail(e_, #call); \
  } while (0)

#define ASC_CHECK(cond, code, msg)        \
  do {                                    \
    if (!(cond)) return asc::fail(code, msg); \
  } while (0)

// ---------------------------------------------------------------------------
// Device-side views of volumes and images.
// ---------------------------------------------------------------------------
// Volume: zero-padded by one voxel on every side, x fastest. pW = W + 2,
// pS = (W + 2) * (H + 2). Texel (i, j, k) of the cropped volume lives at
// ((k + 1) * (H + 2) + (j + 1)) * pW + (i + 1). The zero border realises the
// sampler's CLK_ADDRESS_CLAMP border colour for every in-box sample.
// Layouts (ASC_LAYOUT_*): LINEAR stores one texel per element; QUAD stores
// the 2x2 (x, y) corner block of each element (4 texels), OCTET the 2x2x2
// block (8 texels), so a trilinear sample reads 4, 2 or 1 aligned vectors.
// BRICK stores OCTET cells grouped in bricks of 8 cells (128 bytes for
// 16-bit texels), oriented across the rays: a volume in BRICK keeps one copy
// per ray axis a (x, y, z), whose bricks are 1 cell along a, 2 along the
// next axis and 4 along the third (x: 1x2x4, y: 2x1x4, z: 2x4x1), laid out
// with a fastest; the march reads the copy across each item's central ray
// (ray_axis). A cell's byte offset is the sum over the axes of
// bin[a] * (c_a mod 2^blog[a]) + bstr[a] * (floor(c_a / 2^blog[a]) + 1).
// pair_diff, the encoding of u16 volumes in QUAD, OCTET or BRICK (the
// integers each 2x2 (x, y) block {t00, t10, t01, t11} of a slice holds):
//   0: the texels as they are;
//   1 (every texel <= 32767): the first texel and three signed 16-bit
//      differences: {t00, t10 - t00, t01 - t00, t11 - t01};
//   2 (every texel <= 16383): the row differences too:
//      {t00, t10 - t00, t01 - t00, (t11 - t01) - (t10 - t00)}.
// Each is exactly the integers the march's lerps need (device.hpp).
// The ray march addresses BRICK cells through per-block LDS tables of the
// x, y and z offset terms (W + 1, H + 1 and D + 1 ints); a volume whose
// W + H + D + 3 exceeds this many ints is stored in another layout.
constexpr int kBrickTab = 4096;

struct VolDev {
  const void* data;
  int W, H, D;
  int pW, pS;
  int storage;   // ASC_STORAGE_*
  float vscale;  // raw texel value -> normalised value (1/65535 for U16, else 1)
  int layout;    // ASC_LAYOUT_*
  int esz;       // bytes per element of the layout
  int nbytes;    // bytes of the layout's array (the buffer's extent)
  int pair_diff; // u16 paired layouts: the encoding 0, 1 or 2 above
  // BRICK: per axis the log2 of the brick's extent in cells, the in-brick
  // stride and the brick stride in bytes
  int blog[3], bin[3], bstr[3];
};

// The axis (0 x, 1 y, 2 z) along which the central ray of a view with
// inverse model-view `imv` advances most voxels per unit step: its
// direction is imv's third column, and the unit box spans W, H, D voxels.
__host__ __device__ inline int ray_axis(const VolDev& v, const float* imv)
{
  const float ax = fabsf(imv[2]) * (float)v.W, ay = fabsf(imv[6]) * (float)v.H, az = fabsf(imv[10]) * (float)v.D;
  return ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
}

struct ImgDev {
  const void* data;  // float texels, already normalised (UNORM -> [0, 1])
  int W, H;
  int bps;           // 32
};

}  // namespace asc

// ---------------------------------------------------------------------------
// Host objects behind the opaque handles.
// ---------------------------------------------------------------------------
struct asc_context_s {
  int device;
  hipStream_t own_stream;
  hipStream_t stream;
  // NCC scratch (block partials + result)
  double* ncc_partials = nullptr;
  size_t ncc_partials_cap = 0;  // number of blocks
  double* ncc_result = nullptr;  // 10 doubles
  double* ncc_host = nullptr;    // pinned
  float* conv_wts = nullptr;     // large Gaussian/sharpen weight sets
  // the reference-form NCC (ASC_NCC_TREE_*): per-pixel products and pass
  // buffers, grown on demand
  float* tree_buf = nullptr;
  size_t tree_cap = 0, tree_part = 0;
  int ncc_form = ASC_NCC_MOMENTS;
  // When set, every per-operator entry point waits for its work before
  // returning, as the reference's OpenCL Kernel::launch does (clFinish after
  // each launch, OpenCL.cpp:978). Off by default: the calls stay asynchronous.
  bool blocking = false;
};

struct asc_volume_s {
  asc_context ctx;
  asc_volume_desc desc;
  asc::VolDev dev;
  void* dptr;       // LINEAR copy (always kept)
  size_t bytes;
  // the packed copies of the selected layout: one (QUAD, OCTET) or one per
  // ray axis (BRICK), and their device views; in LINEAR all three views are
  // `dev`
  void* packed[3] = {nullptr, nullptr, nullptr};
  asc::VolDev odev[3];
  size_t packed_bytes = 0;
  int generation = 0;      // bumped whenever `dev` changes
  int u16_max = 0;         // largest stored texel of a u16 volume
  // the view the march reads for rays of a view with inverse model-view imv
  const asc::VolDev& ray_view(const float* imv) const
  {
    return odev[dev.layout == ASC_LAYOUT_BRICK ? asc::ray_axis(dev, imv) : 0];
  }
};

struct asc_image_s {
  asc_context ctx;
  asc::ImgDev dev;  // always 32-bit float texels on the device
  void* dptr;
  size_t bytes;     // device bytes (4 per texel)
  int src_bps;      // bit depth of the caller's data (8, 16 or 32)
};

namespace asc {

// ---------------------------------------------------------------------------
// Kernel launchers (kernels.hip).
// ---------------------------------------------------------------------------
// v: the view of the volume for imv's rays (asc_volume_s::ray_view)
int launch_drr(asc_context ctx, const VolDev& v, const float* imv, const float* vp, float step, float intensity,
               float cutoff, float* out, unsigned w, unsigned h);
int launch_rad(asc_context ctx, const ImgDev& img, const float* plane, const float* vp, float* out, unsigned w,
               unsigned h, bool background, float threshold);
int launch_filter(asc_context ctx, const asc_filter& f, const float* in, float* out, int w, int h);
int launch_binary(asc_context ctx, int op, const float* a, const float* b, float* out, size_t n);
int launch_mask_multiply(asc_context ctx, float* drr, float* rad, const float* bg, float* mask, size_t n);
int launch_composite(asc_context ctx, const float* s1, const float* s2, const float* s3, float* out, size_t n);
int launch_fill_f32(asc_context ctx, float* dst, float v, size_t n);
int launch_copy(asc_context ctx, void* dst, const void* src, size_t bytes);
int launch_pack_volume(asc_context ctx, const VolDev& linear, void* dst, int layout, int pair_diff);
// BRICK copy `dst` described by `brick` (its blog, bin, bstr and nbytes)
int launch_pack_brick(asc_context ctx, const VolDev& linear, const VolDev& brick, void* dst, int pair_diff);
// gpu::ncc as the reference computes it (float tree sums, Ncc.cpp:180-220):
// sums6 = {mask, f, g, products, f dev^2, g dev^2} tree sums (products NaN
// when the denominator is below 1e-5 and the reference skips that sum).
int ncc_tree(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, bool opencl,
             float* sums6, float* ncc);
int ncc_moments(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, double* out9,
                double* ncc);
// asc_ncc_noalloc: the static scratch's address for `device` (current), once;
// then the two moment kernels into the caller's out10, nothing allocated.
int noalloc_scratch_init(int device);
int ncc_noalloc(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, double* out10);

// NCC from the nine moments, combined in double (Ncc.cpp:180-220 semantics).
__host__ __device__ inline double ncc_from_moments(const double* m)
{
  // m = {sum mask, sum f, sum g, count, S_f, S_g, S_ff, S_gg, S_fg}
  const double nb = m[0];
  const double mf = m[1] / nb, mg = m[2] / nb;
  const double n = m[3];
  const double num = m[8] - mg * m[4] - mf * m[5] + n * mf * mg;
  double d1 = m[6] - 2.0 * mf * m[4] + n * mf * mf;
  double d2 = m[7] - 2.0 * mg * m[5] + n * mg * mg;
  // both are sums of squares; rounding in the one-pass form must not make
  // them negative
  d1 = d1 < 0.0 ? 0.0 : d1;
  d2 = d2 < 0.0 ? 0.0 : d2;
  const double den = sqrt(d1 * d2);
  if (den < 1e-5) return 1e5;
  return num / den;
}

enum { BIN_MULTIPLY = 0, BIN_MERGE = 1, BIN_DRR_MASK = 2 };

}  // namespace asc
