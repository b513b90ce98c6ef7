// CPU oracle for Autoscoper's 2D-3D registration cost path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, the arithmetic of
// the reference OpenCL backend and its host code so that the HIP product can
// be checked against it. Only the test-suite, the smoke check and the CPU
// baseline leg of the benchmark load it; the product library never links,
// loads or calls it.
//
// Two arithmetic modes are provided through templates on the real type R:
//   R = float  : "float-faithful". The same float operations, in the same
//                order, as the reference kernels, with no implicit
//                contraction into FMA (built with -ffp-contract=off) and
//                correctly rounded division and sqrt. Two expressions of
//                the ray march are written as the FMA an OpenCL compiler
//                forms by default (FP_CONTRACT on): the sample position
//                o + t*d and the sampler's s*W - 0.5. The OpenCL 1.2 image sampler rules
//                (normalised coordinates, CLK_ADDRESS_CLAMP border 0,
//                CLK_FILTER_LINEAR, UNORM conversion) are written out.
//   R = double : the same algorithm evaluated in double ("FP64 oracle").
//
// Every function cites the reference file:line it restates (paths are
// relative to the reference checkout; kernels live in
// libautoscoper/src/gpu/opencl/kernel/).
//
// Pinning (what makes this oracle trustworthy without running the
// reference): see tests/test_oracle_*.py -- analytic line integrals through
// constant boxes, sample-count closed forms, scipy FP64 filters,
// numpy masked Pearson correlation, rotation invariants, glibc rand()
// sequences and the Numerical Recipes sin(R)/R simplex minimum.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// ---------------------------------------------------------------------------
// Volume preparation (a1): flip -> crop to non-zero -> copy with min/max.
// Reference: libautoscoper/src/VolumeDescription.cpp:59-75 (flipVolume),
// :77-119 (cropVolume), :121-148 (copyVolume), :152-305 (constructor).
// ---------------------------------------------------------------------------
template <class T>
void flip_volume(const T* data, T* dest, int W, int H, int D, bool fx, bool fy, bool fz)
{
  for (int k = 0; k < D; k++)
    for (int i = 0; i < H; i++)
      for (int j = 0; j < W; j++) {
        int x = fx ? (W - 1) - j : j;
        int y = fy ? (H - 1) - i : i;
        int z = fz ? (D - 1) - k : k;
        dest[(size_t)z * W * H + (size_t)y * W + x] = data[(size_t)k * W * H + (size_t)i * W + j];
      }
}

template <class T>
void crop_volume(const T* data, int W, int H, int D, int* mn, int* mx)
{
  mn[0] = W; mn[1] = H; mn[2] = D;
  mx[0] = 0; mx[1] = 0; mx[2] = 0;
  const T* p = data;
  for (int k = 0; k < D; k++) {
    bool nz_slice = false;
    for (int i = 0; i < H; i++) {
      bool nz_row = false;
      for (int j = 0; j < W; j++) {
        if (*p++ != T(0)) {
          if (j < mn[0]) mn[0] = j;
          if (j > mx[0]) mx[0] = j;
          nz_row = true;
        }
      }
      if (nz_row) {
        if (i < mn[1]) mn[1] = i;
        if (i > mx[1]) mx[1] = i;
        nz_slice = true;
      }
    }
    if (nz_slice) {
      if (k < mn[2]) mn[2] = k;
      if (k > mx[2]) mx[2] = k;
    }
  }
}

template <class T>
void copy_volume(T* dest, const T* src, int W, int H, const int* mn, const int* mx, T* vmin, T* vmax)
{
  *vmin = std::numeric_limits<T>::max();
  *vmax = std::numeric_limits<T>::min();
  for (int k = mn[2]; k < mx[2] + 1; k++)
    for (int i = mn[1]; i < mx[1] + 1; i++)
      for (int j = mn[0]; j < mx[0] + 1; j++) {
        T v = src[(size_t)k * W * H + (size_t)i * W + j];
        if (v < *vmin) *vmin = v;
        if (v > *vmax) *vmax = v;
        *dest++ = v;
      }
}

// ---------------------------------------------------------------------------
// OpenCL 1.2 sampler (spec section 8.2) for CLK_NORMALIZED_COORDS_TRUE |
// CLK_ADDRESS_CLAMP | CLK_FILTER_LINEAR, as used by
// RayCaster.cl:13-15, RadRenderer.cl:11-13, BackgroundRenderer.cl:12-14.
//   u = s*W; i0 = floor(u - 0.5); a = frac(u - 0.5) = (u-0.5) - floor(u-0.5)
//   texels outside [0,W-1] read the border colour (0 for CL_R)
//   T = sum over the 8 (4) corners of weight-product * texel, in the spec's
//       corner order.
// UNORM_INT16 texel = u16 / 65535, UNORM_INT8 texel = u8 / 255.
// ---------------------------------------------------------------------------
template <typename R>
inline R unorm16(uint16_t v) { return R(v) / R(65535); }

template <typename R>
inline int floor_index(R x, R& frac)
{
  R f = std::floor(x);
  frac = x - f;
  if (!(f > R(-4))) return -4;          // far outside: any border index
  if (!(f < R(1 << 30))) return 1 << 30;
  return (int)f;
}

template <typename R>
inline R vtex(const uint16_t* v, int W, int H, int D, int i, int j, int k)
{
  if (i < 0 || j < 0 || k < 0 || i >= W || j >= H || k >= D) return R(0);
  return unorm16<R>(v[((size_t)k * H + j) * W + i]);
}

// The un-normalisation s*W - 0.5 as an OpenCL compiler emits it: FP_CONTRACT
// is on by default in OpenCL C, so clang (PoCL's compiler) fuses it into one
// FMA, one rounding. The 2-D sampler (sample2) follows the same convention:
// both are the one read_imagef rule of OpenCL 1.2 section 8.2.
template <typename R>
R sample3(const uint16_t* v, int W, int H, int D, R s, R t, R r)
{
  R x = std::fma(s, R(W), R(-0.5));
  R y = std::fma(t, R(H), R(-0.5));
  R z = std::fma(r, R(D), R(-0.5));
  R a, b, c;
  int i0 = floor_index(x, a), j0 = floor_index(y, b), k0 = floor_index(z, c);
  int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
  R T000 = vtex<R>(v, W, H, D, i0, j0, k0), T100 = vtex<R>(v, W, H, D, i1, j0, k0);
  R T010 = vtex<R>(v, W, H, D, i0, j1, k0), T110 = vtex<R>(v, W, H, D, i1, j1, k0);
  R T001 = vtex<R>(v, W, H, D, i0, j0, k1), T101 = vtex<R>(v, W, H, D, i1, j0, k1);
  R T011 = vtex<R>(v, W, H, D, i0, j1, k1), T111 = vtex<R>(v, W, H, D, i1, j1, k1);
  R one = R(1);
  return (one - a) * (one - b) * (one - c) * T000 + a * (one - b) * (one - c) * T100 +
         (one - a) * b * (one - c) * T010 + a * b * (one - c) * T110 +
         (one - a) * (one - b) * c * T001 + a * (one - b) * c * T101 +
         (one - a) * b * c * T011 + a * b * c * T111;
}

// 2-D image texel: bps 8 -> UNORM8, 16 -> UNORM16, 32 -> float as stored.
template <typename R>
inline R itex(const void* img, int bps, int W, int H, int i, int j)
{
  if (i < 0 || j < 0 || i >= W || j >= H) return R(0);
  size_t idx = (size_t)j * W + i;
  if (bps == 8) return R(((const uint8_t*)img)[idx]) / R(255);
  if (bps == 16) return R(((const uint16_t*)img)[idx]) / R(65535);
  return R(((const float*)img)[idx]);
}

// The same un-normalisation as sample3: one FMA per axis.
template <typename R>
R sample2(const void* img, int bps, int W, int H, R s, R t)
{
  R x = std::fma(s, R(W), R(-0.5));
  R y = std::fma(t, R(H), R(-0.5));
  R a, b;
  int i0 = floor_index(x, a), j0 = floor_index(y, b);
  R T00 = itex<R>(img, bps, W, H, i0, j0), T10 = itex<R>(img, bps, W, H, i0 + 1, j0);
  R T01 = itex<R>(img, bps, W, H, i0, j0 + 1), T11 = itex<R>(img, bps, W, H, i0 + 1, j0 + 1);
  R one = R(1);
  return (one - a) * (one - b) * T00 + a * (one - b) * T10 + (one - a) * b * T01 + a * b * T11;
}

// ---------------------------------------------------------------------------
// DRR ray march (a5). Reference: RayCaster.cl:2-70; launch RayCaster.cpp:158-193.
// min/max on the slab distances use the NaN-ignoring fmin/fmax.
// ---------------------------------------------------------------------------
// margin (optional): per pixel, the DRR contribution step*val/intensity of
// the samples whose value lies within delta*max(val, cutoff) of the cutoff.
// The sampler's arithmetic is implementation-defined in OpenCL, so two
// admissible evaluations of the same sample may fall on either side of the
// cutoff there; the margin bounds what such flips can change.
template <typename R>
int64_t drr(const uint16_t* vol, int W, int H, int D, const float* imv, const float* vp, R step,
            R intensity, R cutoff, unsigned w, unsigned h, R* out, R* margin = nullptr, R delta = R(0))
{
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (int y = 0; y < (int)h; y++) {
    for (unsigned x = 0; x < w; x++) {
      const R u = R(vp[0]) + R(vp[2]) * (R(x) / R(w));
      const R v = R(vp[1]) + R(vp[3]) * (R(y) / R(h));
      const R len = std::sqrt(u * u + v * v + R(-2) * R(-2));
      const R lx = step * (u / len), ly = step * (v / len), lz = step * (R(-2) / len);
      const R ox = R(imv[3]), oy = R(imv[7]), oz = R(imv[11]);
      const R dx = R(imv[0]) * lx + R(imv[1]) * ly + R(imv[2]) * lz;
      const R dy = R(imv[4]) * lx + R(imv[5]) * ly + R(imv[6]) * lz;
      const R dz = R(imv[8]) * lx + R(imv[9]) * ly + R(imv[10]) * lz;
      const R tbx = (R(0) - ox) / dx, tby = (R(0) - oy) / dy, tbz = (R(-1) - oz) / dz;
      const R ttx = (R(1) - ox) / dx, tty = (R(1) - oy) / dy, ttz = (R(0) - oz) / dz;
      const R nx = std::fmin(ttx, tbx), ny = std::fmin(tty, tby), nz = std::fmin(ttz, tbz);
      const R fx = std::fmax(ttx, tbx), fy = std::fmax(tty, tby), fz = std::fmax(ttz, tbz);
      R nearT = std::fmax(std::fmax(nx, ny), nz);
      const R farT = std::fmin(std::fmin(fx, fy), fz);
      R* o = out + (size_t)y * w + x;
      if (margin) margin[(size_t)y * w + x] = R(0);
      if (!(farT > nearT)) {
        *o = R(0);
        continue;
      }
      if (nearT < R(0)) nearT = R(0);
      R t = farT;
      R density = R(0), near_cut = R(0);
      while (t > nearT) {
        // ray.origin + t*ray.direction, contracted as above: fma(t, d, o)
        const R px = std::fma(t, dx, ox), py = std::fma(t, dy, oy), pz = std::fma(t, dz, oz);
        const R val = sample3<R>(vol, W, H, D, px, R(1) - py, -pz);
        density += val > cutoff ? step * val : R(0);
        if (margin && std::fabs(val - cutoff) <= delta * std::max(val, cutoff)) near_cut += step * val;
        t -= R(1);
        total++;
      }
      R r = density / intensity;
      *o = r < R(0) ? R(0) : (r > R(1) ? R(1) : r);
      if (margin) margin[(size_t)y * w + x] = near_cut / intensity;
    }
  }
  return total;
}

// Radiograph resample (a6). Reference: RadRenderer.cl:2-30.
// Background mask (a12). Reference: BackgroundRenderer.cl:2-31.
template <typename R>
void rad_like(const void* img, int bps, int IW, int IH, const float* plane, const float* vp, unsigned w,
              unsigned h, R* out, bool background, R threshold)
{
  for (unsigned y = 0; y < h; y++)
    for (unsigned x = 0; x < w; x++) {
      const R u = R(vp[0]) + R(vp[2]) * (R(x) / R(w));
      const R v = R(vp[1]) + R(vp[3]) * (R(y) / R(h));
      const R s = (u - R(plane[0])) / R(plane[2]) + R(0.5);
      const R t = R(0.5) - (v - R(plane[1])) / R(plane[3]);
      R* o = out + (size_t)y * w + x;
      if (s < R(0) || t < R(0) || s > R(1) || t > R(1)) {
        *o = R(0);
      } else if (background) {
        *o = (threshold <= sample2<R>(img, bps, IW, IH, s, t)) ? R(1) : R(0);
      } else {
        *o = R(1) - sample2<R>(img, bps, IW, IH, s, t);
      }
    }
}

// Sobel (a8). Reference: SobelFilter.cl:2-48.
template <typename R>
void sobel(const R* in, R* out, int w, int h, R scale, R blend)
{
  for (int y1 = 0; y1 < h; y1++)
    for (int x1 = 0; x1 < w; x1++) {
      int x0 = std::max(x1 - 1, 0), y0 = std::max(y1 - 1, 0);
      int x2 = std::min(x1 + 1, w - 1), y2 = std::min(y1 + 1, h - 1);
      R p00 = in[y0 * w + x0], p01 = in[y0 * w + x1], p02 = in[y0 * w + x2];
      R p10 = in[y1 * w + x0], p11 = in[y1 * w + x1], p12 = in[y1 * w + x2];
      R p20 = in[y2 * w + x0], p21 = in[y2 * w + x1], p22 = in[y2 * w + x2];
      R horz = p02 + R(2) * p12 + p22 - p00 - R(2) * p10 - p20;
      R vert = p00 + R(2) * p01 + p02 - p20 - R(2) * p21 - p22;
      R grad = std::sqrt(horz * horz + vert * vert);
      R sum;
      if (blend < R(0.5))
        sum = scale * grad + R(2) * blend * p11;
      else
        sum = R(2) * (R(1) - blend) * scale * grad + p11;
      if (sum < R(0)) sum = R(0);
      else if (sum > R(1)) sum = R(1);
      out[y1 * w + x1] = sum;
    }
}

// Contrast (a9). Reference: ContrastFilter.cl:2-39 (pow is the float pow).
template <typename R>
void contrast(const R* in, R* out, int w, int h, R alpha, R beta, int size)
{
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      R fxy = in[y * w + x];
      R n = R(0), sum = R(0);
      int minI = std::max(y - size / 2, 0), maxI = std::min(y + (size + 1) / 2, h);
      int minJ = std::max(x - size / 2, 0), maxJ = std::min(x + (size + 1) / 2, w);
      for (int i = minI; i < maxI; ++i)
        for (int j = minJ; j < maxJ; ++j) {
          n += R(1);
          sum += in[i * w + j];
        }
      R axy = sum / n;
      R gxy = R(0);
      if (axy > R(0.01f)) gxy = std::pow(axy, alpha - beta) * std::pow(fxy, beta);
      out[y * w + x] = gxy;
    }
}

// Gaussian / sharpen weights. Reference: GaussianFilter.cpp:75-127 and
// SharpenFilter.cpp:103-147: R = (int)(3*radius); N = 2R+1;
// w[i][j] = float(exp(((i-R)^2 + (j-R)^2) / (-2*radius))), normalised by
// the float running sum. Returns N (1 means "copy").
int gauss_weights_f(float radius, std::vector<float>& wts)
{
  if (radius < 0) radius = 0;
  int fr = 3 * radius;
  int n = 2 * fr + 1;
  wts.assign((size_t)n * n, 0.0f);
  if (n == 1) return 1;
  float sum = 0.0f;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      wts[i * n + j] = exp(((i - fr) * (i - fr) + (j - fr) * (j - fr)) / (-2.0 * radius));
      sum = sum + wts[i * n + j];
    }
  for (int i = 0; i < n * n; ++i) wts[i] = wts[i] / sum;
  return n;
}

int gauss_weights_d(double radius, std::vector<double>& wts)
{
  if (radius < 0) radius = 0;
  int fr = (int)(3 * (float)radius);
  int n = 2 * fr + 1;
  wts.assign((size_t)n * n, 0.0);
  if (n == 1) return 1;
  double sum = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      wts[i * n + j] = exp(((i - fr) * (i - fr) + (j - fr) * (j - fr)) / (-2.0 * radius));
      sum += wts[i * n + j];
    }
  for (int i = 0; i < n * n; ++i) wts[i] /= sum;
  return n;
}

// Zero-padded N x N convolution. Reference: GaussianFilter.cl:15-27,
// SharpenFilter.cl:18-30 (i walks x, j walks y; w index i*N+j).
template <typename R>
R conv_at(const R* in, int w, int h, const R* wts, int n, int x, int y)
{
  R acc = R(0);
  int fr = (n - 1) / 2;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      int a = x - fr + i, b = y - fr + j;
      if (!(a < 0 || a >= w || b < 0 || b >= h)) acc = acc + wts[i * n + j] * in[b * w + a];
    }
  return acc;
}

// Gaussian (a10). Reference: GaussianFilter.cl:2-35, GaussianFilter.cpp:129-150.
template <typename R>
void gaussian(const R* in, R* out, int w, int h, const R* wts, int n)
{
  if (n == 1) {
    std::memcpy(out, in, sizeof(R) * (size_t)w * h);
    return;
  }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      R c = conv_at(in, w, h, wts, n, x, y);
      if (c > R(1)) c = R(1);
      if (c < R(0)) c = R(0);
      out[y * w + x] = c;
    }
}

// Sharpen (a11). Reference: SharpenFilter.cl:2-46, SharpenFilter.cpp:80-101,149-173.
template <typename R>
void sharpen(const R* in, R* out, int w, int h, const R* wts, int n, R contrast_, R threshold)
{
  if (n == 1) {
    std::memcpy(out, in, sizeof(R) * (size_t)w * h);
    return;
  }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      R blur = conv_at(in, w, h, wts, n, x, y);
      R f = in[y * w + x];
      if (std::fabs(f - blur) > threshold) {
        R o = f + contrast_ * (f - blur);
        if (o > R(1)) o = R(1);
        if (o < R(0)) o = R(0);
        out[y * w + x] = o;
      } else {
        out[y * w + x] = f;
      }
    }
}

// Filter record shared with the Python side: 6 floats
// {type, enabled, p0, p1, p2, p3}. Types follow Filter.hpp:60-66:
// 0 contrast(alpha, beta, size), 1 sobel(scale, blend), 3 gaussian(radius),
// 4 sharpen(radius, contrast, threshold).
enum { F_CONTRAST = 0, F_SOBEL = 1, F_MEDIAN = 2, F_GAUSSIAN = 3, F_SHARPEN = 4 };

template <typename R>
void apply_filter(const float* rec, const R* in, R* out, int w, int h)
{
  int type = (int)rec[0];
  bool enabled = rec[1] != 0.0f;
  size_t n = (size_t)w * h;
  if (!enabled) {
    std::memcpy(out, in, sizeof(R) * n);
    return;
  }
  switch (type) {
    case F_SOBEL: sobel<R>(in, out, w, h, R(rec[2]), R(rec[3])); break;
    case F_CONTRAST: contrast<R>(in, out, w, h, R(rec[2]), R(rec[3]), (int)rec[4]); break;
    case F_GAUSSIAN:
    case F_SHARPEN: {
      // Sharpen clamps contrast to >= 1 in its setter (SharpenFilter.cpp:92-96).
      int nn;
      std::vector<R> wr;
      if (sizeof(R) == sizeof(float)) {
        std::vector<float> wf;
        nn = gauss_weights_f(rec[2], wf);
        wr.assign(wf.begin(), wf.end());
      } else {
        std::vector<double> wd;
        nn = gauss_weights_d(rec[2], wd);
        wr.assign(wd.begin(), wd.end());
      }
      if (type == F_GAUSSIAN)
        gaussian<R>(in, out, w, h, wr.data(), nn);
      else
        sharpen<R>(in, out, w, h, wr.data(), nn, R(std::max(rec[3], 1.0f)), R(rec[4]));
      break;
    }
    default: std::memcpy(out, in, sizeof(R) * n); break;
  }
}

// View::filter (a7). Reference: View.cpp:475-555. Functionally the chain is
// out = f_k(...f_1(in)); a disabled filter is a copy, an empty chain a copy.
template <typename R>
void filter_chain(const float* recs, int nrec, const R* in, R* out, int w, int h)
{
  size_t n = (size_t)w * h;
  if (nrec == 0) {
    std::memcpy(out, in, sizeof(R) * n);
    return;
  }
  std::vector<R> a(in, in + n), b(n);
  for (int k = 0; k < nrec; k++) {
    apply_filter<R>(recs + 6 * k, a.data(), b.data(), w, h);
    std::swap(a, b);
  }
  std::memcpy(out, a.data(), sizeof(R) * n);
}

// ---------------------------------------------------------------------------
// NCC (a14). Reference: Ncc.cpp:180-220 (host), Ncc.cl:2-27 (pointwise),
// NccSum.cl:2-24 / gpu/cuda/Ncc_kernels.cu:168-187 (tree sums).
// Intended semantics (what Ncc.cpp computes with exact sums):
//   nb = sum(mask); mf = sum(f)/nb; mg = sum(g)/nb
//   num = sum_{mask>0.5} (f-mf)(g-mg); d1, d2 likewise
//   den = sqrt(d1*d2); den < 1e-5 -> 1e5; else num/den
// ---------------------------------------------------------------------------
template <typename R>
double ncc_exact(const R* f, const R* g, const R* m, size_t n, double* moments)
{
  double nb = 0, sf = 0, sg = 0;
  for (size_t i = 0; i < n; i++) {
    nb += (double)m[i];
    sf += (double)f[i];
    sg += (double)g[i];
  }
  double mf = sf / nb, mg = sg / nb;
  double num = 0, d1 = 0, d2 = 0;
  size_t cnt = 0;
  for (size_t i = 0; i < n; i++) {
    if ((double)m[i] > 0.5) {
      double a = (double)f[i] - mf, b = (double)g[i] - mg;
      num += a * b;
      d1 += a * a;
      d2 += b * b;
      cnt++;
    }
  }
  if (moments) {
    moments[0] = nb; moments[1] = sf; moments[2] = sg; moments[3] = num;
    moments[4] = d1; moments[5] = d2; moments[6] = (double)cnt;
  }
  double den = std::sqrt(d1 * d2);
  if (den < 1e-5) return 1e5;
  return num / den;
}

// Tree-sum emulation of the reference's float reductions.
//   mode 0: CUDA indexing, i = blockDim*block + thread (Ncc_kernels.cu:172),
//           256 threads (Ncc_kernels.h:46);
//   mode 1: OpenCL indexing as written, i = gid0 + gid1*gsize1 with the
//           (T,1)x(1,B) launch (NccSum.cl:8, Ncc.cpp:116-117), T = 128
//           (Ncc.cpp:80). Each stage reads all its inputs before any write
//           (the real in-place stage is racy, Ncc.cpp:138).
// Both keep the per-group LDS tree `for s = T/2; s > 0; s >>= 1` which drops
// elements when T is not a power of two.
float tree_sum(const float* f, unsigned n, int mode)
{
  const unsigned maxT = mode == 0 ? 256u : 128u;
  std::vector<float> cur(f, f + n);
  if (n == 0) return 0.0f;
  while (n > 1) {
    unsigned T = n < maxT ? n : maxT;
    unsigned B = (n + T - 1) / T;
    std::vector<float> nxt(B);
    std::vector<float> buf(T);
    for (unsigned b = 0; b < B; b++) {
      for (unsigned t = 0; t < T; t++) {
        unsigned i = mode == 0 ? b * T + t : t + b * B;
        buf[t] = i < n ? cur[i] : 0.0f;
      }
      for (unsigned s = T / 2; s > 0; s >>= 1)
        for (unsigned t = 0; t < s; t++) buf[t] += buf[t + s];
      nxt[b] = buf[0];
    }
    cur.swap(nxt);
    n = B;
  }
  return cur[0];
}

float ncc_tree(const float* f, const float* g, const float* m, unsigned n, int mode)
{
  float nb = tree_sum(m, n, mode);
  float mf = tree_sum(f, n, mode) / nb;
  float mg = tree_sum(g, n, mode) / nb;
  std::vector<float> nums(n), d1s(n), d2s(n);
  for (unsigned i = 0; i < n; i++) {
    if (m[i] > 0.5f) {
      float a = f[i] - mf, b = g[i] - mg;
      nums[i] = a * b;
      d1s[i] = a * a;
      d2s[i] = b * b;
    } else {
      nums[i] = d1s[i] = d2s[i] = 0.0f;
    }
  }
  float den = std::sqrt(tree_sum(d1s.data(), n, mode) * tree_sum(d2s.data(), n, mode));
  if (den < 1e-5) return 1e5f;
  return tree_sum(nums.data(), n, mode) / den;
}

// ---------------------------------------------------------------------------
// Rigid-frame math in double (a2/a3/a16 inputs).
// Reference: libautoscoper/src/CoordFrame.cpp. Frame = {rot[9], trans[3]};
// rot[0..2] is the image of the x axis (column-major 3x3).
// ---------------------------------------------------------------------------
struct Frame {
  double r[9];
  double t[3];
};

Frame frame_identity()
{
  Frame f;
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(f.r, I, sizeof I);
  f.t[0] = f.t[1] = f.t[2] = 0;
  return f;
}

// CoordFrame::rotate, CoordFrame.cpp:262-333 (angle in degrees).
void frame_rotate(Frame& F, const double* caxis, double angle)
{
  double axis[3] = {caxis[0], caxis[1], caxis[2]};
  double c = cos(M_PI * angle / 180.0);
  double s = sin(M_PI * angle / 180.0);
  double mag = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  if (mag == 0) return;
  axis[0] /= mag;
  axis[1] /= mag;
  axis[2] /= mag;
  double xx = axis[0] * axis[0], yy = axis[1] * axis[1], zz = axis[2] * axis[2];
  double xy = axis[0] * axis[1], xz = axis[0] * axis[2], yz = axis[1] * axis[2];
  double sx = s * axis[0], sy = s * axis[1], sz = s * axis[2];
  double oc = 1.0 - c;
  double R[9] = {(oc * xx) + c,  (oc * xy) + sz, (oc * xz) - sy, (oc * xy) - sz, (oc * yy) + c,
                 (oc * yz) + sx, (oc * xz) + sy, (oc * yz) - sx, (oc * zz) + c};
  const double* M = F.r;
  double tmp[9] = {M[0] * R[0] + M[1] * R[3] + M[2] * R[6], M[0] * R[1] + M[1] * R[4] + M[2] * R[7],
                   M[0] * R[2] + M[1] * R[5] + M[2] * R[8], M[3] * R[0] + M[4] * R[3] + M[5] * R[6],
                   M[3] * R[1] + M[4] * R[4] + M[5] * R[7], M[3] * R[2] + M[4] * R[5] + M[5] * R[8],
                   M[6] * R[0] + M[7] * R[3] + M[8] * R[6], M[6] * R[1] + M[7] * R[4] + M[8] * R[7],
                   M[6] * R[2] + M[7] * R[5] + M[8] * R[8]};
  std::memcpy(F.r, tmp, sizeof tmp);
}

// CoordFrame::from_xyzypr, CoordFrame.cpp:100-108.
Frame frame_from_xyzypr(const double* p)
{
  Frame f = frame_identity();
  frame_rotate(f, f.r + 6, p[3]);
  frame_rotate(f, f.r + 3, p[4]);
  frame_rotate(f, f.r + 0, p[5]);
  f.t[0] += p[0];
  f.t[1] += p[1];
  f.t[2] += p[2];
  return f;
}

// CoordFrame::from_xyzAxis_angle, CoordFrame.cpp:163-175.
Frame frame_from_xyz_axis_angle(const double* p)
{
  Frame f = frame_identity();
  double angle = sqrt(p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
  if (angle != 0) frame_rotate(f, p + 3, angle);
  f.t[0] += p[0];
  f.t[1] += p[1];
  f.t[2] += p[2];
  return f;
}

// CoordFrame::to_xyzypr, CoordFrame.cpp:177-189.
void frame_to_xyzypr(const Frame& F, double* p)
{
  const double* r = F.r;
  p[0] = F.t[0];
  p[1] = F.t[1];
  p[2] = F.t[2];
  p[3] = 180.0 * atan2(r[1], r[0]) / M_PI;
  p[4] = 180.0 * atan2(-r[2], sqrt(r[5] * r[5] + r[8] * r[8])) / M_PI;
  p[5] = 180 * atan2(r[5], r[8]) / M_PI;
}

// CoordFrame::inverse, CoordFrame.cpp:335-353.
Frame frame_inverse(const Frame& F)
{
  const double* q = F.r;
  Frame o;
  double r[9] = {q[0], q[3], q[6], q[1], q[4], q[7], q[2], q[5], q[8]};
  std::memcpy(o.r, r, sizeof r);
  o.t[0] = -(r[0] * F.t[0] + r[3] * F.t[1] + r[6] * F.t[2]);
  o.t[1] = -(r[1] * F.t[0] + r[4] * F.t[1] + r[7] * F.t[2]);
  o.t[2] = -(r[2] * F.t[0] + r[5] * F.t[1] + r[8] * F.t[2]);
  return o;
}

// CoordFrame::operator*, CoordFrame.cpp:478-527.
Frame frame_mul(const Frame& A, const Frame& B)
{
  const double* a = A.r;
  const double* b = B.r;
  Frame o;
  o.r[0] = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
  o.r[1] = a[1] * b[0] + a[4] * b[1] + a[7] * b[2];
  o.r[2] = a[2] * b[0] + a[5] * b[1] + a[8] * b[2];
  o.r[3] = a[0] * b[3] + a[3] * b[4] + a[6] * b[5];
  o.r[4] = a[1] * b[3] + a[4] * b[4] + a[7] * b[5];
  o.r[5] = a[2] * b[3] + a[5] * b[4] + a[8] * b[5];
  o.r[6] = a[0] * b[6] + a[3] * b[7] + a[6] * b[8];
  o.r[7] = a[1] * b[6] + a[4] * b[7] + a[7] * b[8];
  o.r[8] = a[2] * b[6] + a[5] * b[7] + a[8] * b[8];
  o.t[0] = a[0] * B.t[0] + a[3] * B.t[1] + a[6] * B.t[2] + A.t[0];
  o.t[1] = a[1] * B.t[0] + a[4] * B.t[1] + a[7] * B.t[2] + A.t[1];
  o.t[2] = a[2] * B.t[0] + a[5] * B.t[1] + a[8] * B.t[2] + A.t[2];
  return o;
}

// CoordFrame::point_to_world_space, CoordFrame.cpp:355-370.
void frame_point(const Frame& F, const double* p, double* q)
{
  const double* r = F.r;
  q[0] = r[0] * p[0] + r[3] * p[1] + r[6] * p[2] + F.t[0];
  q[1] = r[1] * p[0] + r[4] * p[1] + r[7] * p[2] + F.t[1];
  q[2] = r[2] * p[0] + r[5] * p[1] + r[8] * p[2] + F.t[2];
}

// CoordFrame::to_matrix_row_order, CoordFrame.cpp:223-241.
void frame_row_order(const Frame& F, double* m)
{
  const double* r = F.r;
  m[0] = r[0]; m[1] = r[3]; m[2] = r[6]; m[3] = F.t[0];
  m[4] = r[1]; m[5] = r[4]; m[6] = r[7]; m[7] = F.t[1];
  m[8] = r[2]; m[9] = r[5]; m[10] = r[8]; m[11] = F.t[2];
  m[12] = 0.0; m[13] = 0.0; m[14] = 0.0; m[15] = 1.0;
}

// CoordFrame::linear_extrap, CoordFrame.cpp:419-476.
Frame frame_linear_extrap(const Frame& S, const Frame& X)
{
  double t = 2.0;
  double trans[3] = {t * (X.t[0] - S.t[0]), t * (X.t[1] - S.t[1]), t * (X.t[2] - S.t[2])};
  const double* x = X.r;
  const double* r = S.r;
  double A[9] = {x[0] * r[0] + x[3] * r[3] + x[6] * r[6], x[1] * r[0] + x[4] * r[3] + x[7] * r[6],
                 x[2] * r[0] + x[5] * r[3] + x[8] * r[6], x[0] * r[1] + x[3] * r[4] + x[6] * r[7],
                 x[1] * r[1] + x[4] * r[4] + x[7] * r[7], x[2] * r[1] + x[5] * r[4] + x[8] * r[7],
                 x[0] * r[2] + x[3] * r[5] + x[6] * r[8], x[1] * r[2] + x[4] * r[5] + x[7] * r[8],
                 x[2] * r[2] + x[5] * r[5] + x[8] * r[8]};
  double axis[3] = {1.0, 0.0, 0.0};
  double trace = A[0] + A[4] + A[8];
  double ang = acos((trace - 1.0) / 2.0);
  if (ang > 0.0 && ang < M_PI) {
    axis[0] = A[5] - A[7];
    axis[1] = A[6] - A[2];
    axis[2] = A[1] - A[3];
  }
  Frame o = S;
  frame_rotate(o, axis, t * 180.0 * ang / M_PI);
  o.t[0] += trans[0];
  o.t[1] += trans[1];
  o.t[2] += trans[2];
  return o;
}

// Camera record (30 doubles): rot[9], trans[3], image_plane[12],
// viewport[4], size[2]. Reference: Camera.hpp members, Camera.cpp:424-484.
struct Cam {
  Frame frame;
  double image_plane[12];
  double viewport[4];
  double size[2];
};

Cam cam_from_record(const double* c)
{
  Cam k;
  std::memcpy(k.frame.r, c, 9 * sizeof(double));
  std::memcpy(k.frame.t, c + 9, 3 * sizeof(double));
  std::memcpy(k.image_plane, c + 12, 12 * sizeof(double));
  std::memcpy(k.viewport, c + 24, 4 * sizeof(double));
  std::memcpy(k.size, c + 28, 2 * sizeof(double));
  return k;
}

// Camera::calculateViewport / calculateImagePlane, Camera.cpp:424-484.
void cam_fill(Cam& k, double cx, double cy, double fx, double fy, double z)
{
  k.viewport[0] = -(2.0f * cx - k.size[0]) / fx;
  k.viewport[1] = -(2.0f * cy - k.size[1]) / fy;
  k.viewport[2] = 2.0f * k.size[0] / fx;
  k.viewport[3] = 2.0f * k.size[1] / fy;
  const double* T = k.frame.t;
  double distance = sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
  double scale = -1.5 * distance / z;
  double tr[3] = {scale * (k.size[0] / 2.0 - cx), scale * (k.size[1] / 2.0 - cy), scale * z};
  double c[3];
  frame_point(k.frame, tr, c);
  double hw = scale * k.size[0] / 2.0, hh = scale * k.size[1] / 2.0;
  const double* right = k.frame.r;
  const double* up = k.frame.r + 3;
  for (int i = 0; i < 3; i++) {
    k.image_plane[0 + i] = c[i] - hw * right[i] + hh * up[i];
    k.image_plane[3 + i] = c[i] - hw * right[i] - hh * up[i];
    k.image_plane[6 + i] = c[i] + hw * right[i] - hh * up[i];
    k.image_plane[9 + i] = c[i] + hw * right[i] + hh * up[i];
  }
}

// Camera::loadMayaCam2 arithmetic after parsing, Camera.cpp:229-362.
// K, R as they appear in the file (row-major text rows), t the 3 values.
Cam cam_from_mayacam2(const double* Kf, const double* Rf, const double* tf, const double* size)
{
  double K[3][3], rot[3][3], tr[3];
  // file line (4+r) holds K[.][r]; file line (9+r) holds rotation[.][r]
  for (int r = 0; r < 3; r++)
    for (int j = 0; j < 3; j++) {
      K[j][r] = Kf[r * 3 + j];
      rot[j][r] = Rf[r * 3 + j];
    }
  for (int i = 0; i < 3; i++) tr[i] = tf[i];
  Cam k;
  k.size[0] = size[0];
  k.size[1] = size[1];
  tr[0] = -tr[0];
  tr[2] = -tr[2];
  for (int i = 0; i < 3; i++) {
    rot[i][0] = -rot[i][0];
    rot[i][2] = -rot[i][2];
  }
  K[2][1] = (k.size[1] - 1) - K[2][1];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < i; j++) std::swap(rot[i][j], rot[j][i]);
  double tn[3];
  for (int i = 0; i < 3; i++) tn[i] = -(tr[0] * rot[0][i] + tr[1] * rot[1][i] + tr[2] * rot[2][i]);
  for (int i = 0; i < 3; i++) tr[i] = tn[i];
  for (int i = 0; i < 3; i++) {
    rot[0][i] = -rot[0][i];
    rot[1][i] = -rot[1][i];
  }
  std::memcpy(k.frame.r, &rot[0][0], 9 * sizeof(double));
  std::memcpy(k.frame.t, tr, 3 * sizeof(double));
  double z = -0.5 * (K[0][0] + K[1][1]);
  // calculateViewport(K[2][0], K[2][1], K[0][0], K[1][1]); calculateImagePlane(K[2][0], K[2][1], z)
  if (K[0][0] == 0 || K[1][1] == 0) throw 1;
  cam_fill(k, K[2][0], K[2][1], K[0][0], K[1][1], z);
  return k;
}

// RayCaster::setInvModelView (a2), RayCaster.cpp:92-142.
void inv_model_view(const double* m, const float* invScale, const float* invTrans, float* out)
{
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 4; c++)
      out[4 * r + c] = m[4 * r + c] * invScale[r] + m[12 + c] * invTrans[r];
  for (int c = 0; c < 4; c++) out[12 + c] = m[12 + c];
}

// Tracker::calculate_viewport (a3), Tracker.cpp:923-1032.
bool calc_viewport(const Frame& modelview, const Cam& cam, const float* invTrans, const float* invScale,
                   double* vp)
{
  double mm[4] = {1.0, 1.0, -1.0, -1.0};
  double corners[24] = {0, 0, -1, 0, 0, 0, 0, 1, -1, 0, 1, 0, 1, 0, -1, 1, 0, 0, 1, 1, -1, 1, 1, 0};
  for (int j = 0; j < 8; j++) {
    corners[3 * j + 0] = (corners[3 * j + 0] - invTrans[0]) / invScale[0];
    corners[3 * j + 1] = (corners[3 * j + 1] - invTrans[1]) / invScale[1];
    corners[3 * j + 2] = (corners[3 * j + 2] - invTrans[2]) / invScale[2];
    double q[3];
    frame_point(modelview, &corners[3 * j], q);
    double fp0 = -2 * q[0] / q[2];
    double fp1 = -2 * q[1] / q[2];
    if (mm[0] > fp0) mm[0] = fp0;
    if (mm[1] > fp1) mm[1] = fp1;
    if (mm[2] < fp0) mm[2] = fp0;
    if (mm[3] < fp1) mm[3] = fp1;
  }
  Frame ci = frame_inverse(cam.frame);
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  frame_point(ci, cam.image_plane + 3, a);
  frame_point(ci, cam.image_plane + 9, b);
  double rf[4] = {-2 * a[0] / a[2], -2 * a[1] / a[2], -2 * b[0] / b[2], -2 * b[1] / b[2]};
  if (mm[0] > rf[2] && mm[2] > rf[2]) return false;
  if (mm[1] > rf[3] && mm[3] > rf[3]) return false;
  if (mm[0] < rf[0] && mm[2] < rf[0]) return false;
  if (mm[1] < rf[1] && mm[3] < rf[1]) return false;
  if (mm[0] == 1.0 && mm[1] == 1.0 && mm[2] == -1.0 && mm[3] == -1.0) return false;
  if (mm[0] < rf[0]) mm[0] = rf[0];
  if (mm[1] < rf[1]) mm[1] = rf[1];
  if (mm[2] > rf[2]) mm[2] = rf[2];
  if (mm[3] > rf[3]) mm[3] = rf[3];
  vp[0] = mm[0];
  vp[1] = mm[1];
  vp[2] = mm[2] - mm[0];
  vp[3] = mm[3] - mm[1];
  return true;
}

// Quat<float>(yaw, pitch, roll) and toEuler, math/Quaternion.hpp:80-94,220-243.
struct Quatf {
  float w, x, y, z;
};

Quatf quatf_from_ypr(float yaw, float pitch, float roll)
{
  float cr = cosf(roll * float(M_PI) / float(360)), sr = sinf(roll * float(M_PI) / float(360));
  float cp = cosf(pitch * float(M_PI) / float(360)), sp = sinf(pitch * float(M_PI) / float(360));
  float cy = cosf(yaw * float(M_PI) / float(360)), sy = sinf(yaw * float(M_PI) / float(360));
  Quatf q;
  q.w = cr * cp * cy + sr * sp * sy;
  q.x = sr * cp * cy - cr * sp * sy;
  q.y = cr * sp * cy + sr * cp * sy;
  q.z = cr * cp * sy - sr * sp * cy;
  return q;
}

// returns (roll, pitch, yaw) in degrees, as Vec3(x=roll, y=pitch, z=yaw)
void quatf_to_euler(const Quatf& q, float* e)
{
  float sinr = float(2) * (q.w * q.x + q.y * q.z);
  float cosr = float(1) - float(2) * (q.x * q.x + q.y * q.y);
  float roll = atan2f(sinr, cosr);
  float sinp = sqrtf(float(1) + float(2) * (q.w * q.y - q.x * q.z));
  float cosp = sqrtf(float(1) - float(2) * (q.w * q.y - q.x * q.z));
  float pitch = float(2) * atan2f(sinp, cosp) - float(M_PI) / float(2);
  float siny = float(2) * (q.w * q.z + q.x * q.y);
  float cosy = float(1) - float(2) * (q.y * q.y + q.z * q.z);
  float yaw = atan2f(siny, cosy);
  roll *= float(180) / float(M_PI);
  pitch *= float(180) / float(M_PI);
  yaw *= float(180) / float(M_PI);
  e[0] = roll;
  e[1] = pitch;
  e[2] = yaw;
}

// ---------------------------------------------------------------------------
// One pose evaluation (a15 trackFrame + a16 minimizationFunc), in either
// arithmetic mode.
// ---------------------------------------------------------------------------
struct Scene {
  const uint16_t* vol;
  int vdim[3];
  float inv_scale[3], inv_trans[3];
  float step, intensity, cutoff;
  int ncam;
  const double* cams;       // ncam * 30
  const void* const* rads;  // ncam
  const int* rad_info;      // ncam * 3: w, h, bps
  const float* filters;     // ncam * 2 * 8 * 6
  const int* nfilters;      // ncam * 2: drr, rad
  int render_w, render_h;
  float bg_threshold;
  const float* const* bgs;  // ncam or null
  const int* bg_info;       // ncam * 2
};

// Per camera outputs: info[8] = {in_view, render_w, render_h, mask_count,
// (reserved)...}, vp[4], corr (1-ncc or 9999), moments[7].
template <typename R>
void track_frame(const Scene& S, const double* xyzypr, double* corr, int* info, double* vps, double* moments,
                 int64_t* samples_out)
{
  Frame xc = frame_from_xyzypr(xyzypr);
  int64_t samples = 0;
  for (int c = 0; c < S.ncam; c++) {
    Cam cam = cam_from_record(S.cams + 30 * c);
    int* inf = info ? info + 8 * c : nullptr;
    if (inf) std::memset(inf, 0, 8 * sizeof(int));
    Frame mv = frame_mul(frame_inverse(cam.frame), xc);
    double imvd[16];
    frame_row_order(frame_inverse(mv), imvd);
    float imv[16];
    inv_model_view(imvd, S.inv_scale, S.inv_trans, imv);
    double vp[4];
    if (!calc_viewport(mv, cam, S.inv_trans, S.inv_scale, vp)) {
      corr[c] = 9999.0;
      continue;
    }
    unsigned rw = vp[2] * S.render_w / cam.viewport[2];
    unsigned rh = vp[3] * S.render_h / cam.viewport[3];
    if (vps) std::memcpy(vps + 4 * c, vp, sizeof vp);
    if (rw > (unsigned)S.render_w || rh > (unsigned)S.render_h) {
      corr[c] = -1.0;  // the reference throws (Tracker.cpp:625-629)
      if (inf) inf[0] = -1;
      continue;
    }
    float vpf[4] = {(float)vp[0], (float)vp[1], (float)vp[2], (float)vp[3]};
    size_t n = (size_t)rw * rh;
    std::vector<R> drr_raw(n), drr_f(n), rad_raw(n), rad_f(n), mask(n, R(1)), bg(n, R(1));
    samples += drr<R>(S.vol, S.vdim[0], S.vdim[1], S.vdim[2], imv, vpf, R(S.step), R(S.intensity), R(S.cutoff),
                      rw, rh, drr_raw.data());
    const float* fl = S.filters + (size_t)c * 2 * 8 * 6;
    filter_chain<R>(fl, S.nfilters[2 * c], drr_raw.data(), drr_f.data(), rw, rh);
    float plane[4] = {(float)cam.viewport[0], (float)cam.viewport[1], (float)cam.viewport[2],
                      (float)cam.viewport[3]};
    const int* ri = S.rad_info + 3 * c;
    rad_like<R>(S.rads[c], ri[2], ri[0], ri[1], plane, vpf, rw, rh, rad_raw.data(), false, R(0));
    filter_chain<R>(fl + 8 * 6, S.nfilters[2 * c + 1], rad_raw.data(), rad_f.data(), rw, rh);
    if (S.bg_threshold >= 0.0f && S.bgs && S.bgs[c])
      rad_like<R>(S.bgs[c], 32, S.bg_info[2 * c], S.bg_info[2 * c + 1], plane, vpf, rw, rh, bg.data(), true,
                  R(S.bg_threshold));
    int cnt = 0;
    for (size_t i = 0; i < n; i++) {
      R m = drr_f[i] != R(0) ? R(1) : R(0);
      m = bg[i] * m;
      mask[i] = m;
      rad_f[i] = rad_f[i] * m;
      drr_f[i] = drr_f[i] * m;
      cnt += m > R(0.5);
    }
    double mom[7];
    double ncc = ncc_exact<R>(drr_f.data(), rad_f.data(), mask.data(), n, mom);
    corr[c] = 1.0 - ncc;
    if (moments) std::memcpy(moments + 7 * c, mom, sizeof mom);
    if (inf) {
      inf[0] = 1;
      inf[1] = (int)rw;
      inf[2] = (int)rh;
      inf[3] = cnt;
    }
  }
  if (samples_out) *samples_out = samples;
}

// Pose store: float x, y, z and a Quatf, as the reference's KeyCurve<float>
// and KeyCurve<Quatf> hold them (Tracker.cpp:777-784, KeyCurve.hpp:330-356).
struct PoseStore {
  float x, y, z;
  Quatf q;
};

void pose_store_get(const PoseStore& P, double* xyzypr)
{
  float e[3];
  quatf_to_euler(P.q, e);
  xyzypr[0] = P.x; xyzypr[1] = P.y; xyzypr[2] = P.z;
  xyzypr[3] = e[2]; xyzypr[4] = e[1]; xyzypr[5] = e[0];
}

void pose_store_set(PoseStore& P, const double* xyzypr)
{
  P.x = (float)xyzypr[0];
  P.y = (float)xyzypr[1];
  P.z = (float)xyzypr[2];
  P.q = quatf_from_ypr((float)xyzypr[3], (float)xyzypr[4], (float)xyzypr[5]);
}

// Tracker::minimizationFunc composition (a16), Tracker.cpp:775-827.
void compose_pose(const PoseStore& P, const Frame& pivot, const double* manip, double* out_xyzypr)
{
  double base[6];
  pose_store_get(P, base);
  Frame xc = frame_from_xyzypr(base);
  Frame mf = frame_from_xyz_axis_angle(manip);
  xc = frame_mul(frame_mul(frame_mul(xc, frame_inverse(pivot)), mf), pivot);
  frame_to_xyzypr(xc, out_xyzypr);
}

template <typename R>
double min_func(const Scene& S, const PoseStore& P, const Frame& pivot, const double* manip)
{
  double xyzypr[6];
  compose_pose(P, pivot, manip, xyzypr);
  std::vector<double> corr(S.ncam);
  track_frame<R>(S, xyzypr, corr.data(), nullptr, nullptr, nullptr, nullptr);
  double c = corr[0];
  for (int i = 1; i < S.ncam; i++) c += corr[i];
  if (c < 0) c = 9999;
  return c;
}

// ---------------------------------------------------------------------------
// PSO (a17), PSO.cpp:83-159,198-295, PSO.hpp:17-20, with glibc rand().
// ---------------------------------------------------------------------------
typedef double (*cost_fn)(const double* x, void* user);

struct Particle {
  float ncc;
  float pos[6];
  float vel[6];
};

float rand_clamped() { return (float)rand() / (float)RAND_MAX; }
float rand_range(float lo, float hi) { return lo + rand_clamped() * (hi - lo); }

// synchronous = true is the batched variant of the product (not in the
// reference): every particle of an epoch moves against the gBest of the
// previous epoch, all are evaluated, then pBest/gBest are updated in
// particle order. The random draws are taken in the same order as the
// reference's.
int pso(cost_fn fn, void* user, float lo, float hi, unsigned max_epochs, unsigned max_stall, int nparticles,
        float* best_pos, float* best_ncc, bool synchronous = false)
{
  const float c1 = 1.5f, c2 = 1.5f;
  int stall = 0;
  bool go = true;
  unsigned counter = 0;
  float omega = 0.8f;
  std::vector<Particle> p(nparticles);
  for (int i = 0; i < nparticles; i++) {
    p[i].ncc = FLT_MAX;
    for (int d = 0; d < 6; d++) {
      p[i].pos[d] = 0.f;
      p[i].vel[d] = 0.f;
    }
    if (i > 0)
      for (int d = 0; d < 6; d++) p[i].pos[d] = rand_range(lo, hi);
  }
  Particle g;
  g.ncc = FLT_MAX;
  for (int d = 0; d < 6; d++) g.pos[d] = g.vel[d] = 0.f;
  std::vector<Particle> pb = p;
  while (go) {
    if (counter >= max_epochs) go = false;
    Particle cur = g;
    const Particle g_epoch = g;
    for (int i = 0; i < nparticles; i++) {
      const Particle& gb = synchronous ? g_epoch : g;
      for (int d = 0; d < 6; d++) {
        float rp = rand_clamped();
        float rg = rand_clamped();
        p[i].vel[d] = omega * p[i].vel[d] + c1 * rp * (pb[i].pos[d] - p[i].pos[d]) +
                      c2 * rg * (gb.pos[d] - p[i].pos[d]);
        p[i].pos[d] += p[i].vel[d];
      }
      double x[6];
      for (int d = 0; d < 6; d++) x[d] = (double)p[i].pos[d];
      p[i].ncc = (float)fn(x, user);
      if (p[i].ncc < pb[i].ncc) pb[i] = p[i];
      if (p[i].ncc < g.ncc) g = p[i];
    }
    omega = omega * 0.9f;
    float diff = std::fabs(g.ncc - cur.ncc);
    if (diff < 1e-4f)
      stall++;
    else if (diff > 0.001f)
      stall = 0;
    if (stall == (int)max_stall) go = false;
    counter++;
  }
  std::memcpy(best_pos, g.pos, sizeof g.pos);
  *best_ncc = g.ncc;
  return (int)counter;
}

// ---------------------------------------------------------------------------
// Downhill simplex (a18), DownhillSimplex.cpp:39-172. P is (ndim+1) x ndim,
// row-major, 0-indexed here; the algorithm is the 1-indexed original.
// ---------------------------------------------------------------------------
int amoeba(cost_fn fn, void* user, double* P0, double* Y0, int ndim, double ftol)
{
  const int ITMAX = 1000;
  const double ALPHA = 1.0, BETA = 0.5, GAMMA = 2.0;
  int mpts = ndim + 1;
  auto P = [&](int i, int j) -> double& { return P0[(i - 1) * ndim + (j - 1)]; };
  auto Y = [&](int i) -> double& { return Y0[i - 1]; };
  std::vector<double> PR(ndim + 1), PRR(ndim + 1), PBAR(ndim + 1), CENT(ndim + 1);
  int iter = 0;
  for (;;) {
    int ilo = 1, ihi, inhi;
    if (Y(1) > Y(2)) {
      ihi = 1;
      inhi = 2;
    } else {
      ihi = 2;
      inhi = 1;
    }
    for (int i = 1; i <= mpts; i++) {
      if (Y(i) < Y(ilo)) ilo = i;
      if (Y(i) > Y(ihi)) {
        inhi = ihi;
        ihi = i;
      } else if (Y(i) > Y(inhi))
        if (i != ihi) inhi = i;
    }
    for (int j = 1; j <= ndim; j++) {
      CENT[j] = 0;
      for (int i = 1; i <= mpts; i++) CENT[j] += P(i, j);
      CENT[j] /= mpts;
    }
    double rtol = 0;
    for (int i = 1; i <= mpts; i++) {
      double ss = 0;
      for (int j = 1; j <= ndim; j++) ss += (CENT[j] - P(i, j)) * (CENT[j] - P(i, j));
      rtol += sqrt(ss);
    }
    rtol /= mpts;
    if (rtol < ftol) return iter;
    if (iter == ITMAX) return iter;
    iter++;
    for (int j = 1; j <= ndim; j++) PBAR[j] = 0.0;
    for (int i = 1; i <= mpts; i++)
      if (i != ihi)
        for (int j = 1; j <= ndim; j++) PBAR[j] += P(i, j);
    for (int j = 1; j <= ndim; j++) {
      PBAR[j] /= 1.0 * ndim;
      PR[j] = (1.0 + ALPHA) * PBAR[j] - ALPHA * P(ihi, j);
    }
    double ypr = fn(&PR[1], user);
    if (ypr <= Y(ilo)) {
      for (int j = 1; j <= ndim; j++) PRR[j] = GAMMA * PR[j] + (1.0 - GAMMA) * PBAR[j];
      double yprr = fn(&PRR[1], user);
      if (yprr < Y(ilo)) {
        for (int j = 1; j <= ndim; j++) P(ihi, j) = PRR[j];
        Y(ihi) = yprr;
      } else {
        for (int j = 1; j <= ndim; j++) P(ihi, j) = PR[j];
        Y(ihi) = ypr;
      }
    } else if (ypr >= Y(inhi)) {
      if (ypr < Y(ihi)) {
        for (int j = 1; j <= ndim; j++) P(ihi, j) = PR[j];
        Y(ihi) = ypr;
      }
      for (int j = 1; j <= ndim; j++) PRR[j] = BETA * P(ihi, j) + (1.0 - BETA) * PBAR[j];
      double yprr = fn(&PRR[1], user);
      if (yprr < Y(ihi)) {
        for (int j = 1; j <= ndim; j++) P(ihi, j) = PRR[j];
        Y(ihi) = yprr;
      } else {
        for (int i = 1; i <= mpts; i++)
          if (i != ilo) {
            for (int j = 1; j <= ndim; j++) {
              PR[j] = 0.5 * (P(i, j) + P(ilo, j));
              P(i, j) = PR[j];
            }
            Y(i) = fn(&PR[1], user);
          }
      }
    } else {
      for (int j = 1; j <= ndim; j++) P(ihi, j) = PR[j];
      Y(ihi) = ypr;
    }
  }
}

struct MinCtx {
  const Scene* S;
  const PoseStore* P;
  Frame pivot;
  int fp64;
  int64_t evals;
};

double min_func_cb(const double* manip, void* user)
{
  MinCtx* m = (MinCtx*)user;
  m->evals++;
  return m->fp64 ? min_func<double>(*m->S, *m->P, m->pivot, manip)
                 : min_func<float>(*m->S, *m->P, m->pivot, manip);
}

}  // namespace

// ===========================================================================
// C entry points (loaded through ctypes by the tests and the CPU baseline).
// ===========================================================================

ORC_API int orc_volume_prep(const void* data, int W, int H, int D, int bps, const int* flip,
                            const double* scale, uint16_t* out, int* dims, float* inv_scale, float* inv_trans,
                            double* trans_center, float* minmax, int* crop)
{
  int mn[3], mx[3];
  if (bps != 8 && bps != 16) return -1;
  size_t nvox = (size_t)W * H * D;
  if (bps == 16) {
    std::vector<uint16_t> fl(nvox);
    flip_volume((const uint16_t*)data, fl.data(), W, H, D, flip[0], flip[1], flip[2]);
    crop_volume(fl.data(), W, H, D, mn, mx);
    if (mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]) return -2;
    uint16_t vmin, vmax;
    copy_volume(out, fl.data(), W, H, mn, mx, &vmin, &vmax);
    minmax[0] = vmin / (float)65535;
    minmax[1] = vmax / (float)65535;
  } else {
    std::vector<uint8_t> fl(nvox);
    flip_volume((const uint8_t*)data, fl.data(), W, H, D, flip[0], flip[1], flip[2]);
    crop_volume(fl.data(), W, H, D, mn, mx);
    if (mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]) return -2;
    std::vector<uint8_t> tmp((size_t)(mx[0] - mn[0] + 1) * (mx[1] - mn[1] + 1) * (mx[2] - mn[2] + 1));
    uint8_t vmin, vmax;
    copy_volume(tmp.data(), fl.data(), W, H, mn, mx, &vmin, &vmax);
    // 8-bit volumes are widened to the 16-bit grid the sampler reads
    // (UNORM8 v/255 == UNORM16 257v/65535 exactly).
    for (size_t i = 0; i < tmp.size(); i++) out[i] = (uint16_t)(tmp[i] * 257);
    minmax[0] = vmin / (float)255;
    minmax[1] = vmax / (float)255;
  }
  int dim[3] = {mx[0] - mn[0] + 1, mx[1] - mn[1] + 1, mx[2] - mn[2] + 1};
  for (int i = 0; i < 3; i++) {
    dims[i] = dim[i];
    crop[i] = mn[i];
    crop[3 + i] = mx[i];
  }
  trans_center[0] = -scale[0] * 0.5 * (mn[0] + mx[0]);
  trans_center[1] = -scale[1] * 0.5 * (mn[1] + mx[1]);
  trans_center[2] = scale[2] * 0.5 * (mn[2] + mx[2]);
  inv_scale[0] = 1.0f / (float)(scale[0] * dim[0]);
  inv_scale[1] = 1.0f / (float)(scale[1] * dim[1]);
  inv_scale[2] = 1.0f / (float)(scale[2] * dim[2]);
  inv_trans[0] = -mn[0] / (float)dim[0];
  inv_trans[1] = -((H - mx[1] - 1) / (float)dim[1]);
  inv_trans[2] = mn[2] / (float)dim[2];
  return 0;
}

ORC_API int64_t orc_drr_f32(const uint16_t* vol, const int* vdim, const float* imv, const float* vp, float step,
                            float intensity, float cutoff, unsigned w, unsigned h, float* out)
{
  return drr<float>(vol, vdim[0], vdim[1], vdim[2], imv, vp, step, intensity, cutoff, w, h, out);
}

// orc_drr_f32 plus the cutoff margin of every pixel (see drr)
ORC_API int64_t orc_drr_margin_f32(const uint16_t* vol, const int* vdim, const float* imv, const float* vp,
                                   float step, float intensity, float cutoff, float delta, unsigned w, unsigned h,
                                   float* out, float* margin)
{
  return drr<float>(vol, vdim[0], vdim[1], vdim[2], imv, vp, step, intensity, cutoff, w, h, out, margin, delta);
}

ORC_API int64_t orc_drr_f64(const uint16_t* vol, const int* vdim, const float* imv, const float* vp, float step,
                            float intensity, float cutoff, unsigned w, unsigned h, double* out)
{
  return drr<double>(vol, vdim[0], vdim[1], vdim[2], imv, vp, step, intensity, cutoff, w, h, out);
}

ORC_API void orc_rad_f32(const void* img, int bps, int iw, int ih, const float* plane, const float* vp, unsigned w,
                         unsigned h, float* out)
{
  rad_like<float>(img, bps, iw, ih, plane, vp, w, h, out, false, 0.0f);
}

ORC_API void orc_background_f32(const float* img, int iw, int ih, const float* plane, const float* vp,
                                float threshold, unsigned w, unsigned h, float* out)
{
  rad_like<float>(img, 32, iw, ih, plane, vp, w, h, out, true, threshold);
}

ORC_API void orc_sobel_f32(const float* in, float* out, int w, int h, float scale, float blend)
{
  sobel<float>(in, out, w, h, scale, blend);
}

ORC_API void orc_contrast_f32(const float* in, float* out, int w, int h, float alpha, float beta, int size)
{
  contrast<float>(in, out, w, h, alpha, beta, size);
}

ORC_API int orc_gaussian_weights(float radius, float* out, int cap)
{
  std::vector<float> w;
  int n = gauss_weights_f(radius, w);
  if (n * n > cap) return -n;
  std::memcpy(out, w.data(), sizeof(float) * n * n);
  return n;
}

ORC_API void orc_gaussian_f32(const float* in, float* out, int w, int h, float radius)
{
  std::vector<float> wt;
  int n = gauss_weights_f(radius, wt);
  gaussian<float>(in, out, w, h, wt.data(), n);
}

ORC_API void orc_sharpen_f32(const float* in, float* out, int w, int h, float radius, float contrast_,
                             float threshold)
{
  std::vector<float> wt;
  int n = gauss_weights_f(radius, wt);
  sharpen<float>(in, out, w, h, wt.data(), n, std::max(contrast_, 1.0f), threshold);
}

ORC_API void orc_filter_chain_f32(const float* recs, int n, const float* in, float* out, int w, int h)
{
  filter_chain<float>(recs, n, in, out, w, h);
}

ORC_API void orc_filter_chain_f64(const float* recs, int n, const double* in, double* out, int w, int h)
{
  filter_chain<double>(recs, n, in, out, w, h);
}

ORC_API double orc_ncc_exact_f32in(const float* f, const float* g, const float* m, size_t n, double* moments)
{
  return ncc_exact<float>(f, g, m, n, moments);
}

ORC_API double orc_ncc_exact_f64in(const double* f, const double* g, const double* m, size_t n, double* moments)
{
  return ncc_exact<double>(f, g, m, n, moments);
}

ORC_API float orc_ncc_tree(const float* f, const float* g, const float* m, unsigned n, int mode)
{
  return ncc_tree(f, g, m, n, mode);
}

ORC_API float orc_tree_sum(const float* f, unsigned n, int mode) { return tree_sum(f, n, mode); }

// Frames are passed as 12 doubles {rot[9], trans[3]}.
ORC_API void orc_frame_from_xyzypr(const double* p, double* f)
{
  Frame F = frame_from_xyzypr(p);
  std::memcpy(f, F.r, 9 * sizeof(double));
  std::memcpy(f + 9, F.t, 3 * sizeof(double));
}

ORC_API void orc_frame_from_xyz_axis_angle(const double* p, double* f)
{
  Frame F = frame_from_xyz_axis_angle(p);
  std::memcpy(f, F.r, 9 * sizeof(double));
  std::memcpy(f + 9, F.t, 3 * sizeof(double));
}

static Frame frame_load(const double* f)
{
  Frame F;
  std::memcpy(F.r, f, 9 * sizeof(double));
  std::memcpy(F.t, f + 9, 3 * sizeof(double));
  return F;
}

ORC_API void orc_frame_to_xyzypr(const double* f, double* p) { frame_to_xyzypr(frame_load(f), p); }

ORC_API void orc_frame_mul(const double* a, const double* b, double* o)
{
  Frame F = frame_mul(frame_load(a), frame_load(b));
  std::memcpy(o, F.r, 9 * sizeof(double));
  std::memcpy(o + 9, F.t, 3 * sizeof(double));
}

ORC_API void orc_frame_inverse(const double* a, double* o)
{
  Frame F = frame_inverse(frame_load(a));
  std::memcpy(o, F.r, 9 * sizeof(double));
  std::memcpy(o + 9, F.t, 3 * sizeof(double));
}

ORC_API void orc_frame_linear_extrap(const double* a, const double* b, double* o)
{
  Frame F = frame_linear_extrap(frame_load(a), frame_load(b));
  std::memcpy(o, F.r, 9 * sizeof(double));
  std::memcpy(o + 9, F.t, 3 * sizeof(double));
}

ORC_API int orc_camera_from_mayacam2(const double* K, const double* R, const double* t, const double* size,
                                     double* rec)
{
  Cam k;
  try {
    k = cam_from_mayacam2(K, R, t, size);
  } catch (...) {
    return -1;
  }
  std::memcpy(rec, k.frame.r, 9 * sizeof(double));
  std::memcpy(rec + 9, k.frame.t, 3 * sizeof(double));
  std::memcpy(rec + 12, k.image_plane, 12 * sizeof(double));
  std::memcpy(rec + 24, k.viewport, 4 * sizeof(double));
  std::memcpy(rec + 28, k.size, 2 * sizeof(double));
  return 0;
}

// (a2) + (a3) + render-size truncation (Tracker.cpp:604-629) for one camera.
// Returns 1 in view, 0 not in view; imv_out[16] floats; vp_out[4];
// size_out[2] = render width/height.
ORC_API int orc_view_geometry(const double* cam_rec, const double* xyzypr, const float* inv_scale,
                              const float* inv_trans, int render_w, int render_h, float* imv_out, double* vp_out,
                              unsigned* size_out)
{
  Cam cam = cam_from_record(cam_rec);
  Frame xc = frame_from_xyzypr(xyzypr);
  Frame mv = frame_mul(frame_inverse(cam.frame), xc);
  double imvd[16];
  frame_row_order(frame_inverse(mv), imvd);
  inv_model_view(imvd, inv_scale, inv_trans, imv_out);
  if (!calc_viewport(mv, cam, inv_trans, inv_scale, vp_out)) return 0;
  size_out[0] = vp_out[2] * render_w / cam.viewport[2];
  size_out[1] = vp_out[3] * render_h / cam.viewport[3];
  return 1;
}

ORC_API void orc_quatf_roundtrip(const double* xyzypr, double* out)
{
  PoseStore P;
  pose_store_set(P, xyzypr);
  pose_store_get(P, out);
}

static Scene make_scene(const uint16_t* vol, const int* vdim, const float* vdesc, const float* drr_params, int ncam,
                        const double* cams, const void* const* rads, const int* rad_info, const float* filters,
                        const int* nfilters, int rw, int rh, float bg_threshold, const float* const* bgs,
                        const int* bg_info)
{
  Scene S;
  S.vol = vol;
  for (int i = 0; i < 3; i++) {
    S.vdim[i] = vdim[i];
    S.inv_scale[i] = vdesc[i];
    S.inv_trans[i] = vdesc[3 + i];
  }
  S.step = drr_params[0];
  S.intensity = drr_params[1];
  S.cutoff = drr_params[2];
  S.ncam = ncam;
  S.cams = cams;
  S.rads = rads;
  S.rad_info = rad_info;
  S.filters = filters;
  S.nfilters = nfilters;
  S.render_w = rw;
  S.render_h = rh;
  S.bg_threshold = bg_threshold;
  S.bgs = bgs;
  S.bg_info = bg_info;
  return S;
}

#define SCENE_ARGS                                                                                            \
  const uint16_t *vol, const int *vdim, const float *vdesc, const float *drr_params, int ncam,               \
    const double *cams, const void *const *rads, const int *rad_info, const float *filters,                  \
    const int *nfilters, int rw, int rh, float bg_threshold, const float *const *bgs, const int *bg_info
#define SCENE_PASS vol, vdim, vdesc, drr_params, ncam, cams, rads, rad_info, filters, nfilters, rw, rh, bg_threshold, bgs, bg_info

ORC_API int64_t orc_track_frame(SCENE_ARGS, const double* xyzypr, int fp64, double* corr, int* info, double* vps,
                                double* moments)
{
  Scene S = make_scene(SCENE_PASS);
  int64_t samples = 0;
  if (fp64)
    track_frame<double>(S, xyzypr, corr, info, vps, moments, &samples);
  else
    track_frame<float>(S, xyzypr, corr, info, vps, moments, &samples);
  return samples;
}

// Pose store helpers: store[7] = {x, y, z, qw, qx, qy, qz} as floats.
static PoseStore store_load(const float* s)
{
  PoseStore P;
  P.x = s[0]; P.y = s[1]; P.z = s[2];
  P.q.w = s[3]; P.q.x = s[4]; P.q.y = s[5]; P.q.z = s[6];
  return P;
}
static void store_save(const PoseStore& P, float* s)
{
  s[0] = P.x; s[1] = P.y; s[2] = P.z;
  s[3] = P.q.w; s[4] = P.q.x; s[5] = P.q.y; s[6] = P.q.z;
}

ORC_API void orc_pose_store_set(float* store, const double* xyzypr)
{
  PoseStore P;
  pose_store_set(P, xyzypr);
  store_save(P, store);
}

ORC_API void orc_pose_store_get(const float* store, double* xyzypr) { pose_store_get(store_load(store), xyzypr); }

ORC_API void orc_compose_pose(const float* store, const double* pivot, const double* manip, double* out)
{
  compose_pose(store_load(store), frame_load(pivot), manip, out);
}

ORC_API double orc_min_func(SCENE_ARGS, const float* store, const double* pivot, const double* manip, int fp64)
{
  Scene S = make_scene(SCENE_PASS);
  PoseStore P = store_load(store);
  return fp64 ? min_func<double>(S, P, frame_load(pivot), manip) : min_func<float>(S, P, frame_load(pivot), manip);
}

// Tracker::optimize for one frame with guess 0 (Tracker.cpp:337-580):
// method 0 = PSO then simplex, 1 = simplex only. store[7] is updated in
// place. Returns the number of cost evaluations; stats[4] = {pso epochs,
// simplex iterations, pso best cost, final cost}.
ORC_API int64_t orc_optimize_frame(SCENE_ARGS, float* store, const double* pivot, int method, unsigned max_iter,
                                   double min_limit, double max_limit, unsigned max_stall, const double* offsets,
                                   unsigned seed, int nparticles, int fp64, int synchronous, double* stats)
{
  Scene S = make_scene(SCENE_PASS);
  srand(seed);
  PoseStore P = store_load(store);
  MinCtx ctx{&S, &P, frame_load(pivot), fp64, 0};
  double xyzypr[6];
  pose_store_get(P, xyzypr);
  double init_manip[6] = {0, 0, 0, 0, 0, 0};
  double init_ncc = min_func_cb(init_manip, &ctx);
  Frame manip = frame_from_xyz_axis_angle(init_manip);
  double Pm[7 * 6], Y[7];
  int iters = 0;
  if (method == 0) {
    float best[6], best_ncc;
    int epochs = pso(min_func_cb, &ctx, (float)min_limit, (float)max_limit, max_iter, max_stall, nparticles, best,
                     &best_ncc, synchronous != 0);
    double bm[6];
    for (int d = 0; d < 6; d++) bm[d] = best[d];
    manip = frame_from_xyz_axis_angle(bm);
    Frame xc = frame_from_xyzypr(xyzypr);
    xc = frame_mul(frame_mul(frame_mul(xc, frame_inverse(ctx.pivot)), manip), ctx.pivot);
    frame_to_xyzypr(xc, xyzypr);
    pose_store_set(P, xyzypr);
    if (stats) {
      stats[0] = epochs;
      stats[2] = best_ncc;
    }
    for (int i = 0; i < 7; i++)
      for (int j = 0; j < 6; j++) Pm[i * 6 + j] = (i == j + 1) ? offsets[j] : 0.0;
    for (int i = 0; i < 7; i++) Y[i] = min_func_cb(Pm + 6 * i, &ctx);
    iters = amoeba(min_func_cb, &ctx, Pm, Y, 6, 1e-3);
    double final_ncc = min_func_cb(Pm, &ctx);
    if (final_ncc < init_ncc)
      manip = frame_from_xyz_axis_angle(Pm);
    else
      manip = frame_from_xyz_axis_angle(init_manip);
    if (stats) stats[3] = final_ncc;
  } else {
    for (int i = 0; i < 7; i++)
      for (int j = 0; j < 6; j++) Pm[i * 6 + j] = (i == j + 1) ? offsets[j] : 0.0;
    for (int i = 0; i < 7; i++) Y[i] = min_func_cb(Pm + 6 * i, &ctx);
    iters = amoeba(min_func_cb, &ctx, Pm, Y, 6, 1e-3);
    if (stats) stats[3] = min_func_cb(Pm, &ctx);
    manip = frame_from_xyz_axis_angle(Pm);
  }
  Frame xc = frame_from_xyzypr(xyzypr);
  xc = frame_mul(frame_mul(frame_mul(xc, frame_inverse(ctx.pivot)), manip), ctx.pivot);
  frame_to_xyzypr(xc, xyzypr);
  pose_store_set(P, xyzypr);
  store_save(P, store);
  if (stats) stats[1] = iters;
  return ctx.evals;
}

// PSO and AMOEBA on closed-form cost functions for pinning (sphere bowl,
// sin(R)/R), and on the exported test objectives below.
static double cost_sinc(const double* x, void*)
{
  double r = sqrt(x[0] * x[0] + x[1] * x[1]);
  return sin(r) / r;
}
static double cost_bowl(const double* x, void* user)
{
  const double* c = (const double*)user;
  double s = 0;
  for (int i = 0; i < 6; i++) s += (x[i] - c[i]) * (x[i] - c[i]);
  return s;
}

// Test objectives, exported so that the reference's own optimisers (built
// into oracle/_ref) can be driven with the very same function.
// user = double[6] centre; a weighted bowl with a ripple, and a 9999 plateau
// outside |x_0 - c_0| > 4 as the tracker returns for out-of-view poses.
ORC_API double orc_cost_ripple(const double* x, void* user)
{
  const double* c = (const double*)user;
  if (std::fabs(x[0] - c[0]) > 4.0) return 9999.0;
  double s = 0;
  for (int i = 0; i < 6; i++) s += (1.0 + 0.5 * i) * (x[i] - c[i]) * (x[i] - c[i]);
  return s + 0.3 * sin(3.0 * x[0]) * cos(2.0 * x[1]);
}
ORC_API double orc_cost_bowl(const double* x, void* user) { return cost_bowl(x, user); }
ORC_API double orc_cost_sinc(const double* x, void* user) { return cost_sinc(x, user); }

ORC_API int orc_pso_fn(cost_fn fn, void* user, unsigned seed, float lo, float hi, unsigned max_epochs,
                       unsigned max_stall, int nparticles, int synchronous, float* best, float* best_ncc)
{
  srand(seed);
  return pso(fn, user, lo, hi, max_epochs, max_stall, nparticles, best, best_ncc, synchronous != 0);
}

ORC_API int orc_amoeba_fn(cost_fn fn, void* user, double* P, double* Y, int ndim, double ftol)
{
  return amoeba(fn, user, P, Y, ndim, ftol);
}

ORC_API int orc_amoeba_sinc(double* P, double* Y, double ftol)
{
  for (int i = 0; i < 3; i++) Y[i] = cost_sinc(P + 2 * i, nullptr);
  return amoeba(cost_sinc, nullptr, P, Y, 2, ftol);
}

ORC_API int orc_pso_bowl(const double* centre, unsigned seed, float lo, float hi, unsigned max_epochs,
                         unsigned max_stall, int nparticles, float* best, float* best_ncc)
{
  srand(seed);
  return pso(cost_bowl, (void*)centre, lo, hi, max_epochs, max_stall, nparticles, best, best_ncc);
}

ORC_API int orc_glibc_rand_first(unsigned seed, int n, int* out)
{
  srand(seed);
  for (int i = 0; i < n; i++) out[i] = rand();
  return RAND_MAX;
}

ORC_API void orc_merge_f32(const float* a, const float* b, float* o, size_t n)
{
  for (size_t i = 0; i < n; i++) o[i] = std::min(a[i] + b[i], 1.0f);
}

// Compositor.cl:2-24: dest = {src1, mask3>=0.5 ? src1/2 + src2/2 : 0, src2}.
ORC_API void orc_composite_f32(const float* s1, const float* s2, const float* s3, float* o, size_t n)
{
  for (size_t i = 0; i < n; i++) {
    float multi = (s3[i] < 0.5f) ? 0.0f : 1.0f;
    o[3 * i + 0] = s1[i];
    o[3 * i + 1] = multi * (s1[i] / 2.f + s2[i] / 2.f);
    o[3 * i + 2] = s2[i];
  }
}

// Direct access to the sampler (normalised coordinates) for pinning tests.
ORC_API double orc_sample3(const uint16_t* vol, const int* vdim, double s, double t, double r, int fp64)
{
  if (fp64) return sample3<double>(vol, vdim[0], vdim[1], vdim[2], s, t, r);
  return sample3<float>(vol, vdim[0], vdim[1], vdim[2], (float)s, (float)t, (float)r);
}
