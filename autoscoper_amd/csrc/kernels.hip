// Per-operator kernels: one launch per reference operator (the "plain"
// form that gpu::View and Tracker::trackFrame drive one call at a time).
#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>

#include "device.hpp"

namespace asc {
namespace {

constexpr int kBlock = 256;

// 16 x 16 pixel block made of four 8 x 8 wave tiles: the 64 rays of a wave
// are spatial neighbours, so their samples hit neighbouring voxels.
__device__ __forceinline__ void tile_pixel(unsigned bx, unsigned by, unsigned& x, unsigned& y)
{
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  x = bx * 16 + (wave & 1) * 8 + (lane & 7);
  y = by * 16 + (wave >> 1) * 8 + (lane >> 3);
}

struct DrrArgs {
  float imv[12];
  float vp[4];
  float step, intensity, cutoff, step_vscale;
  unsigned w, h;
};

__global__ __launch_bounds__(kBlock) void drr_kernel(VolDev v, DrrArgs a, float* __restrict__ out)
{
  __shared__ int s_tab[kBrickTab];
  dev::brick_tables(v, s_tab);
  __syncthreads();
  unsigned x, y;
  tile_pixel(blockIdx.x, blockIdx.y, x, y);
  if (x >= a.w || y >= a.h) return;
  dev::Ray r;
  float val = 0.0f;
  if (dev::ray_setup(a.imv, a.vp, a.step, x, y, a.w, a.h, r)) {
    int n = 0;
    val = dev::drr_finish(dev::march<2>(v, r, a.cutoff, n, s_tab), a.step_vscale, a.intensity);
  }
  out[(size_t)y * a.w + x] = val;
}

struct RadArgs {
  float plane[4];
  float vp[4];
  unsigned w, h;
  float threshold;
  int background;
};

__global__ __launch_bounds__(kBlock) void rad_kernel(ImgDev im, RadArgs a, float* __restrict__ out)
{
  unsigned x, y;
  tile_pixel(blockIdx.x, blockIdx.y, x, y);
  if (x >= a.w || y >= a.h) return;
  out[(size_t)y * a.w + x] = dev::rad_pixel(im, a.plane, a.vp, x, y, a.w, a.h, a.background != 0, a.threshold);
}

struct GlobalGet {
  const float* __restrict__ p;
  int w;
  __device__ __forceinline__ float operator()(int x, int y) const { return p[y * w + x]; }
};

__global__ __launch_bounds__(kBlock) void sobel_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                       int h, float scale, float blend)
{
  unsigned x, y;
  tile_pixel(blockIdx.x, blockIdx.y, x, y);
  if ((int)x >= w || (int)y >= h) return;
  out[y * w + x] = dev::sobel_px(GlobalGet{in, w}, x, y, w, h, scale, blend);
}

__global__ __launch_bounds__(kBlock) void contrast_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                          int w, int h, float alpha, float beta, int size)
{
  unsigned x, y;
  tile_pixel(blockIdx.x, blockIdx.y, x, y);
  if ((int)x >= w || (int)y >= h) return;
  out[y * w + x] = dev::contrast_px(GlobalGet{in, w}, x, y, w, h, alpha, beta, size);
}

constexpr int kArgTaps = 19 * 19;   // radius <= 3: weights travel in the kernel arguments
constexpr int kMaxTaps = 61 * 61;   // radius <= 10: weights from a device buffer

struct ConvArgs {
  int n;
  int sharpen;
  float contrast, threshold;
  const float* wts_dev;  // used when n*n > kArgTaps
  float wts[kArgTaps];
};

__global__ __launch_bounds__(kBlock) void conv_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                      int h, ConvArgs a)
{
  __shared__ float sw[kMaxTaps];
  const int nn = a.n * a.n;
  for (int i = threadIdx.x; i < nn; i += kBlock) sw[i] = nn > kArgTaps ? a.wts_dev[i] : a.wts[i];
  __syncthreads();
  unsigned x, y;
  tile_pixel(blockIdx.x, blockIdx.y, x, y);
  if ((int)x >= w || (int)y >= h) return;
  const float c = dev::conv_px(GlobalGet{in, w}, x, y, w, h, sw, a.n);
  out[y * w + x] = a.sharpen ? dev::sharpen_finish(in[y * w + x], c, a.contrast, a.threshold)
                             : dev::gaussian_finish(c);
}

__global__ __launch_bounds__(kBlock) void binary_kernel(int op, const float* __restrict__ a,
                                                        const float* __restrict__ b, float* __restrict__ o,
                                                        size_t n)
{
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case BIN_MULTIPLY: o[i] = a[i] * b[i]; break;                    // Mult.cl:16
    case BIN_MERGE: o[i] = fminf(a[i] + b[i], 1.0f); break;          // Merger.cl:16
    default: o[i] = (a[i] != 0.0f) ? 1.0f : 0.0f; break;             // DRRBackground.cl:15
  }
}

// Tracker.cpp:747-752 fused: m = bg * (drr != 0); rad = rad * m; drr = drr * m
__global__ __launch_bounds__(kBlock) void mask_multiply_kernel(float* __restrict__ drr, float* __restrict__ rad,
                                                               const float* __restrict__ bg,
                                                               float* __restrict__ mask, size_t n)
{
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float m = (drr[i] != 0.0f) ? 1.0f : 0.0f;
  if (bg) m = bg[i] * m;
  if (mask) mask[i] = m;
  rad[i] = rad[i] * m;
  drr[i] = drr[i] * m;
}

// Compositor.cl:17-23
__global__ __launch_bounds__(kBlock) void composite_kernel(const float* __restrict__ s1,
                                                           const float* __restrict__ s2,
                                                           const float* __restrict__ s3, float* __restrict__ o,
                                                           size_t n)
{
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float multi = (s3[i] < 0.5f) ? 0.0f : 1.0f;
  o[3 * i + 0] = s1[i];
  o[3 * i + 1] = multi * (s1[i] / 2.f + s2[i] / 2.f);
  o[3 * i + 2] = s2[i];
}

// Buffer::copy as a kernel (stream-ordered, capturable in a graph); 16-byte
// vectors when both pointers and the size allow it.
__global__ __launch_bounds__(kBlock) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      size_t n16)
{
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (size_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void copy_bytes_kernel(const uint8_t* __restrict__ src,
                                                            uint8_t* __restrict__ dst, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void fill_kernel(float* __restrict__ o, float v, size_t n)
{
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) o[i] = v;
}

// NCC moments: stage 1 writes one 9-double partial per block, stage 2 sums
// the partials in a fixed order (deterministic) and forms the NCC.
constexpr int kNccPerThread = 8;

__global__ __launch_bounds__(kBlock) void ncc_partial_kernel(const float* __restrict__ f,
                                                             const float* __restrict__ g,
                                                             const float* __restrict__ m, unsigned n,
                                                             double* __restrict__ partials)
{
  __shared__ double lds[4 * 9];
  dev::Moments acc;
  const size_t base = (size_t)blockIdx.x * kBlock * kNccPerThread + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kNccPerThread; k++) {
    const size_t i = base + (size_t)k * kBlock;
    if (i < n) acc.add(f[i], g[i], m[i]);
  }
  double v[9];
  acc.to_double(v);
  dev::block_sum9(v, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 9; k++) partials[(size_t)blockIdx.x * 9 + k] = v[k];
}

__global__ __launch_bounds__(kBlock) void ncc_final_kernel(const double* __restrict__ partials, unsigned nblocks,
                                                           double* __restrict__ result)
{
  __shared__ double lds[4 * 9];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (unsigned b = threadIdx.x; b < nblocks; b += kBlock)
    for (int k = 0; k < 9; k++) v[k] += partials[(size_t)b * 9 + k];
  dev::block_sum9(v, lds);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 9; k++) result[k] = v[k];
    result[9] = ncc_from_moments(v);
  }
}

// Stores the corner texels v[0..per) of one element (x fastest, then y,
// then z: 2x2 (x, y) blocks of one slice each) in the encoding `enc` of
// VolDev::pair_diff. Differences are stored as signed 16-bit integers (the
// host picks an encoding whose differences fit); enc is 0 for F32 storage.
template <typename T>
__device__ __forceinline__ void store_element(T* __restrict__ dst, const T* v, int per, int enc)
{
  for (int s = 0; s < per; s += 4) {
    if (per == 1 || enc == 0) {
      for (int q = s; q < s + 4 && q < per; q++) dst[q] = v[q];
      continue;
    }
    const int t00 = (int)v[s], t10 = (int)v[s + 1], t01 = (int)v[s + 2], t11 = (int)v[s + 3];
    dst[s] = (T)t00;
    dst[s + 1] = (T)(t10 - t00);
    dst[s + 2] = (T)(t01 - t00);
    dst[s + 3] = (T)(enc == 2 ? (t11 - t01) - (t10 - t00) : t11 - t01);
  }
}

// QUAD / OCTET packing: element (i, j, k) of the padded grid receives the
// texels (i+di, j+dj, k+dk), di, dj (, dk) in {0, 1}, in the order
// gather8() expects; texels beyond the padded grid are 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void pack_kernel(const T* __restrict__ src, T* __restrict__ dst, int pW, int pH,
                                                      int pD, int per, int diff)
{
  const size_t n = (size_t)pW * pH * pD;
  for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (size_t)gridDim.x * kBlock) {
    const int i = (int)(e % pW), j = (int)((e / pW) % pH), k = (int)(e / ((size_t)pW * pH));
    T v[8];
    for (int q = 0; q < per; q++) {
      const int di = q & 1, dj = (q >> 1) & 1, dk = (q >> 2) & 1;
      const int ii = i + di, jj = j + dj, kk = k + dk;
      v[q] = T(0);
      if (ii < pW && jj < pH && kk < pD) v[q] = src[((size_t)kk * pH + jj) * pW + ii];
    }
    store_element(dst + e * per, v, per, diff);
  }
}

// A BRICK copy from the padded linear volume (pW = W+2, pH = H+2, pD = D+2,
// voxel x at padded x+1). Every element of the copy is visited once: the
// bricks of an axis of extent 2^l cover cells -2^l .. nb 2^l - 2^l - 1, and
// cell c of the axis lies at bin (c mod 2^l) + bstr (floor(c / 2^l) + 1)
// (the march's offset tables hold the same terms). The element receives
// the 2x2x2 corner texels of its cell (0 outside the padded grid) in the
// order trilinear() reads them.
struct BrickArgs {
  int pW, pH, pD;  // padded linear grid
  int cells[3];    // covered cells per axis (bricks x extent)
  int blog[3], bin[3], bstr[3];
  int esz;         // bytes per cell
};

template <typename T>
__global__ __launch_bounds__(kBlock) void pack_brick_kernel(const T* __restrict__ src, T* __restrict__ dst, BrickArgs a,
                                                            int diff)
{
  const size_t n = (size_t)a.cells[0] * a.cells[1] * a.cells[2];
  for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (size_t)gridDim.x * kBlock) {
    int c[3];
    c[0] = (int)(e % a.cells[0]);
    c[1] = (int)((e / a.cells[0]) % a.cells[1]);
    c[2] = (int)(e / ((size_t)a.cells[0] * a.cells[1]));
    size_t off = 0;
    for (int k = 0; k < 3; k++) {
      c[k] -= 1 << a.blog[k];
      off += (size_t)a.bin[k] * (size_t)(c[k] & ((1 << a.blog[k]) - 1)) +
             (size_t)a.bstr[k] * (size_t)((c[k] >> a.blog[k]) + 1);
    }
    T v[8];
    for (int q = 0; q < 8; q++) {
      // padded coordinates of corner q
      const int ii = c[0] + (q & 1) + 1, jj = c[1] + ((q >> 1) & 1) + 1, kk = c[2] + (q >> 2) + 1;
      v[q] = T(0);
      if (ii >= 0 && jj >= 0 && kk >= 0 && ii < a.pW && jj < a.pH && kk < a.pD)
        v[q] = src[((size_t)kk * a.pH + jj) * a.pW + ii];
    }
    store_element(dst + off / a.esz * 8, v, 8, diff);
  }
}

dim3 grid2(unsigned w, unsigned h) { return dim3((w + 15) / 16, (h + 15) / 16); }

}  // namespace

int launch_drr(asc_context ctx, const VolDev& v, const float* imv, const float* vp, float step, float intensity,
               float cutoff, float* out, unsigned w, unsigned h)
{
  if (w == 0 || h == 0) return ASC_OK;
  DrrArgs a;
  for (int i = 0; i < 12; i++) a.imv[i] = imv[i];
  for (int i = 0; i < 4; i++) a.vp[i] = vp[i];
  a.step = step;
  a.intensity = intensity;
  a.cutoff = cutoff;
  a.step_vscale = step * v.vscale;
  a.w = w;
  a.h = h;
  if (v.storage < ASC_STORAGE_U16 || v.storage > ASC_STORAGE_F32)
    return fail(ASC_ERR_UNSUPPORTED, "drr: unknown volume storage");
  drr_kernel<<<grid2(w, h), kBlock, 0, ctx->stream>>>(v, a, out);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_rad(asc_context ctx, const ImgDev& img, const float* plane, const float* vp, float* out, unsigned w,
               unsigned h, bool background, float threshold)
{
  if (w == 0 || h == 0) return ASC_OK;
  RadArgs a;
  for (int i = 0; i < 4; i++) {
    a.plane[i] = plane[i];
    a.vp[i] = vp[i];
  }
  a.w = w;
  a.h = h;
  a.threshold = threshold;
  a.background = background ? 1 : 0;
  rad_kernel<<<grid2(w, h), kBlock, 0, ctx->stream>>>(img, a, out);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_filter(asc_context ctx, const asc_filter& f, const float* in, float* out, int w, int h)
{
  if (w <= 0 || h <= 0) return ASC_OK;
  const size_t bytes = sizeof(float) * (size_t)w * h;
  if (!f.enabled) return launch_copy(ctx, out, in, bytes);
  switch (f.type) {
    case ASC_FILTER_SOBEL:
      sobel_kernel<<<grid2(w, h), kBlock, 0, ctx->stream>>>(in, out, w, h, f.p[0], f.p[1]);
      break;
    case ASC_FILTER_CONTRAST:
      contrast_kernel<<<grid2(w, h), kBlock, 0, ctx->stream>>>(in, out, w, h, f.p[0], f.p[1], (int)f.p[2]);
      break;
    case ASC_FILTER_GAUSSIAN:
    case ASC_FILTER_SHARPEN: {
      std::vector<float> wts;
      const int n = gauss_weights(f.p[0], wts);
      if (n == 1) return launch_copy(ctx, out, in, bytes);  // GaussianFilter.cpp:131-133
      if (n * n > kMaxTaps) return fail(ASC_ERR_UNSUPPORTED, "filter radius too large (max 10)");
      ConvArgs a;
      a.n = n;
      a.sharpen = f.type == ASC_FILTER_SHARPEN;
      a.contrast = f.p[1] < 1.0f ? 1.0f : f.p[1];  // SharpenFilter::set_contrast clamps to >= 1
      a.threshold = f.p[2];
      a.wts_dev = nullptr;
      if (n * n <= kArgTaps) {
        for (int i = 0; i < n * n; i++) a.wts[i] = wts[i];
      } else {
        ASC_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->conv_wts) ASC_HIP(hipFree(ctx->conv_wts));
        ctx->conv_wts = nullptr;
        ASC_HIP(hipMalloc((void**)&ctx->conv_wts, sizeof(float) * n * n));
        ASC_HIP(hipMemcpy(ctx->conv_wts, wts.data(), sizeof(float) * n * n, hipMemcpyHostToDevice));
        a.wts_dev = ctx->conv_wts;
      }
      conv_kernel<<<grid2(w, h), kBlock, 0, ctx->stream>>>(in, out, w, h, a);
      ASC_HIP(hipGetLastError());
      return ASC_OK;
    }
    default: return fail(ASC_ERR_UNSUPPORTED, "filter type not implemented by the reference (median)");
  }
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_binary(asc_context ctx, int op, const float* a, const float* b, float* out, size_t n)
{
  if (n == 0) return ASC_OK;
  binary_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(op, a, b, out, n);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_mask_multiply(asc_context ctx, float* drr, float* rad, const float* bg, float* mask, size_t n)
{
  if (n == 0) return ASC_OK;
  mask_multiply_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(drr, rad, bg, mask, n);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_composite(asc_context ctx, const float* s1, const float* s2, const float* s3, float* out, size_t n)
{
  if (n == 0) return ASC_OK;
  composite_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(s1, s2, s3, out, n);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_pack_volume(asc_context ctx, const VolDev& v, void* dst, int layout, int pair_diff)
{
  if (pair_diff && (v.storage != ASC_STORAGE_U16 || layout == ASC_LAYOUT_LINEAR))
    return fail(ASC_ERR_INVALID, "pair differences need a u16 volume in a paired layout");
  const int per = layout == ASC_LAYOUT_QUAD ? 4 : 8;
  const int pH = v.pS / v.pW, pD = v.D + 2;
  const unsigned grid = 256 * 8;
  if (layout == ASC_LAYOUT_BRICK) return fail(ASC_ERR_INVALID, "BRICK copies are packed by launch_pack_brick");
  if (v.storage == ASC_STORAGE_F32)
    pack_kernel<uint32_t><<<grid, kBlock, 0, ctx->stream>>>((const uint32_t*)v.data, (uint32_t*)dst, v.pW, pH, pD,
                                                            per, 0);
  else
    pack_kernel<uint16_t><<<grid, kBlock, 0, ctx->stream>>>((const uint16_t*)v.data, (uint16_t*)dst, v.pW, pH, pD,
                                                            per, pair_diff);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_pack_brick(asc_context ctx, const VolDev& v, const VolDev& b, void* dst, int pair_diff)
{
  if (pair_diff && v.storage != ASC_STORAGE_U16)
    return fail(ASC_ERR_INVALID, "texel differences need a u16 volume");
  BrickArgs a;
  a.pW = v.pW;
  a.pH = v.pS / v.pW;
  a.pD = v.D + 2;
  const int n[3] = {v.W, v.H, v.D};
  for (int k = 0; k < 3; k++) {
    a.blog[k] = b.blog[k];
    a.bin[k] = b.bin[k];
    a.bstr[k] = b.bstr[k];
    a.cells[k] = (((n[k] - 1) >> b.blog[k]) + 2) << b.blog[k];
  }
  a.esz = b.esz;
  const unsigned grid = 256 * 8;
  if (v.storage == ASC_STORAGE_F32)
    pack_brick_kernel<uint32_t><<<grid, kBlock, 0, ctx->stream>>>((const uint32_t*)v.data, (uint32_t*)dst, a, 0);
  else
    pack_brick_kernel<uint16_t><<<grid, kBlock, 0, ctx->stream>>>((const uint16_t*)v.data, (uint16_t*)dst, a,
                                                                  pair_diff);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_copy(asc_context ctx, void* dst, const void* src, size_t bytes)
{
  if (bytes == 0) return ASC_OK;
  const bool vec = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0) && (bytes % 16 == 0);
  const size_t n = vec ? bytes / 16 : bytes;
  const unsigned grid = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 256 * 16);
  if (vec)
    copy_kernel<<<grid, kBlock, 0, ctx->stream>>>((const uint4*)src, (uint4*)dst, n);
  else
    copy_bytes_kernel<<<grid, kBlock, 0, ctx->stream>>>((const uint8_t*)src, (uint8_t*)dst, n);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int launch_fill_f32(asc_context ctx, float* dst, float v, size_t n)
{
  if (n == 0) return ASC_OK;
  fill_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(dst, v, n);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

int ncc_moments(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, double* out9,
                double* ncc)
{
  const unsigned per_block = kBlock * kNccPerThread;
  const unsigned nblocks = n == 0 ? 1 : (n + per_block - 1) / per_block;
  if (ctx->ncc_partials_cap < nblocks) {
    if (ctx->ncc_partials) ASC_HIP(hipFree(ctx->ncc_partials));
    ctx->ncc_partials = nullptr;
    ASC_HIP(hipMalloc((void**)&ctx->ncc_partials, sizeof(double) * 9 * nblocks));
    ctx->ncc_partials_cap = nblocks;
  }
  if (!ctx->ncc_result) ASC_HIP(hipMalloc((void**)&ctx->ncc_result, sizeof(double) * 10));
  if (!ctx->ncc_host) ASC_HIP(hipHostMalloc((void**)&ctx->ncc_host, sizeof(double) * 10, hipHostMallocDefault));
  ncc_partial_kernel<<<nblocks, kBlock, 0, ctx->stream>>>(f, g, m, n, ctx->ncc_partials);
  ASC_HIP(hipGetLastError());
  ncc_final_kernel<<<1, kBlock, 0, ctx->stream>>>(ctx->ncc_partials, nblocks, ctx->ncc_result);
  ASC_HIP(hipGetLastError());
  ASC_HIP(hipMemcpyAsync(ctx->ncc_host, ctx->ncc_result, sizeof(double) * 10, hipMemcpyDeviceToHost, ctx->stream));
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  if (out9)
    for (int k = 0; k < 9; k++) out9[k] = ctx->ncc_host[k];
  if (ncc) *ncc = ctx->ncc_host[9];
  return ASC_OK;
}

// ---- the moments NCC without allocation (asc_ncc_noalloc) -------------------
// Its block partials live in `scratch`, a static device array: every GPU the
// library runs on holds its own copy (ASC_NCC_NOALLOC_MAX_N elements need
// kNoAllocBlocks partials of 9 doubles, 316 KB) until the process ends. Host
// code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): noalloc_scratch_init takes the address of
// the current device's copy once, and the launches use that address.
constexpr unsigned kNoAllocBlocks = (ASC_NCC_NOALLOC_MAX_N + kBlock * kNccPerThread - 1) / (kBlock * kNccPerThread);
static __device__ __align__(256) double scratch[9 * kNoAllocBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device)
{
  ASC_CHECK(device >= 0 && device < kMaxDevices, ASC_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return ASC_OK;
  void* p = nullptr;
  ASC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double*>(p);
  return ASC_OK;
}

int ncc_noalloc(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, double* out10)
{
  ASC_CHECK(n <= ASC_NCC_NOALLOC_MAX_N, ASC_ERR_TOO_LARGE,
            "asc_ncc_noalloc: n exceeds ASC_NCC_NOALLOC_MAX_N (9000000 elements)");
  double* part;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    part = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  }
  ASC_CHECK(part, ASC_ERR_INVALID, "asc_ncc_noalloc: scratch address not set up for this device");
  const unsigned per_block = kBlock * kNccPerThread;
  const unsigned nblocks = n == 0 ? 1 : (n + per_block - 1) / per_block;
  // every partial the final kernel reads is written by the first: nothing
  // left by an earlier call is read, so nothing needs clearing
  ncc_partial_kernel<<<nblocks, kBlock, 0, ctx->stream>>>(f, g, m, n, part);
  ASC_HIP(hipGetLastError());
  ncc_final_kernel<<<1, kBlock, 0, ctx->stream>>>(part, nblocks, out10);
  ASC_HIP(hipGetLastError());
  return ASC_OK;
}

// ---- the reference's float NCC, launch by launch (gpu::ncc, Ncc.cpp:180-220)

// One pass of ncc_sum (NccSum.cl:2-31): block b of T threads loads element
// i (CUDA indexing b*T + t, or the OpenCL kernel's t + b*B as written) or 0,
// then halves `for s = T/2; s > 0; s >>= 1` in LDS (an odd width drops its
// last element, as the reference's tree does) and writes its sum.
__global__ __launch_bounds__(256) void tree_sum_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                      unsigned n, unsigned nblocks, int opencl)
{
  __shared__ float buf[256];
  const unsigned t = threadIdx.x, b = blockIdx.x, T = blockDim.x;
  const unsigned i = opencl ? t + b * nblocks : b * T + t;
  buf[t] = i < n ? in[i] : 0.0f;
  __syncthreads();
  for (unsigned s = T / 2; s > 0; s >>= 1) {
    if (t < s) buf[t] += buf[t + s];
    __syncthreads();
  }
  if (t == 0) out[b] = buf[0];
}

// ncc_kernel (Ncc.cl:2-31): per-pixel products about the float means.
__global__ __launch_bounds__(128) void ncc_products_kernel(const float* __restrict__ f, float mf,
                                                          const float* __restrict__ g, float mg,
                                                          const float* __restrict__ m, float* __restrict__ nums,
                                                          float* __restrict__ d1s, float* __restrict__ d2s,
                                                          unsigned n)
{
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (m[i] > 0.5f) {
    const float a = f[i] - mf, c = g[i] - mg;
    nums[i] = a * c;
    d1s[i] = a * a;
    d2s[i] = c * c;
  } else {
    nums[i] = d1s[i] = d2s[i] = 0.0f;
  }
}

namespace {

// A blocking context waits after every launch (the reference's clFinish).
int after_launch(asc_context ctx)
{
  ASC_HIP(hipGetLastError());
  if (ctx->blocking) ASC_HIP(hipStreamSynchronize(ctx->stream));
  return ASC_OK;
}

// ncc_sum (Ncc.cpp:103-146): passes of at most maxT-thread blocks until one
// value is left, then a blocking read of it. Each pass writes a buffer of
// its own (the reference's OpenCL passes run in place, a race).
int tree_sum(asc_context ctx, const float* in, unsigned n, bool opencl, float* result)
{
  const unsigned maxT = opencl ? 128u : 256u;
  if (n == 0) {
    *result = 0.0f;
    return ASC_OK;
  }
  const float* cur = in;
  float* bufs[2] = {ctx->tree_buf + ctx->tree_cap - 2 * ctx->tree_part, ctx->tree_buf + ctx->tree_cap - ctx->tree_part};
  int which = 0;
  int rc;
  while (n > 1) {
    const unsigned T = n < maxT ? n : maxT;
    const unsigned B = (n + T - 1) / T;
    tree_sum_kernel<<<B, T, 0, ctx->stream>>>(cur, bufs[which], n, B, opencl ? 1 : 0);
    if ((rc = after_launch(ctx))) return rc;
    cur = bufs[which];
    which ^= 1;
    n = B;
  }
  float* host = reinterpret_cast<float*>(ctx->ncc_host);
  ASC_HIP(hipMemcpyAsync(host, cur, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  *result = host[0];
  return ASC_OK;
}

}  // namespace

int ncc_tree(asc_context ctx, const float* f, const float* g, const float* m, unsigned n, bool opencl,
             float* sums6, float* ncc)
{
  // scratch: nums, den1s, den2s (n each) and two pass buffers
  const size_t part = n / 128 + 2;
  const size_t need = 3 * (size_t)n + 2 * part;
  if (ctx->tree_cap < need) {
    if (ctx->tree_buf) ASC_HIP(hipFree(ctx->tree_buf));
    ctx->tree_buf = nullptr;
    ctx->tree_cap = 0;
    ASC_HIP(hipMalloc((void**)&ctx->tree_buf, sizeof(float) * need));
    ctx->tree_cap = need;
  }
  ctx->tree_part = part;
  if (!ctx->ncc_host) ASC_HIP(hipHostMalloc((void**)&ctx->ncc_host, sizeof(double) * 10, hipHostMallocDefault));
  float* nums = ctx->tree_buf;
  float* d1s = nums + n;
  float* d2s = d1s + n;
  int rc;
  float nb, sf, sg, s1, s2, sn;
  if ((rc = tree_sum(ctx, m, n, opencl, &nb))) return rc;
  if ((rc = tree_sum(ctx, f, n, opencl, &sf))) return rc;
  if ((rc = tree_sum(ctx, g, n, opencl, &sg))) return rc;
  const float mf = sf / nb, mg = sg / nb;
  if (n > 0) {
    ncc_products_kernel<<<(n + 127) / 128, 128, 0, ctx->stream>>>(f, mf, g, mg, m, nums, d1s, d2s, n);
    if ((rc = after_launch(ctx))) return rc;
  }
  if ((rc = tree_sum(ctx, d1s, n, opencl, &s1))) return rc;
  if ((rc = tree_sum(ctx, d2s, n, opencl, &s2))) return rc;
  const float den = std::sqrt(s1 * s2);
  sn = std::numeric_limits<float>::quiet_NaN();
  if (den < 1e-5) {  // (a double comparison, as written)
    *ncc = 1e5f;
  } else {
    if ((rc = tree_sum(ctx, nums, n, opencl, &sn))) return rc;
    *ncc = sn / den;
  }
  if (sums6) {
    sums6[0] = nb, sums6[1] = sf, sums6[2] = sg, sums6[3] = sn, sums6[4] = s1, sums6[5] = s2;
  }
  return ASC_OK;
}

}  // namespace asc
