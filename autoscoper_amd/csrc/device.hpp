// Device helpers shared by the per-operator kernels and the fused batch
// kernels. Only included by translation units built with -ffp-contract=off:
// plain `a * b + c` below is two IEEE operations, as in the reference
// kernels; contraction is written explicitly (fmaf) where it is allowed.
#pragma once

#include "internal.hpp"

namespace asc {
namespace dev {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Ray setup, RayCaster.cl:20-57. Exact operation order: the slab distances
// decide the sample count and hence the DRR's zero pattern (the mask).
// ---------------------------------------------------------------------------
struct Ray {
  float ox, oy, oz;
  float dx, dy, dz;
  float t_near, t_far;
};

__device__ __forceinline__ bool ray_setup(const float* __restrict__ imv, const float* __restrict__ vp, float step,
                                          unsigned x, unsigned y, unsigned w, unsigned h, Ray& r)
{
  const float u = vp[0] + vp[2] * ((float)x / (float)w);
  const float v = vp[1] + vp[3] * ((float)y / (float)h);
  // normalize((u, v, -2)) as p / sqrt(dot(p, p))
  const float len = sqrtf(u * u + v * v + 4.0f);
  const float lx = step * (u / len);
  const float ly = step * (v / len);
  const float lz = step * (-2.0f / len);
  r.ox = imv[3];
  r.oy = imv[7];
  r.oz = imv[11];
  r.dx = imv[0] * lx + imv[1] * ly + imv[2] * lz;
  r.dy = imv[4] * lx + imv[5] * ly + imv[6] * lz;
  r.dz = imv[8] * lx + imv[9] * ly + imv[10] * lz;
  // box [0,1] x [0,1] x [-1,0]
  const float bx = (0.0f - r.ox) / r.dx, by = (0.0f - r.oy) / r.dy, bz = (-1.0f - r.oz) / r.dz;
  const float tx = (1.0f - r.ox) / r.dx, ty = (1.0f - r.oy) / r.dy, tz = (0.0f - r.oz) / r.dz;
  float t_near = fmaxf(fmaxf(fminf(tx, bx), fminf(ty, by)), fminf(tz, bz));
  const float t_far = fminf(fminf(fmaxf(tx, bx), fmaxf(ty, by)), fmaxf(tz, bz));
  if (!(t_far > t_near)) return false;
  if (t_near < 0.0f) t_near = 0.0f;
  r.t_near = t_near;
  r.t_far = t_far;
  return true;
}

// ---------------------------------------------------------------------------
// Volume texel loads (raw value; normalised = raw * VolDev::vscale).
// ---------------------------------------------------------------------------
// Volume reads go through buffer loads: the descriptor is scalar, the two
// y/z strides ride in scalar offsets, and an offset outside the allocation
// reads 0 instead of faulting.
struct VolRsrc {
  __amdgpu_buffer_rsrc_t rsrc;
  int sy, sz;  // byte strides of +1 row and +1 slice (< 2^23, checked on the host)
  int c0;      // byte offset of padded element (1, 1, 1): voxel (0, 0, 0)
  const int* tab;  // BRICK: the block's LDS offset tables (brick_tables)
  int ty, tz;      // indices of the y and z tables in `tab` (W + 1, W + H + 2)
};

// log2 of the bytes per element of layout LY over storage ST: one texel
// (LINEAR), a 2x2 block (QUAD) or a 2x2x2 cell (OCTET, BRICK).
template <int ST, int LY>
constexpr int elem_shift()
{
  return (ST == ASC_STORAGE_F32 ? 2 : 1) + (LY == ASC_LAYOUT_LINEAR ? 0 : LY == ASC_LAYOUT_QUAD ? 2 : 3);
}

// floor(x) as an int in one instruction (finite x within int range).
__device__ __forceinline__ int cvt_flr(float x)
{
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// a * b + c with a, b signed 24-bit (b wave-uniform) and c 32-bit.
__device__ __forceinline__ int mad_i24(int a, int b, int c)
{
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

// The volume of a ray is uniform across the wave (one volume per tile);
// readfirstlane states that, so the descriptor and strides are scalar.
__device__ __forceinline__ VolRsrc vol_rsrc(const VolDev& v, const int* tab)
{
  const uint64_t base = (uint64_t)(uintptr_t)v.data;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
  void* p = (void*)(uintptr_t)(((uint64_t)hi << 32) | lo);
  const int bytes = __builtin_amdgcn_readfirstlane(v.nbytes);
  VolRsrc r;
  // dword 3 = 0x00020000: DATA_FORMAT_32, as raw (untyped) buffer loads need
  r.rsrc = __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, 0x00020000);
  r.sy = __builtin_amdgcn_readfirstlane(v.pW * v.esz);
  r.sz = __builtin_amdgcn_readfirstlane(v.pS * v.esz);
  // byte offset of the element of voxel (0, 0, 0): padded element (1, 1, 1)
  // (BRICK copies use their offset tables instead)
  r.c0 = r.sy + r.sz + __builtin_amdgcn_readfirstlane(v.esz);
  r.tab = tab;
  r.ty = __builtin_amdgcn_readfirstlane(v.W + 1);
  r.tz = __builtin_amdgcn_readfirstlane(v.W + v.H + 2);
  return r;
}

// The BRICK copy's offset terms for cells i = -1 .. W-1, j = -1 .. H-1 and
// k = -1 .. D-1 (every cell an in-box sample can touch), filled by the
// whole block into `tab` (kBrickTab ints of LDS): cell c of axis a lies at
// bin[a] (c mod 2^blog[a]) + bstr[a] (floor(c / 2^blog[a]) + 1) bytes of
// that axis, and a cell's offset is the sum of its three terms, tab[i + 1]
// + tab[W + 2 + j] + tab[W + H + 3 + k] (cell_offset; the pack kernel
// places cells by the same terms). Other layouts need no table. The caller
// synchronises the block before the tables are read.
__device__ __forceinline__ void brick_tables(const VolDev& v, int* tab)
{
  if (v.layout != ASC_LAYOUT_BRICK) return;
  const int n[3] = {v.W, v.H, v.D};
  int base = 0;
  for (int a = 0; a < 3; a++) {
    const int l = v.blog[a], m = (1 << l) - 1, bi = v.bin[a], bs = v.bstr[a];
    for (int e = (int)threadIdx.x; e <= n[a]; e += (int)blockDim.x) {
      const int c = e - 1;
      tab[base + e] = bi * (c & m) + bs * ((c >> l) + 1);
    }
    base += n[a] + 1;
  }
}

// Internal storage codes of u16 volumes in the encodings 1 and 2 of
// VolDev::pair_diff (texel differences); never seen by the C ABI.
constexpr int kStorageU16Diff = 16;
constexpr int kStorageU16Grad = 17;

// Two 16-bit texels packed in a dword -> raw floats.
template <int ST>
__device__ __forceinline__ void unpack2(uint32_t p, float& lo, float& hi)
{
  if constexpr (ST == ASC_STORAGE_U16) {
    lo = (float)(p & 0xFFFFu);
    hi = (float)(p >> 16);
  } else if constexpr (ST == ASC_STORAGE_F16) {
    lo = __half2float(__ushort_as_half((unsigned short)(p & 0xFFFFu)));
    hi = __half2float(__ushort_as_half((unsigned short)(p >> 16)));
  } else {
    lo = __uint_as_float(p << 16);
    hi = __uint_as_float(p & 0xFFFF0000u);
  }
}

// 16-bit storages: the 4 dwords holding the x-pairs {t000|t100, t010|t110,
// t001|t101, t011|t111} (low half = x0, high half = x1) of element voff.
template <int LY>
__device__ __forceinline__ void gather_pairs16(const VolRsrc& r, int voff, uint32_t* q)
{
  if constexpr (LY == ASC_LAYOUT_LINEAR) {
    const int so[4] = {0, r.sy, r.sz, r.sz + r.sy};
#pragma unroll
    for (int e = 0; e < 4; e++) q[e] = __builtin_amdgcn_raw_buffer_load_b32(r.rsrc, voff, so[e], 0);
  } else if constexpr (LY == ASC_LAYOUT_QUAD) {
    const auto p0 = __builtin_amdgcn_raw_buffer_load_b64(r.rsrc, voff, 0, 0);
    const auto p1 = __builtin_amdgcn_raw_buffer_load_b64(r.rsrc, voff, r.sz, 0);
    q[0] = p0[0];
    q[1] = p0[1];
    q[2] = p1[0];
    q[3] = p1[1];
  } else {
    const auto p = __builtin_amdgcn_raw_buffer_load_b128(r.rsrc, voff, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; e++) q[e] = p[e];
  }
}

__device__ __forceinline__ float f16_lo(uint32_t p) { return __half2float(__ushort_as_half((unsigned short)(p & 0xFFFFu))); }
__device__ __forceinline__ float f16_hi(uint32_t p) { return __half2float(__ushort_as_half((unsigned short)(p >> 16))); }

// The 8 corner texels {t000, t100, t010, t110, t001, t101, t011, t111} of
// element voff (byte offset of corner (i0, j0, k0)) in layout LY.
template <int ST, int LY>
__device__ __forceinline__ void gather8(const VolRsrc& r, int voff, float* t)
{
  constexpr bool F32 = ST == ASC_STORAGE_F32;
  if constexpr (LY == ASC_LAYOUT_LINEAR) {
    // x-pairs: 4 loads (y, z), (y+1, z), (y, z+1), (y+1, z+1)
    const int so[4] = {0, r.sy, r.sz, r.sz + r.sy};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if constexpr (F32) {
        const auto p = __builtin_amdgcn_raw_buffer_load_b64(r.rsrc, voff, so[q], 0);
        t[2 * q] = __uint_as_float(p[0]);
        t[2 * q + 1] = __uint_as_float(p[1]);
      } else {
        unpack2<ST>(__builtin_amdgcn_raw_buffer_load_b32(r.rsrc, voff, so[q], 0), t[2 * q], t[2 * q + 1]);
      }
    }
  } else if constexpr (LY == ASC_LAYOUT_QUAD) {
    // one 2x2 block per slice: slices k and k+1
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int so = q ? r.sz : 0;
      if constexpr (F32) {
        const auto p = __builtin_amdgcn_raw_buffer_load_b128(r.rsrc, voff, so, 0);
#pragma unroll
        for (int e = 0; e < 4; e++) t[4 * q + e] = __uint_as_float(p[e]);
      } else {
        const auto p = __builtin_amdgcn_raw_buffer_load_b64(r.rsrc, voff, so, 0);
        unpack2<ST>(p[0], t[4 * q], t[4 * q + 1]);
        unpack2<ST>(p[1], t[4 * q + 2], t[4 * q + 3]);
      }
    }
  } else {
    if constexpr (F32) {
      const auto p = __builtin_amdgcn_raw_buffer_load_b128(r.rsrc, voff, 0, 0);
      const auto q = __builtin_amdgcn_raw_buffer_load_b128(r.rsrc, voff, 16, 0);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        t[e] = __uint_as_float(p[e]);
        t[4 + e] = __uint_as_float(q[e]);
      }
    } else {
      const auto p = __builtin_amdgcn_raw_buffer_load_b128(r.rsrc, voff, 0, 0);
#pragma unroll
      for (int e = 0; e < 4; e++) unpack2<ST>(p[e], t[2 * e], t[2 * e + 1]);
    }
  }
}

// Byte offset of the element holding cell (i, j, k) (corner (i, j, k) of
// the trilinear footprint) in layout LY. In-box samples have
// -1 <= i <= W-1 (likewise j, k), so the padded element (i+1, j+1, k+1) is
// in the array and its offset i*esz + j*sy + k*sz + c0 is >= 0; the +1s are
// folded into c0.
template <int ST, int LY>
__device__ __forceinline__ int cell_offset(const VolRsrc& r, int i, int j, int k)
{
  if constexpr (LY == ASC_LAYOUT_BRICK) {
    // the copy's per-axis offset terms from the block's LDS tables
    // (brick_tables); in-box samples have -1 <= i <= W-1 (likewise j, k)
    return r.tab[i + 1] + r.tab[r.ty + 1 + j] + r.tab[r.tz + 1 + k];
  } else {
    return mad_i24(k, r.sz, mad_i24(j, r.sy, (i << elem_shift<ST, LY>()) + r.c0));
  }
}

// Trilinear sample at voxel-space position (X, Y, Z) = (s*W-0.5, t*H-0.5,
// r*D-0.5) with the sampler's weights (OpenCL 1.2 section 8.2): i0 =
// floor(X), a = X - floor(X). Each lerp is one FMA: fma(a, T1 - T0, T0) is
// exact when T1 == 0 or T0 == 0, so the result is > 0 exactly when a texel
// with positive weight is > 0 (the zero pattern of the weighted-sum form).
template <int ST, int LY>
__device__ __forceinline__ float trilinear(const VolDev& v, const VolRsrc& r, float X, float Y, float Z)
{
  const float a = __builtin_amdgcn_fractf(X);
  const float b = __builtin_amdgcn_fractf(Y);
  const float c = __builtin_amdgcn_fractf(Z);
  // in-box samples have -1 <= i0 <= W-1, so the padded element
  // (i0+1, j0+1, k0+1) is in the array and its byte offset
  // i0*esz + j0*sy + k0*sz + c0 is >= 0; the +1s are folded into c0
  const int voff = cell_offset<ST, LY>(r, cvt_flr(X), cvt_flr(Y), cvt_flr(Z));
  if constexpr (ST == kStorageU16Grad || ST == kStorageU16Diff || ST == ASC_STORAGE_U16) {
    // u16 volumes (integer texels). Per slice, with the slice's 2 x 2
    // block {t00, t10, t01, t11}: the x-lerp of row 0 is
    // r0 = fma(a, t10 - t00, t00), and the y-lerp's difference of the two
    // x-lerps, (t01 + a (t11 - t01)) - (t00 + a (t10 - t00)), is the exact
    // dy + a dxy with the integers dy = t01 - t00 and
    // dxy = (t11 - t01) - (t10 - t00): one FMA rounds it once. So the
    // slice's value is fma(b, fma(a, dxy, dy), r0). Since that difference
    // is >= -r0 before rounding, it is >= -r0 after (rounding is monotonic
    // and odd), and the slice value is >= 0; it is 0 exactly when every
    // texel of the slice with a positive weight is 0. The z-lerp is
    // fma(c, v1 - v0, v0) on the two slice values, which keeps both
    // properties. Every u16 encoding and layout computes these same
    // integers (differences of integers below 2^24 are exact in float), so
    // all give the same bits.
    float x0[2], dx[2], dy[2], dxy[2];  // per slice
    if constexpr (ST == kStorageU16Grad) {
      // {t00 | t10 - t00, t01 - t00 | dxy} per slice
      uint32_t q[4];
      gather_pairs16<LY>(r, voff, q);
#pragma unroll
      for (int sl = 0; sl < 2; sl++) {
        x0[sl] = (float)(q[2 * sl] & 0xFFFFu);
        dx[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl] >> 16);
        dy[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl + 1] & 0xFFFFu);
        dxy[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl + 1] >> 16);
      }
    } else if constexpr (ST == kStorageU16Diff) {
      // {t00 | t10 - t00, t01 - t00 | t11 - t01} per slice
      uint32_t q[4];
      gather_pairs16<LY>(r, voff, q);
#pragma unroll
      for (int sl = 0; sl < 2; sl++) {
        x0[sl] = (float)(q[2 * sl] & 0xFFFFu);
        dx[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl] >> 16);
        dy[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl + 1] & 0xFFFFu);
        dxy[sl] = (float)(int)(int16_t)(uint16_t)(q[2 * sl + 1] >> 16) - dx[sl];
      }
    } else {
      float t[8];
      gather8<ST, LY>(r, voff, t);
#pragma unroll
      for (int sl = 0; sl < 2; sl++) {
        const float* u = t + 4 * sl;
        x0[sl] = u[0];
        dx[sl] = u[1] - u[0];
        dy[sl] = u[2] - u[0];
        dxy[sl] = (u[3] - u[2]) - dx[sl];
      }
    }
    const float v0 = fmaf(b, fmaf(a, dxy[0], dy[0]), fmaf(a, dx[0], x0[0]));
    const float v1 = fmaf(b, fmaf(a, dxy[1], dy[1]), fmaf(a, dx[1], x0[1]));
    return fmaf(c, v1 - v0, v0);
  }
  if constexpr (ST == ASC_STORAGE_F16) {
    // F16 volumes: the x-lerps read the half-precision texels directly in
    // FP32 mixed-precision FMAs (v_fma_mix_f32), as (1-a)*T0 + a*T1 with one
    // rounding per FMA; y and z lerps as below. Zero weights and zero
    // texels still give exact zeros.
    uint32_t q[4];
    gather_pairs16<LY>(r, voff, q);
    const float oma = 1.0f - a;
    const float c00 = fmaf(a, f16_hi(q[0]), fmaf(oma, f16_lo(q[0]), 0.0f));
    const float c10 = fmaf(a, f16_hi(q[1]), fmaf(oma, f16_lo(q[1]), 0.0f));
    const float c01 = fmaf(a, f16_hi(q[2]), fmaf(oma, f16_lo(q[2]), 0.0f));
    const float c11 = fmaf(a, f16_hi(q[3]), fmaf(oma, f16_lo(q[3]), 0.0f));
    const float c0 = fmaf(b, c10 - c00, c00);
    const float c1 = fmaf(b, c11 - c01, c01);
    return fmaf(c, c1 - c0, c0);
  }
  float t[8];
  gather8<ST, LY>(r, voff, t);
  const float c00 = fmaf(a, t[1] - t[0], t[0]);
  const float c10 = fmaf(a, t[3] - t[2], t[2]);
  const float c01 = fmaf(a, t[5] - t[4], t[4]);
  const float c11 = fmaf(a, t[7] - t[6], t[6]);
  const float c0 = fmaf(b, c10 - c00, c00);
  const float c1 = fmaf(b, c11 - c01, c01);
  return fmaf(c, c1 - c0, c0);
}

// Ray march, RayCaster.cl:59-69: t runs from t_far down by 1 while t >
// t_near (the same float loop, so the same sample count). Returns the sum
// of raw texel values above the raw cutoff.
//
// Exact volumes (u16, F32) form the sample position as the reference's
// expressions compile in OpenCL C, where FP_CONTRACT is on by default and
// clang fuses `ray.origin + t*ray.direction` and the sampler's `s*W - 0.5`
// into one FMA each: p = fma(t, d, o), X = fma(p, W, -0.5). The origin lies
// far outside the unit box, so o + t*d cancels, and a different association
// (the entry-point form below) moves near-zero silhouette pixels by more
// than 1e-4 relative and can flip an exact zero of the DRR (and so a mask
// count).
// Half-precision volumes (F16, BF16), whose storage rounding already
// dominates their error against the all-double DRR, form positions from the
// ray's entry point in voxel space when there is no cutoff:
// X_k = A - k*B, one FMA per axis, where A = X(t_far) and B = dX/dt are
// computed in double from the same float o, d and t_far and rounded once.
// That position error (<= half an ulp of A plus k half-ulps of B) is below
// the reference's own rounding of o + t*d: pixels >= 1e-5 move by <= 2.5e-5
// relative on small scenes and <= 6.5e-5 on a 256x256x320 volume at 0.3 mm
// (p99.9 2.3e-5), against an F16 storage rounding of 2^-11.
// With a cutoff (CUT = true) every storage keeps the reference order: a
// sample whose value lies near the cutoff would otherwise be kept by one
// and dropped by the other.
// CUT = false is the cutoff <= 0 case: trilinear values are >= 0, so
// `val > cutoff ? val : 0` adds val unchanged.
// Samples visited by `for (t = far; t > near; t -= 1)`: t_k = far - k is
// exact in float for every visited k (far < 2^24, near >= 0), so the count
// is #{k >= 0 : far - k > near} = ceil(far - near) in exact arithmetic.
__device__ __forceinline__ int sample_count(const Ray& r)
{
  // (the cap only matters for degenerate geometry, where the reference's
  // own loop would not terminate)
  return (int)fmin(ceil((double)r.t_far - (double)r.t_near), 4194304.0);
}

template <int ST, int LY, bool CUT>
__device__ __forceinline__ void accumulate(const VolDev& v, const VolRsrc& rs, const Ray& r, float t, float W,
                                           float H, float D, float cutoff, float vscale, float& acc)
{
  const float px = fmaf(t, r.dx, r.ox);
  const float py = fmaf(t, r.dy, r.oy);
  const float pz = fmaf(t, r.dz, r.oz);
  // sampler coordinates (p.x, 1 - p.y, -p.z) in voxel space
  const float X = fmaf(px, W, -0.5f);
  const float Y = fmaf(1.0f - py, H, -0.5f);
  const float Z = fmaf(-pz, D, -0.5f);
  const float val = trilinear<ST, LY>(v, rs, X, Y, Z);
  // the reference's `val > cutoff` on the normalised value (raw * vscale),
  // while the sum stays raw: step * vscale is applied once per ray
  if constexpr (CUT)
    acc += val * vscale > cutoff ? val : 0.0f;
  else
    acc += val;
}

// The voxel-space line of a ray from its entry point: sample k (t = t_far -
// k) is at (fma(k, nbx, ax), fma(k, nby, ay), fma(k, nbz, az)).
struct VoxelLine {
  float ax, ay, az;     // X, Y, Z at t = t_far
  float nbx, nby, nbz;  // -dX/dt, -dY/dt, -dZ/dt
};

__device__ __forceinline__ VoxelLine voxel_line(const Ray& r, float W, float H, float D)
{
  // sampler coordinates (p.x, 1 - p.y, -p.z) in voxel space, p = o + t*d
  const double tf = r.t_far;
  VoxelLine l;
  l.ax = (float)((double)W * ((double)r.ox + tf * (double)r.dx) - 0.5);
  l.ay = (float)((double)H * (1.0 - ((double)r.oy + tf * (double)r.dy)) - 0.5);
  l.az = (float)(-(double)D * ((double)r.oz + tf * (double)r.dz) - 0.5);
  l.nbx = (float)(-(double)W * (double)r.dx);
  l.nby = (float)((double)H * (double)r.dy);
  l.nbz = (float)((double)D * (double)r.dz);
  return l;
}

template <int ST, int LY>
__device__ __forceinline__ void accumulate_line(const VolDev& v, const VolRsrc& rs, const VoxelLine& l, float k,
                                                float& acc)
{
  acc += trilinear<ST, LY>(v, rs, fmaf(k, l.nbx, l.ax), fmaf(k, l.nby, l.ay), fmaf(k, l.nbz, l.az));
}

// UNROLL samples per iteration (t, t-1, ...): their loads are independent,
// so a wave keeps 4*UNROLL loads in flight; the sum keeps the reference's
// sample order.
template <int ST, int LY, bool CUT, int UNROLL>
__device__ __forceinline__ float march_loop(const VolDev& v, const Ray& r, float cutoff, int& nsamples,
                                            const int* tab)
{
  const VolRsrc rs = vol_rsrc(v, tab);
  const float vscale = v.vscale;
  const float W = (float)v.W, H = (float)v.H, D = (float)v.D;
  const int n = sample_count(r);
  float acc = 0.0f;
  int k = 0;
  if constexpr (CUT || !(ST == ASC_STORAGE_F16 || ST == ASC_STORAGE_BF16)) {
    float t = r.t_far;
    for (; k + UNROLL <= n; k += UNROLL) {
#pragma unroll
      for (int u = 0; u < UNROLL; u++) accumulate<ST, LY, CUT>(v, rs, r, t - (float)u, W, H, D, cutoff, vscale, acc);
      t -= (float)UNROLL;
    }
    for (; k < n; k++) {
      accumulate<ST, LY, CUT>(v, rs, r, t, W, H, D, cutoff, vscale, acc);
      t -= 1.0f;
    }
  } else {
    const VoxelLine l = voxel_line(r, W, H, D);
    float kf = 0.0f;  // == k, exact (n <= 2^22)
    for (; k + UNROLL <= n; k += UNROLL) {
#pragma unroll
      for (int u = 0; u < UNROLL; u++) accumulate_line<ST, LY>(v, rs, l, kf + (float)u, acc);
      kf += (float)UNROLL;
    }
    for (; k < n; k++) {
      accumulate_line<ST, LY>(v, rs, l, kf, acc);
      kf += 1.0f;
    }
  }
  nsamples += n;
  return acc;
}

// cutoff is on normalised values, as the reference compares them
// Gather probe: the march's memory access stream with as little arithmetic
// as it needs. The same rays and sample counts visit the same cells (their
// positions from the ray's entry point, one FMA per axis), one 16-byte cell
// load per sample, 4 in flight per lane, folded by XOR so that no load is
// dead. What it returns is not a DRR; its duration measures what the memory
// hierarchy delivers for this access pattern (16-bit storages in the OCTET
// and BRICK layouts, whose cells are 16 bytes; other volumes return 0 and
// count no samples).
template <int LY>
__device__ __forceinline__ uint32_t gather_probe_st(const VolDev& v, const Ray& r, int& nsamples, const int* tab)
{
  const VolRsrc rs = vol_rsrc(v, tab);
  const VoxelLine l = voxel_line(r, (float)v.W, (float)v.H, (float)v.D);
  const int n = sample_count(r);
  uint32_t acc = 0;
  float kf = 0.0f;
  int k = 0;
  auto one = [&](float kk) {
    const float X = fmaf(kk, l.nbx, l.ax), Y = fmaf(kk, l.nby, l.ay), Z = fmaf(kk, l.nbz, l.az);
    const int voff = cell_offset<ASC_STORAGE_U16, LY>(rs, cvt_flr(X), cvt_flr(Y), cvt_flr(Z));
    const auto q = __builtin_amdgcn_raw_buffer_load_b128(rs.rsrc, voff, 0, 0);
    acc ^= q[0] ^ q[1] ^ q[2] ^ q[3];
  };
  for (; k + 4 <= n; k += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) one(kf + (float)u);
    kf += 4.0f;
  }
  for (; k < n; k++) {
    one(kf);
    kf += 1.0f;
  }
  nsamples += n;
  return acc;
}

__device__ __forceinline__ float gather_probe(const VolDev& v, const Ray& r, int& nsamples, const int* tab)
{
  uint32_t acc = 0;
  if (v.storage != ASC_STORAGE_F32) {
    if (v.layout == ASC_LAYOUT_BRICK)
      acc = gather_probe_st<ASC_LAYOUT_BRICK>(v, r, nsamples, tab);
    else if (v.layout == ASC_LAYOUT_OCTET)
      acc = gather_probe_st<ASC_LAYOUT_OCTET>(v, r, nsamples, tab);
  }
  return __uint_as_float(acc & 0x007FFFFFu);  // a finite value depending on every load
}

template <int ST, int LY, int UNROLL>
__device__ __forceinline__ float march_st(const VolDev& v, const Ray& r, float cutoff, int& nsamples,
                                          const int* tab)
{
  if (cutoff > 0.0f) return march_loop<ST, LY, true, UNROLL>(v, r, cutoff, nsamples, tab);
  return march_loop<ST, LY, false, UNROLL>(v, r, cutoff, nsamples, tab);
}

// Dispatch on the volume's storage and layout (uniform per wave).
// tab: the block's BRICK tables for v (brick_tables, synchronised)
template <int UNROLL>
__device__ __forceinline__ float march(const VolDev& v, const Ray& r, float cutoff, int& nsamples, const int* tab)
{
#define ASC_MARCH_ST(ST)                                                                  \
  switch (v.layout) {                                                                     \
    case ASC_LAYOUT_QUAD: return march_st<ST, ASC_LAYOUT_QUAD, UNROLL>(v, r, cutoff, nsamples, tab);   \
    case ASC_LAYOUT_OCTET: return march_st<ST, ASC_LAYOUT_OCTET, UNROLL>(v, r, cutoff, nsamples, tab); \
    case ASC_LAYOUT_BRICK: return march_st<ST, ASC_LAYOUT_BRICK, UNROLL>(v, r, cutoff, nsamples, tab); \
    default: return march_st<ST, ASC_LAYOUT_LINEAR, UNROLL>(v, r, cutoff, nsamples, tab);              \
  }
  switch (v.storage) {
    case ASC_STORAGE_U16:
      if (v.pair_diff == 2) {
        switch (v.layout) {
          case ASC_LAYOUT_QUAD: return march_st<kStorageU16Grad, ASC_LAYOUT_QUAD, UNROLL>(v, r, cutoff, nsamples, tab);
          case ASC_LAYOUT_OCTET: return march_st<kStorageU16Grad, ASC_LAYOUT_OCTET, UNROLL>(v, r, cutoff, nsamples, tab);
          default: return march_st<kStorageU16Grad, ASC_LAYOUT_BRICK, UNROLL>(v, r, cutoff, nsamples, tab);
        }
      }
      if (v.pair_diff) {
        switch (v.layout) {
          case ASC_LAYOUT_QUAD: return march_st<kStorageU16Diff, ASC_LAYOUT_QUAD, UNROLL>(v, r, cutoff, nsamples, tab);
          case ASC_LAYOUT_OCTET: return march_st<kStorageU16Diff, ASC_LAYOUT_OCTET, UNROLL>(v, r, cutoff, nsamples, tab);
          default: return march_st<kStorageU16Diff, ASC_LAYOUT_BRICK, UNROLL>(v, r, cutoff, nsamples, tab);
        }
      }
      ASC_MARCH_ST(ASC_STORAGE_U16)
    case ASC_STORAGE_F16: ASC_MARCH_ST(ASC_STORAGE_F16)
    case ASC_STORAGE_BF16: ASC_MARCH_ST(ASC_STORAGE_BF16)
    default: ASC_MARCH_ST(ASC_STORAGE_F32)
  }
#undef ASC_MARCH_ST
}

// density = sum(step * val); out = clamp(density / intensity, 0, 1)
__device__ __forceinline__ float drr_finish(float acc_raw, float step_vscale, float intensity)
{
  const float d = (acc_raw * step_vscale) / intensity;
  return d < 0.0f ? 0.0f : (d > 1.0f ? 1.0f : d);
}

// ---------------------------------------------------------------------------
// 2-D image sampling, RadRenderer.cl / BackgroundRenderer.cl (exact order).
// ---------------------------------------------------------------------------
// Images hold texels already normalised at upload (raw / (2^bits - 1),
// correctly rounded): the sampler's UNORM read is a plain load here.
// Bilinear sample at sampler coordinates (s, t) in [0, 1] (every caller
// returns before sampling outside), OpenCL 1.2 section 8.2: x = fma(s, W, -0.5),
// i0 = floor(x), a = x - i0, texels outside the image read the border 0
// (CLK_ADDRESS_CLAMP). The texels come through buffer loads, whose
// out-of-range offsets read 0: a texel whose row or column lies outside the
// image is sent out of range explicitly (the image holds at most 2^28
// texels, checked on the host).
// A buffer resource over an image's float texels (range-checked loads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t img_rsrc(const ImgDev& im)
{
  const int W = __builtin_amdgcn_readfirstlane(im.W), H = __builtin_amdgcn_readfirstlane(im.H);
  const uint64_t base = (uint64_t)(uintptr_t)im.data;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)(((uint64_t)hi << 32) | lo), 0, W * H * 4, 0x00020000);
}

__device__ __forceinline__ float img_sample(const ImgDev& im, float s, float t)
{
  const int W = __builtin_amdgcn_readfirstlane(im.W), H = __builtin_amdgcn_readfirstlane(im.H);
  const __amdgpu_buffer_rsrc_t rs = img_rsrc(im);
  // the un-normalisation as in the 3-D sampler: one FMA per axis
  const float x = fmaf(s, (float)W, -0.5f);
  const float y = fmaf(t, (float)H, -0.5f);
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  const int i0 = (int)fx, j0 = (int)fy;
  const int row0 = (j0 * W + i0) * 4, row1 = row0 + W * 4;
  const bool c0 = (unsigned)i0 < (unsigned)W, c1 = (unsigned)(i0 + 1) < (unsigned)W;
  const bool r0 = (unsigned)j0 < (unsigned)H, r1 = (unsigned)(j0 + 1) < (unsigned)H;
  // a texel outside the image is loaded at offset 2^31: past the image
  // (at most 2^30 bytes), and far enough from 2^32 that the range check
  // (offset plus the access size, in 32 bits) cannot wrap into range, as
  // an offset such as -1 would
  constexpr int kOut = (int)0x80000000u;
  const float T00 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, c0 && r0 ? row0 : kOut, 0, 0));
  const float T10 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, c1 && r0 ? row0 + 4 : kOut, 0, 0));
  const float T01 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, c0 && r1 ? row1 : kOut, 0, 0));
  const float T11 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, c1 && r1 ? row1 + 4 : kOut, 0, 0));
  return (1.0f - a) * (1.0f - b) * T00 + a * (1.0f - b) * T10 + (1.0f - a) * b * T01 + a * b * T11;
}

// One pixel of rad_render_kernel (background = false) or
// background_render_kernel (background = true).
__device__ __forceinline__ float rad_pixel(const ImgDev& im, const float* __restrict__ plane,
                                           const float* __restrict__ vp, unsigned x, unsigned y, unsigned w,
                                           unsigned h, bool background, float threshold)
{
  const float u = vp[0] + vp[2] * ((float)x / (float)w);
  const float v = vp[1] + vp[3] * ((float)y / (float)h);
  const float s = (u - plane[0]) / plane[2] + 0.5f;
  const float t = 0.5f - (v - plane[1]) / plane[3];
  if (s < 0.0f || t < 0.0f || s > 1.0f || t > 1.0f) return 0.0f;
  const float val = img_sample(im, s, t);
  if (background) return (threshold <= val) ? 1.0f : 0.0f;
  return 1.0f - val;
}

// ---------------------------------------------------------------------------
// Filter pixel functions over an accessor get(x, y) that returns the input
// at in-image coordinates. Same operation order as the reference kernels.
// ---------------------------------------------------------------------------
// SobelFilter.cl:10-47, from the 3 x 3 neighbourhood p[row][column].
__device__ __forceinline__ float sobel_core(float p00, float p01, float p02, float p10, float p11, float p12,
                                            float p20, float p21, float p22, float scale, float blend)
{
  const float horz = p02 + 2.0f * p12 + p22 - p00 - 2.0f * p10 - p20;
  const float vert = p00 + 2.0f * p01 + p02 - p20 - 2.0f * p21 - p22;
  const float grad = sqrtf(horz * horz + vert * vert);
  float sum;
  if (blend < 0.5f)
    sum = scale * grad + 2.0f * blend * p11;
  else
    sum = 2.0f * (1.0f - blend) * scale * grad + p11;
  if (sum < 0.0f)
    sum = 0.0f;
  else if (sum > 1.0f)
    sum = 1.0f;
  return sum;
}

// The neighbourhood clamped to the image, as the reference kernel reads it.
template <class G>
__device__ __forceinline__ float sobel_px(const G& get, int x1, int y1, int w, int h, float scale, float blend)
{
  const int x0 = x1 - 1 < 0 ? 0 : x1 - 1, y0 = y1 - 1 < 0 ? 0 : y1 - 1;
  const int x2 = x1 + 1 > w - 1 ? w - 1 : x1 + 1, y2 = y1 + 1 > h - 1 ? h - 1 : y1 + 1;
  return sobel_core(get(x0, y0), get(x1, y0), get(x2, y0), get(x0, y1), get(x1, y1), get(x2, y1), get(x0, y2),
                    get(x1, y2), get(x2, y2), scale, blend);
}

// x^y of the Contrast filter. For a positive normal x and y != 0 it is
// 2^(y * log2 x) on the hardware base-2 logarithm and exponential, when the
// result stays a normal number; everything else (zero, negative, subnormal
// or non-finite x, y == 0, results outside the normal range) goes to powf.
// The fast form is within a few units in the last place for the inputs the
// filter sees (its error grows with |y * log2 x|: about 2^-24 * ln 2 of it,
// plus an ulp from each hardware step), inside the OpenCL pow's 16-ulp
// allowance that the reference kernel runs under.
__device__ __forceinline__ float contrast_pow(float x, float y)
{
  if (x >= 1.17549435e-38f && x <= 3.40282347e+38f && y != 0.0f) {
    const float e = y * __builtin_amdgcn_logf(x);
    if (e > -126.0f && e < 128.0f) return __builtin_amdgcn_exp2f(e);
  }
  return powf(x, y);
}

__device__ __forceinline__ float contrast_finish(float fxy, float axy, float alpha, float beta)
{
  float gxy = 0.0f;
  if (axy > 0.01f) gxy = contrast_pow(axy, alpha - beta) * contrast_pow(fxy, beta);
  return gxy;
}

// ContrastFilter.cl:11-38
template <class G>
__device__ __forceinline__ float contrast_px(const G& get, int x, int y, int w, int h, float alpha, float beta,
                                             int size)
{
  const float fxy = get(x, y);
  float n = 0.0f, sum = 0.0f;
  const int minI = max(y - size / 2, 0), maxI = min(y + (size + 1) / 2, h);
  const int minJ = max(x - size / 2, 0), maxJ = min(x + (size + 1) / 2, w);
  for (int i = minI; i < maxI; ++i)
    for (int j = minJ; j < maxJ; ++j) {
      n += 1.0f;
      sum += get(j, i);
    }
  return contrast_finish(fxy, sum / n, alpha, beta);
}

// The 3 x 3 window wholly inside the image: nine terms in the kernel's
// order (rows outer, columns inner), n = 9.
__device__ __forceinline__ float contrast3_core(float p00, float p01, float p02, float p10, float p11, float p12,
                                                float p20, float p21, float p22, float alpha, float beta)
{
  float sum = 0.0f;
  sum += p00;
  sum += p01;
  sum += p02;
  sum += p10;
  sum += p11;
  sum += p12;
  sum += p20;
  sum += p21;
  sum += p22;
  return contrast_finish(p11, sum / 9.0f, alpha, beta);
}

// GaussianFilter.cl:15-27 / SharpenFilter.cl:18-30: zero-padded N x N
// convolution, i walks x (outer), j walks y (inner), weight index i*N+j.
template <class G>
__device__ __forceinline__ float conv_px(const G& get, int x, int y, int w, int h, const float* __restrict__ wts,
                                         int n)
{
  const int fr = (n - 1) / 2;
  float acc = 0.0f;
  for (int i = 0; i < n; ++i) {
    const int a = x - fr + i;
    if (a < 0 || a >= w) continue;
    for (int j = 0; j < n; ++j) {
      const int b = y - fr + j;
      if (!(b < 0 || b >= h)) acc = acc + wts[i * n + j] * get(a, b);
    }
  }
  return acc;
}

__device__ __forceinline__ float gaussian_finish(float c)
{
  if (c > 1.0f) c = 1.0f;
  if (c < 0.0f) c = 0.0f;
  return c;
}

// SharpenFilter.cl:32-45
__device__ __forceinline__ float sharpen_finish(float f, float blur, float contrast, float threshold)
{
  if (fabsf(f - blur) > threshold) {
    float o = f + contrast * (f - blur);
    if (o > 1.0f) o = 1.0f;
    if (o < 0.0f) o = 0.0f;
    return o;
  }
  return f;
}

// Halo (pixels) a filter needs on each side.
__host__ __device__ inline int filter_radius(const asc_filter& f, int gauss_n)
{
  if (!f.enabled) return 0;
  switch (f.type) {
    case ASC_FILTER_SOBEL: return 1;
    case ASC_FILTER_CONTRAST: {
      const int size = (int)f.p[2];
      const int lo = size / 2, hi = (size + 1) / 2 - 1;
      return lo > hi ? lo : hi;
    }
    case ASC_FILTER_GAUSSIAN:
    case ASC_FILTER_SHARPEN: return gauss_n > 1 ? (gauss_n - 1) / 2 : 0;
    default: return 0;
  }
}

// ---------------------------------------------------------------------------
// NCC moments: per-thread fp32 partials with compensation (the running sum and
// its exact rounding errors), combined
// in double across lanes, waves and blocks.
// ---------------------------------------------------------------------------
struct Comp {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x)
  {
    // Knuth's TwoSum: the exact rounding error of s + x, without a branch
    // (Neumaier's two cases give the same value)
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

struct Moments {
  // {sum mask, sum f, sum g, count, S_f, S_g, S_ff, S_gg, S_fg}
  Comp m, f, g, sf, sg, sff, sgg, sfg;
  int count = 0;
  __device__ __forceinline__ void add(float fv, float gv, float mv)
  {
    m.add(mv);
    f.add(fv);
    g.add(gv);
    if (mv > 0.5f) {
      ++count;
      sf.add(fv);
      sg.add(gv);
      sff.add(fv * fv);
      sgg.add(gv * gv);
      sfg.add(fv * gv);
    }
  }
  __device__ __forceinline__ void to_double(double* o) const
  {
    o[0] = m.value();
    o[1] = f.value();
    o[2] = g.value();
    o[3] = (double)count;
    o[4] = sf.value();
    o[5] = sg.value();
    o[6] = sff.value();
    o[7] = sgg.value();
    o[8] = sfg.value();
  }
};

// Moments of a post-processing tile, where the mask is 0 or 1 and f and g
// are already multiplied by it: a masked-out pixel adds zeros, which leave
// a compensated sum as it was, so sum f = S_f, sum g = S_g and sum mask =
// count, and six sums give the nine of Moments with the same bits.
struct MaskedMoments {
  int m = 0;  // the mask is 0 or 1: its sum is a count
  Comp f, g, ff, gg, fg;
  __device__ __forceinline__ void add(float fv, float gv, float mv)
  {
    m += mv > 0.5f;
    f.add(fv);
    g.add(gv);
    ff.add(fv * fv);
    gg.add(gv * gv);
    fg.add(fv * gv);
  }
  __device__ __forceinline__ void to_double6(double* o) const
  {
    o[0] = (double)m;
    o[1] = f.value();
    o[2] = g.value();
    o[3] = ff.value();
    o[4] = gg.value();
    o[5] = fg.value();
  }
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum of N doubles (blockDim.x == 256). Result valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(double* v, double* lds /* 4*N */)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; k++) lds[wave * N + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = (lds[k] + lds[N + k]) + (lds[2 * N + k] + lds[3 * N + k]);
}

__device__ __forceinline__ void block_sum9(double* v, double* lds /* 4*9 */) { block_sum<9>(v, lds); }

}  // namespace dev
}  // namespace asc
