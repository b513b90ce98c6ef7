// Scene + fused batched evaluation of Tracker::trackFrame (Tracker.cpp:583-773)
// for many poses x cameras at once.
//
// One batch = n poses x C cameras = n*C "items". Per item the host computes
// the exact double-precision geometry (inverse model-view, projected
// bounding box, render size). The device then runs three kernels:
//   A  drr_tiles   : persistent; 16x16-pixel tiles of every item's bounding
//                    box; ray march -> DRR scratch (one image per item).
//   B  post_tiles  : persistent; per 32x32 tile: DRR filter chain and
//                    radiograph resample + filter chain in LDS with halos,
//                    DRR mask, multiplies, nine NCC moments (four pixels
//                    per thread, compensated) -> per-tile partial
//                    (9 doubles).
//   C  finalize    : one block per item: fixed-order sum of its partials,
//                    NCC in double; resets the work counters for the next
//                    batch.
// The three launches have fixed grids and read all sizes from device memory,
// so a batch can be recorded once in a hipGraph and replayed.
//
// Tile order and L2 locality: tiles are enumerated as (camera, tile row,
// tile column, pose), so consecutive tiles are the same image region of
// different poses -- nearly the same rays through the same voxels when the
// poses are a PSO swarm. The enumeration is cut into 8 contiguous
// partitions, one per XCD (workgroups are dealt round-robin to the 8 XCDs,
// so block b runs on XCD b % 8 and pulls from partition b % 8 first); each
// XCD's L2 therefore holds one slab of the volume. A block whose partition
// is drained steals from the next ones.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "device.hpp"

namespace asc {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 16;   // ray-march tile side
constexpr int kPTile = 32;  // post-processing tile side (4 pixels per thread)
constexpr int kChainCopies = 0, kChainSobel = 1, kChainGeneral = 2;
// pixels of a thread whose Sobel neighbourhoods are loaded together (9 loads each)
constexpr int kDirectGroup = 2;
constexpr int kMaxChain = 8;
constexpr int kMaxHalo = 12;
constexpr int kSmallHalo = 4;  // chains up to this halo use the smaller-LDS post kernel

struct ChainDev {
  int n;
  int halo;
  // the chain's stages other than copies: kChainCopies (none: the output is
  // the input), kChainSobel (one Sobel) or kChainGeneral
  int shape;
  asc_filter f[kMaxChain];
  int gn[kMaxChain];    // Gaussian/sharpen size N (1 = copy)
  int woff[kMaxChain];  // offset of the weights in the weight pool
  int radius[kMaxChain];
};

struct CamDev {
  ImgDev bg;
  int has_bg;
  float bg_threshold;
  float plane[4];  // camera viewport = RadRenderer image plane (Tracker.cpp:327-329)
  ChainDev chain[2];
};

struct VolParams {
  VolDev v[3];  // the volume's view for rays along x, y, z (asc_volume_s::odev)
  float step, intensity, cutoff, step_vscale;
};

struct ItemDesc {
  float imv[12];
  float vp[4];
  ImgDev rad;  // the radiograph of this item's frame and camera
  int w, h;
  int vol, cam;
  int axis;                // view of the volume: ray_axis of imv
  int tiles_x, tiles_y;    // ray-march tiles (kTile)
  int ptiles_x, ptiles_y;  // post tiles (kPTile)
  int tile_base;  // first tile of this item in the batch's tile list
  int part_off;   // first partial slot (one per post tile)
  long long drr_off;
};

constexpr int kMaxCams = 64;    // camera slots per scene (all jobs of a server)
constexpr int kMaxSetCams = 8;  // cameras per camera set (one job's views)
constexpr int kParts = 8;       // XCDs per MI355X

// A (camera, tile row, tile column, item) enumeration of the batch's tiles,
// over the cameras that have items in the batch ("active" cameras). Items
// are stored camera-major: the items of active camera a are
// item_base[a] .. item_base[a] + cam_items[a] - 1.
struct TileEnum {
  int total;
  int pad;
  int cam_base[kMaxCams];  // first enumeration index of each active camera
  int tiles_x[kMaxCams];   // widest item of the camera, in tiles
  int tiles_y[kMaxCams];
};

struct BatchHeader {
  int n_items;
  int debug;
  int n_acams;                // active cameras
  int pad;
  int cam_items[kMaxCams];    // items per active camera
  int item_base[kMaxCams];    // first item of each active camera
  TileEnum drr;   // kTile tiles
  TileEnum post;  // kPTile tiles
};

// Work counters, one per XCD partition and kernel, each on a 256-byte
// line of its own: device-scope atomics to one address are serialised
// where they are performed, so counters sharing a line would queue every
// claim of the launch behind one another.
constexpr int kCtrStride = 64;  // unsigned ints between counters (256 B)
struct Counters {
  unsigned int drr_next[kParts * kCtrStride];
  unsigned int post_next[kParts * kCtrStride];
};

struct BatchArgs {
  const BatchHeader* hdr;
  const ItemDesc* items;
  const CamDev* cams;
  const VolParams* vols;
  const float* weights;
  float* drr;
  double* partials;
  double* results;  // n_items * 10, in the submission slot's pinned host buffer
  unsigned long long* samples;  // n_items
  unsigned long long* samples_host;  // n_items, in the submission slot's pinned host buffer
  Counters* counters;
  float* dbg_drr;
  float* dbg_rad;
  float* dbg_mask;
  int wave_shift;  // log2 of the width of one wave's ray patch in a 16x16 tile: 3 (8x8), 4 (16x4) or 2 (4x16)
};

// Next enumeration index for the whole block, or -1 once every partition
// is drained. `part` counts the partitions this block has exhausted; it
// starts with its own XCD's. One atomic per tile: claiming several tiles at
// once, or the next tile ahead, widens the range of tiles in flight and so
// the volume region the L2s must hold (both measured slower).
__device__ __forceinline__ int next_tile(unsigned int* counters, int total, int* slot, int& part)
{
  const int chunk = (total + kParts - 1) / kParts;
  while (part < kParts) {
    const int p = (int)((blockIdx.x + part) % kParts);
    __syncthreads();
    if (threadIdx.x == 0) *slot = p * chunk + (int)atomicAdd(&counters[p * kCtrStride], 1u);
    __syncthreads();
    // the slot is block-uniform: say so, so that everything derived from it
    // (item, volume, descriptors) lives in scalar registers
    const int t = __builtin_amdgcn_readfirstlane(*slot);
    if (t < min((p + 1) * chunk, total)) return t;
    ++part;
  }
  return -1;
}

// Enumeration index -> (item, tile column, tile row); false when the tile
// lies outside that item's bounding box.
template <bool POST>
__device__ __forceinline__ bool decode_tile(const BatchHeader& hd, const ItemDesc* __restrict__ items, int t,
                                            int& ii, int& tx, int& ty)
{
  const TileEnum& en = POST ? hd.post : hd.drr;
  // last active camera whose range starts at or before t
  int lo = 0, hi = hd.n_acams - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (en.cam_base[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int c = lo;
  const int r = t - en.cam_base[c];
  const int n_c = hd.cam_items[c];
  const int e = r % n_c;
  const int q = r / n_c;
  tx = q % en.tiles_x[c];
  ty = q / en.tiles_x[c];
  ii = hd.item_base[c] + e;
  const ItemDesc& it = items[ii];
  return POST ? (tx < it.ptiles_x && ty < it.ptiles_y) : (tx < it.tiles_x && ty < it.tiles_y);
}

template <int UNROLL>
__device__ __forceinline__ float march_item(const VolParams& vp, const ItemDesc& it, unsigned x, unsigned y,
                                            int& nsamp, const int* tab)
{
  dev::Ray r;
  if (!dev::ray_setup(it.imv, it.vp, vp.step, x, y, (unsigned)it.w, (unsigned)it.h, r)) return 0.0f;
  return dev::drr_finish(dev::march<UNROLL>(vp.v[it.axis], r, vp.cutoff, nsamp, tab), vp.step_vscale,
                         vp.intensity);
}

__device__ __forceinline__ float probe_item(const VolParams& vp, const ItemDesc& it, unsigned x, unsigned y,
                                            int& nsamp, const int* tab)
{
  dev::Ray r;
  if (!dev::ray_setup(it.imv, it.vp, vp.step, x, y, (unsigned)it.w, (unsigned)it.h, r)) return 0.0f;
  return dev::gather_probe(vp.v[it.axis], r, nsamp, tab);
}

// PROBE: the gather probe in place of the march (same tiles, rays and
// sample counts; the images it leaves are not DRRs).
template <int UNROLL, bool PROBE = false>
__global__ __launch_bounds__(kBlock) void drr_tiles(BatchArgs a)
{
  __shared__ int s_tile;
  __shared__ unsigned long long s_samples;
  __shared__ int s_tab[kBrickTab];  // BRICK offset tables of view tab_view
  const BatchHeader& hd = *a.hdr;
  const int total = hd.drr.total;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int part = 0, tab_view = -1;
  for (;;) {
    const int t = next_tile(a.counters->drr_next, total, &s_tile, part);
    if (t < 0) break;
    int ii, tx, ty;
    if (!decode_tile<false>(hd, a.items, t, ii, tx, ty)) continue;
    const ItemDesc& it = a.items[ii];
    if (it.vol * 3 + it.axis != tab_view) {
      // every wave has passed next_tile's barrier, so none still reads the
      // tables of the previous view
      dev::brick_tables(a.vols[it.vol].v[it.axis], s_tab);
      __syncthreads();
      tab_view = it.vol * 3 + it.axis;
    }
    // each of the tile's 4 waves marches a pw x (64 / pw) patch of rays,
    // pw = 2^wave_shift; the patches cover the 16 x 16 tile row by row.
    // Inside a patch, lanes 4q..4q+3 take the q-th 2 x 2 block of rays (row
    // by row): the vector memory unit serves a quad of lanes per cycle for
    // each cache line its four addresses touch, and the samples of a 2 x 2
    // block of neighbouring rays share a brick's line more often than those
    // of a row of four.
    const int wsh = a.wave_shift, pw = 1 << wsh, ph = 64 >> wsh;
    const int wpr = kTile >> wsh;  // waves per tile row
    const unsigned q = lane >> 2, qsh = (unsigned)wsh - 1u;
    const unsigned lx = (lane & 1u) | ((q & ((1u << qsh) - 1u)) << 1);
    const unsigned ly = ((lane >> 1) & 1u) | ((q >> qsh) << 1);
    const unsigned x = tx * kTile + (wave % wpr) * pw + lx;
    const unsigned y = ty * kTile + (wave / wpr) * ph + ly;
    if (threadIdx.x == 0) s_samples = 0;
    __syncthreads();
    int nsamp = 0;
    if (x < (unsigned)it.w && y < (unsigned)it.h) {
      const VolParams& vp = a.vols[it.vol];
      const float val =
        PROBE ? probe_item(vp, it, x, y, nsamp, s_tab) : march_item<UNROLL>(vp, it, x, y, nsamp, s_tab);
      a.drr[it.drr_off + (long long)y * it.w + x] = val;
    }
    // per-tile sample count (for the algorithmic-bytes figure)
    int ws = nsamp;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ws += __shfl_xor(ws, off, 64);
    if (lane == 0) atomicAdd(&s_samples, (unsigned long long)ws);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&a.samples[ii], s_samples);
  }
}

// LDS region accessor: region origin (ox, oy) in image coordinates, row
// stride R.
template <int R>
struct LdsGet {
  const float* buf;
  int ox, oy;
  __device__ __forceinline__ float operator()(int x, int y) const { return buf[(y - oy) * R + (x - ox)]; }
};

// Visit every in-image pixel of the side x side square at (rx0, ry0), one
// flat index k per thread, stepping by the block: thread t starts at
// (t % side, t / side) and each step adds kBlock = q rows + r columns (the
// float quotients below are exact: side <= 56 leaves their floors a margin
// of 1 / (2 side), far above the rounding error).
template <class F>
__device__ __forceinline__ void for_region(int side, int rx0, int ry0, int w, int h, F&& fn)
{
  const int n = side * side;
  const float inv = 1.0f / (float)side;
  int yy = (int)(((float)threadIdx.x + 0.5f) * inv);
  int xx = (int)threadIdx.x - yy * side;
  const int q = (int)(((float)kBlock + 0.5f) * inv);
  const int r = kBlock - q * side;
  for (int k = threadIdx.x; k < n; k += kBlock) {
    const int px = rx0 + xx, py = ry0 + yy;
    if ((unsigned)px < (unsigned)w && (unsigned)py < (unsigned)h) fn(xx, yy, px, py);
    xx += r;
    yy += q;
    if (xx >= side) {
      xx -= side;
      ++yy;
    }
  }
}

// for_region for a square wholly inside the image: every position is
// visited, fn gets the LDS index c = base + yy * R + xx (stepped along with
// xx and yy, no multiply) and (xx, yy).
template <int R, class F>
__device__ __forceinline__ void for_square(int side, int base, F&& fn)
{
  const int n = side * side;
  const float inv = 1.0f / (float)side;
  int yy = (int)(((float)threadIdx.x + 0.5f) * inv);
  int xx = (int)threadIdx.x - yy * side;
  const int q = (int)(((float)kBlock + 0.5f) * inv);
  const int r = kBlock - q * side;
  int c = base + yy * R + xx;
  for (int k = threadIdx.x; k < n; k += kBlock) {
    fn(c, xx, yy);
    xx += r;
    yy += q;
    c += q * R + r;
    if (xx >= side) {
      xx -= side;
      ++yy;
      c += R - side;
    }
  }
}

// Run a filter chain over the region in LDS (row stride R). On entry `cur`
// holds stage-0 values at every in-image position of the post tile grown by
// the chain's halo; on exit the returned buffer holds the chain output for
// the tile. Each stage's kind is uniform, so each has a loop of its own;
// Sobel and the 3 x 3 Contrast read a pixel's neighbourhood straight from
// LDS where it lies inside the image (the same values the clamped form
// reads there).
template <int R>
__device__ __forceinline__ float* run_chain(const ChainDev& ch, const float* __restrict__ weights, float* cur, float* nxt, int x0,
                            int y0, int w, int h)
{
  const int H0 = ch.halo;
  const int ox = x0 - H0, oy = y0 - H0;  // region origin
  int remaining = H0;
  for (int k = 0; k < ch.n; k++) {
    const asc_filter& f = ch.f[k];
    remaining -= ch.radius[k];
    // output rectangle: tile grown by `remaining`
    const int rx0 = x0 - remaining, ry0 = y0 - remaining;
    const int side = kPTile + 2 * remaining;
    const LdsGet<R> get{cur, ox, oy};
    float* const out = nxt;
    const float* const in = cur;
    auto at = [&](int px, int py) { return (py - oy) * R + (px - ox); };
    // the stage's outputs and their 3 x 3 neighbourhoods all inside the
    // image (a block-uniform test): no per-pixel bounds or border tests
    const bool inner = rx0 >= 1 && ry0 >= 1 && rx0 + side <= w - 1 && ry0 + side <= h - 1;
    const int base = at(rx0, ry0);
    if (f.enabled && f.type == ASC_FILTER_SOBEL && inner) {
      const float scale = f.p[0], blend = f.p[1];
      for_square<R>(side, base, [&](int c, int, int) {
        out[c] = dev::sobel_core(in[c - R - 1], in[c - R], in[c - R + 1], in[c - 1], in[c], in[c + 1], in[c + R - 1],
                                 in[c + R], in[c + R + 1], scale, blend);
      });
    } else if (f.enabled && f.type == ASC_FILTER_CONTRAST && inner && (int)f.p[2] == 3) {
      const float alpha = f.p[0], beta = f.p[1];
      for_square<R>(side, base, [&](int c, int, int) {
        out[c] = dev::contrast3_core(in[c - R - 1], in[c - R], in[c - R + 1], in[c - 1], in[c], in[c + 1],
                                     in[c + R - 1], in[c + R], in[c + R + 1], alpha, beta);
      });
    } else if (f.enabled && f.type == ASC_FILTER_SOBEL) {
      const float scale = f.p[0], blend = f.p[1];
      for_region(side, rx0, ry0, w, h, [&](int, int, int px, int py) {
        const int c = at(px, py);
        float v;
        if (px >= 1 && px <= w - 2 && py >= 1 && py <= h - 2)
          v = dev::sobel_core(in[c - R - 1], in[c - R], in[c - R + 1], in[c - 1], in[c], in[c + 1], in[c + R - 1],
                              in[c + R], in[c + R + 1], scale, blend);
        else
          v = dev::sobel_px(get, px, py, w, h, scale, blend);
        out[c] = v;
      });
    } else if (f.enabled && f.type == ASC_FILTER_CONTRAST) {
      const float alpha = f.p[0], beta = f.p[1];
      const int size = (int)f.p[2];
      for_region(side, rx0, ry0, w, h, [&](int, int, int px, int py) {
        const int c = at(px, py);
        float v;
        if (size == 3 && px >= 1 && px <= w - 2 && py >= 1 && py <= h - 2)
          v = dev::contrast3_core(in[c - R - 1], in[c - R], in[c - R + 1], in[c - 1], in[c], in[c + 1],
                                  in[c + R - 1], in[c + R], in[c + R + 1], alpha, beta);
        else
          v = dev::contrast_px(get, px, py, w, h, alpha, beta, size);
        out[c] = v;
      });
    } else if (f.enabled && ch.gn[k] > 1 && (f.type == ASC_FILTER_GAUSSIAN || f.type == ASC_FILTER_SHARPEN)) {
      const float* wts = weights + ch.woff[k];
      const int gn = ch.gn[k];
      const bool sharpen = f.type == ASC_FILTER_SHARPEN;
      const float contrast = f.p[1] < 1.0f ? 1.0f : f.p[1], threshold = f.p[2];
      for_region(side, rx0, ry0, w, h, [&](int, int, int px, int py) {
        const float c = dev::conv_px(get, px, py, w, h, wts, gn);
        out[at(px, py)] = sharpen ? dev::sharpen_finish(get(px, py), c, contrast, threshold) : dev::gaussian_finish(c);
      });
    } else {
      // disabled, or a Gaussian/Sharpen of size <= 1: a copy
      for_region(side, rx0, ry0, w, h, [&](int, int, int px, int py) { out[at(px, py)] = in[at(px, py)]; });
    }
    __syncthreads();
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  return cur;
}

// Radiograph sampler coordinate of pixel column x / row y of a w x h render
// (RadRenderer.cl:8-16): s depends on x only and t on y only, so a tile
// computes them once per column and row. Outside [0, 1] -> -1 (the kernel
// returns 0 there).
__device__ __forceinline__ float rad_s(const float* __restrict__ plane, const float* __restrict__ vp, unsigned x,
                                       unsigned w)
{
  const float u = vp[0] + vp[2] * ((float)x / (float)w);
  const float s = (u - plane[0]) / plane[2] + 0.5f;
  return (s < 0.0f || s > 1.0f) ? -1.0f : s;
}

__device__ __forceinline__ float rad_t(const float* __restrict__ plane, const float* __restrict__ vp, unsigned y,
                                       unsigned h)
{
  const float v = vp[1] + vp[3] * ((float)y / (float)h);
  const float t = 0.5f - (v - plane[1]) / plane[3];
  return (t < 0.0f || t > 1.0f) ? -1.0f : t;
}

// The radiograph sampler's arithmetic (device.hpp img_sample) for one
// column or one row of a tile region, done once per column and row: the
// texel index i0 = floor(s n - 1/2) becomes the byte offsets of texel
// columns i0 and i0 + 1 (of texel rows j0 and j0 + 1 for a row), kSampOut
// where they lie outside the image, and the weight a with 1 - a. A texel's
// offset is the sum of its column's and its row's: in the image when both
// are, otherwise at least 2^30 (past the image, which holds at most 2^30
// bytes) and at most 2^31 (where the range check, offset plus access size
// in 32 bits, cannot wrap), so that it reads the border 0. A coordinate
// outside [0, 1] (-1 from rad_s / rad_t) gets w = -1: its pixels are 0.
constexpr unsigned kSampOut = 0x40000000u;
struct SampAxis {
  unsigned o0, o1;
  float w, nw;  // a and 1 - a
};

__device__ __forceinline__ SampAxis samp_axis(float s, int n, unsigned stride)
{
  SampAxis r;
  const float x = fmaf(s, (float)n, -0.5f);
  const float fx = floorf(x);
  const int i0 = (int)fx;
  r.w = x - fx;
  r.nw = 1.0f - r.w;
  r.o0 = (unsigned)i0 < (unsigned)n ? (unsigned)i0 * stride : kSampOut;
  r.o1 = (unsigned)(i0 + 1) < (unsigned)n ? (unsigned)(i0 + 1) * stride : kSampOut;
  if (s < 0.0f) {
    r.o0 = r.o1 = kSampOut;
    r.w = -1.0f;
  }
  return r;
}

// Eight waves per SIMD (at most 64 VGPRs, two Sobel neighbourhoods loaded
// together): a block waits on loads and barriers about half of its time,
// and more resident blocks cover it (5 -> 8 waves: 0.49 -> 0.44 ms per
// launch of 100 poses in two 1024^2 cameras). With halos above kSmallHalo
// the LDS holds five blocks per CU whatever the registers.
template <int MAXH>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void post_tiles(BatchArgs a)
{
  constexpr int R = kPTile + 2 * MAXH;  // region row stride
  __shared__ float buf0[R * R];
  __shared__ float buf1[R * R];
  __shared__ float sdrr[kPTile * kPTile];
  __shared__ float s_col[R], s_row[R];
  __shared__ SampAxis s_scol[R], s_srow[R];
  __shared__ double lds_red[4 * 9];
  __shared__ int s_tile;
  const BatchHeader& hd = *a.hdr;
  const int total = hd.post.total;
  const int debug = hd.debug;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  int part = 0;
  for (;;) {
    const int t = next_tile(a.counters->post_next, total, &s_tile, part);
    if (t < 0) break;
    int ii, ptx, pty;
    if (!decode_tile<true>(hd, a.items, t, ii, ptx, pty)) continue;
    const ItemDesc& it = a.items[ii];
    const CamDev& cam = a.cams[it.cam];
    const int local = pty * it.ptiles_x + ptx;
    const int x0 = ptx * kPTile, y0 = pty * kPTile;
    const int w = it.w, h = it.h;

    // --- DRR chain: stage 0 = DRR scratch ---
    if (cam.chain[0].shape != kChainGeneral) {
      // no stage, or one Sobel: each thread's 4 outputs straight from the
      // DRR scratch (Sobel's neighbourhood clamped to the image as the
      // reference's kernel clamps it, the same values the LDS form reads),
      // every load issued before any result is stored; positions outside
      // the image read an in-image pixel and are not stored
      const ChainDev& ch = cam.chain[0];
      const float* src = a.drr + it.drr_off;
      int sob = 0;
      for (int k = 0; k < ch.n; k++)
        if (ch.f[k].enabled && ch.f[k].type == ASC_FILTER_SOBEL) sob = k;
      const float scale = ch.f[sob].p[0], blend = ch.f[sob].p[1];
      const int cx = min(x0 + tx, w - 1), xm = max(cx - 1, 0), xp = min(cx + 1, w - 1);
      float v[4];
      if (ch.shape == kChainCopies) {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = src[(long long)min(y0 + ty + 8 * k, h - 1) * w + cx];
      } else {
#pragma unroll
        for (int k0 = 0; k0 < 4; k0 += kDirectGroup) {
          float p[kDirectGroup][9];
#pragma unroll
          for (int k = k0; k < k0 + kDirectGroup; k++) {
            const int cy = min(y0 + ty + 8 * k, h - 1), ym = max(cy - 1, 0), yp = min(cy + 1, h - 1);
            const float* r0 = src + (long long)ym * w;
            const float* r1 = src + (long long)cy * w;
            const float* r2 = src + (long long)yp * w;
            float* q = p[k - k0];
            q[0] = r0[xm], q[1] = r0[cx], q[2] = r0[xp];
            q[3] = r1[xm], q[4] = r1[cx], q[5] = r1[xp];
            q[6] = r2[xm], q[7] = r2[cx], q[8] = r2[xp];
          }
#pragma unroll
          for (int k = k0; k < k0 + kDirectGroup; k++) {
            const float* q = p[k - k0];
            v[k] = dev::sobel_core(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], scale, blend);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int ly = ty + 8 * k;
        if (x0 + tx < w && y0 + ly < h) sdrr[ly * kPTile + tx] = v[k];
      }
    } else {
      const ChainDev& ch = cam.chain[0];
      const int H0 = ch.halo, side = kPTile + 2 * H0;
      const int ox = x0 - H0, oy = y0 - H0;
      const float* src = a.drr + it.drr_off;
      if (ox >= 0 && oy >= 0 && ox + side <= w && oy + side <= h) {
        // the region wholly inside the image: rows of the DRR straight in
        const float* s0 = src + (long long)oy * w + ox;
        for_square<R>(side, 0, [&](int c, int xx, int yy) { buf0[c] = s0[yy * w + xx]; });
      } else {
        for_region(side, ox, oy, w, h,
                   [&](int xx, int yy, int px, int py) { buf0[yy * R + xx] = src[(long long)py * w + px]; });
      }
      __syncthreads();
      const float* outb = run_chain<R>(ch, a.weights, buf0, buf1, x0, y0, w, h);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int ly = ty + 8 * k;
        sdrr[ly * kPTile + tx] = outb[(ly + H0) * R + (tx + H0)];
      }
      __syncthreads();
    }
    // --- radiograph chain: stage 0 = rad resample (RadRenderer.cl) ---
    const float* radb;
    const int rH0 = cam.chain[1].halo;
    {
      const ChainDev& ch = cam.chain[1];
      const int side = kPTile + 2 * rH0;
      const int ox = x0 - rH0, oy = y0 - rH0;
      for (int k = threadIdx.x; k < 2 * side; k += kBlock) {
        if (k < side) {
          const int px = ox + k;
          if (px >= 0 && px < w) {
            const float ss = rad_s(cam.plane, it.vp, (unsigned)px, (unsigned)w);
            s_col[k] = ss;
            s_scol[k] = samp_axis(ss, it.rad.W, 4u);
          }
        } else {
          const int py = oy + (k - side);
          if (py >= 0 && py < h) {
            const float tt = rad_t(cam.plane, it.vp, (unsigned)py, (unsigned)h);
            s_row[k - side] = tt;
            s_srow[k - side] = samp_axis(tt, it.rad.H, 4u * (unsigned)it.rad.W);
          }
        }
      }
      __syncthreads();
      // img_sample's texels and weighted sum, in its order, from the
      // column's and the row's sampler values
      const __amdgpu_buffer_rsrc_t rrs = dev::img_rsrc(it.rad);
      auto resample = [&](int xx, int yy) {
        const SampAxis c = s_scol[xx], r = s_srow[yy];
        const float T00 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rrs, (int)(c.o0 + r.o0), 0, 0));
        const float T10 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rrs, (int)(c.o1 + r.o0), 0, 0));
        const float T01 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rrs, (int)(c.o0 + r.o1), 0, 0));
        const float T11 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rrs, (int)(c.o1 + r.o1), 0, 0));
        const float v = (c.nw * r.nw) * T00 + (c.w * r.nw) * T10 + (c.nw * r.w) * T01 + (c.w * r.w) * T11;
        return (c.w < 0.0f || r.w < 0.0f) ? 0.0f : 1.0f - v;
      };
      if (ox >= 0 && oy >= 0 && ox + side <= w && oy + side <= h)
        for_square<R>(side, 0, [&](int c, int xx, int yy) { buf0[c] = resample(xx, yy); });
      else
        for_region(side, ox, oy, w, h, [&](int xx, int yy, int, int) { buf0[yy * R + xx] = resample(xx, yy); });
      __syncthreads();
      radb = run_chain<R>(ch, a.weights, buf0, buf1, x0, y0, w, h);
    }
    // --- mask, multiplies, moments (Tracker.cpp:744-752, Ncc.cpp) ---
    dev::MaskedMoments acc;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int ly = ty + 8 * k;
      const int px = x0 + tx, py = y0 + ly;
      if (px < w && py < h) {
        float d = sdrr[ly * kPTile + tx];
        float g = radb[(ly + rH0) * R + (tx + rH0)];
        float m = (d != 0.0f) ? 1.0f : 0.0f;
        if (cam.has_bg) {
          // BackgroundRenderer.cl: same sampler coordinates as the radiograph
          const float ss = s_col[tx + rH0], tt = s_row[ly + rH0];
          const float bgv =
            (ss < 0.0f || tt < 0.0f) ? 0.0f : (cam.bg_threshold <= dev::img_sample(cam.bg, ss, tt) ? 1.0f : 0.0f);
          m = bgv * m;
        }
        g = g * m;
        d = d * m;
        acc.add(d, g, m);
        if (debug) {
          const long long o = it.drr_off + (long long)py * w + px;
          a.dbg_drr[o] = d;
          a.dbg_rad[o] = g;
          a.dbg_mask[o] = m;
        }
      }
    }
    double v[6];
    acc.to_double6(v);
    dev::block_sum<6>(v, lds_red);
    if (threadIdx.x == 0) {
      // {sum mask, sum f, sum g, count, S_f, S_g, S_ff, S_gg, S_fg}
      double* dst = a.partials + (size_t)(it.part_off + local) * 9;
      const double o[9] = {v[0], v[1], v[2], v[0], v[1], v[2], v[3], v[4], v[5]};
#pragma unroll
      for (int k = 0; k < 9; k++) dst[k] = o[k];
    }
  }
}

__global__ __launch_bounds__(kBlock) void finalize_items(BatchArgs a)
{
  __shared__ double lds[4 * 9];
  const int n_items = a.hdr->n_items;
  for (int ii = blockIdx.x; ii < n_items; ii += gridDim.x) {
    const ItemDesc& it = a.items[ii];
    const int nt = it.ptiles_x * it.ptiles_y;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < nt; k += kBlock) {
      const double* p = a.partials + (size_t)(it.part_off + k) * 9;
#pragma unroll
      for (int j = 0; j < 9; j++) v[j] += p[j];
    }
    dev::block_sum9(v, lds);
    if (threadIdx.x == 0) {
      double* r = a.results + (size_t)ii * 10;
      for (int j = 0; j < 9; j++) r[j] = v[j];
      r[9] = ncc_from_moments(v);
      a.samples_host[ii] = a.samples[ii];  // drr_tiles has finished its counts
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x < kParts) {
    a.counters->drr_next[threadIdx.x * kCtrStride] = 0;
    a.counters->post_next[threadIdx.x * kCtrStride] = 0;
  }
}

// The batch's descriptors from pinned host memory to the device, and the
// per-item sample counters cleared: one launch where a copy and a fill
// would each be a stream operation of their own.
__global__ void stage_batch(const uint4* __restrict__ src, uint4* __restrict__ dst, int n16,
                            unsigned long long* __restrict__ samples, int n_items)
{
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += stride) samples[i] = 0;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  int ensure(size_t n)
  {
    if (n <= cap) return ASC_OK;
    if (p) ASC_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    const size_t want = std::max(n, (size_t)16);
    hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
    if (e != hipSuccess) return fail(ASC_ERR_NOMEM, std::string("scene scratch allocation: ") + hipGetErrorString(e));
    cap = want;
    return ASC_OK;
  }
  void release()
  {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

template <class T>
struct PinBuf {
  T* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n)
  {
    if (n <= cap) return ASC_OK;
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max(n, (size_t)16);
    // coherent (fine-grained): kernels read and write it directly, and no
    // cache keeps a line of it across batches
    ASC_HIP(hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocCoherent));
    cap = want;
    return ASC_OK;
  }
  void release()
  {
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace
}  // namespace asc

using namespace asc;

struct SceneVolume {
  bool in_use = false;
  asc_volume vol = nullptr;
  int generation_seen = -1;
  float step = 0, intensity = 1, cutoff = 0;
  Frame pivot;
};

struct SceneCamera {
  bool in_use = false;
  asc_camera cam;
  asc_image rad = nullptr;  // used by items without a frame
  asc_image bg = nullptr;
  float bg_threshold = -1.0f;
  std::vector<asc_filter> chain[2];
};

// The views of one job: cameras (scene slots) and their render size
// (Tracker's render_width/height).
struct CameraSet {
  bool in_use = false;
  std::vector<int> cams;
  int RW = 0, RH = 0;
};

// One time point of a camera set: a radiograph per camera of the set.
struct SceneFrame {
  bool in_use = false;
  int set = -1;
  std::vector<asc_image> rads;
};

struct asc_scene_s {
  asc_context ctx;
  int RW, RH;  // render size of the implicit all-camera set
  std::vector<SceneVolume> vols;
  std::vector<SceneCamera> cams;
  std::vector<CameraSet> sets;
  std::vector<SceneFrame> frames;
  bool dirty = true;
  int grid_persistent = 0;  // ray-march grid (blocks)
  bool grid_user = false;   // grid_persistent set through asc_scene_set_tuning
  int grid_post = 0;        // post-processing grid (blocks)
  int post_small = 1;       // every chain's halo <= kSmallHalo: smaller-LDS post kernel
  // device
  DevBuf<CamDev> d_cams;
  DevBuf<VolParams> d_vols;
  DevBuf<float> d_weights;
  DevBuf<unsigned char> d_batch;  // header + items
  DevBuf<float> d_drr;
  DevBuf<double> d_partials;
  DevBuf<unsigned long long> d_samples;
  DevBuf<Counters> d_counters;
  DevBuf<float> d_dbg;
  // host
  // pinned host buffers of the batch in each of the two submission slots
  PinBuf<unsigned char> h_batch[2];
  PinBuf<double> h_results[2];
  PinBuf<unsigned long long> h_samples[2];
  // items of the last batch in request order (request i, k-th camera of its
  // set at req_base[i] + k); kernel order is camera-major: korder[logical]
  std::vector<ItemDesc> items;
  std::vector<int> in_view;
  std::vector<int> korder;
  std::vector<int> req_base, req_ncam;
  int n_items = 0;
  size_t drr_total = 0;  // floats of DRR scratch used by the last batch
  // Batches submitted and not yet finished (asc_scene_submit_requests): the
  // host side of each (the fields above, moved here at submission and back
  // at finish) and an event after its last kernel, finalize_items, which
  // writes the slot's pinned results. The device buffers are shared: the
  // batches run one after another on the scene's stream.
  struct InFlight {
    bool busy = false;
    int n = 0, stride = 0;
    std::vector<ItemDesc> items;
    std::vector<int> in_view, korder, req_base, req_ncam;
    int n_items = 0;
    size_t drr_total = 0;
    bool timed = false;
    unsigned long long seq = 0;  // submission number
    hipEvent_t done = nullptr;
  } fly[2];
  // submission numbers of the last batch submitted and of the last one
  // finished: the debug images on the device are the last submitted batch's
  unsigned long long submitted_seq = 0, finished_seq = 0;
  int last_slot = 0;  // slot of the last finished batch (its h_results)
  std::vector<CameraView> cam_views;  // per batch: camera_view of each camera in it
  std::vector<char> cam_view_ok;
  std::vector<BoxCorners> vol_boxes;  // per batch: box_corners of each volume in it
  std::vector<char> vol_box_ok;
  int debug = 0;
  // graph
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_items = -1;
  // plain-path scratch (one eval at a time)
  DevBuf<float> p_bufs;
  // per-kernel timing of the last eager batch (HIP events on the scene stream)
  bool timing = false;
  int unroll = 4;  // ray samples per loop iteration (1, 2 or 4)
  int wave_shift = 3;  // a wave marches an 8 x 8 ray patch
  int gather_probe = 0;  // 1: the gather probe replaces the march (measurement)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float last_ms[3] = {0, 0, 0};
};

namespace {

void destroy_graph(asc_scene s);

// Cameras in use, in slot order: the implicit set of eval_poses/eval_manips.
std::vector<int> all_cameras(const asc_scene s)
{
  std::vector<int> c;
  for (size_t i = 0; i < s->cams.size(); i++)
    if (s->cams[i].in_use) c.push_back((int)i);
  return c;
}

int scene_upload_config(asc_scene s)
{
  for (const SceneVolume& sv : s->vols)
    if (sv.in_use && sv.generation_seen != sv.vol->generation) s->dirty = true;
  if (!s->dirty) return ASC_OK;
  std::vector<CamDev> cams(std::max<size_t>(s->cams.size(), 1));
  std::vector<float> pool;
  int max_halo = 0;
  for (size_t c = 0; c < s->cams.size(); c++) {
    const SceneCamera& sc = s->cams[c];
    CamDev& cd = cams[c];
    std::memset(&cd, 0, sizeof cd);
    if (!sc.in_use) continue;
    cd.has_bg = (sc.bg && sc.bg_threshold >= 0.0f) ? 1 : 0;  // View.cpp:173-174
    if (sc.bg) cd.bg = sc.bg->dev;
    cd.bg_threshold = sc.bg_threshold;
    for (int i = 0; i < 4; i++) cd.plane[i] = (float)sc.cam.viewport[i];
    for (int which = 0; which < 2; which++) {
      const std::vector<asc_filter>& fs = sc.chain[which];
      ChainDev& ch = cd.chain[which];
      if (fs.size() > (size_t)kMaxChain) return fail(ASC_ERR_UNSUPPORTED, "filter chain longer than 8");
      ch.n = (int)fs.size();
      ch.halo = 0;
      for (int k = 0; k < ch.n; k++) {
        ch.f[k] = fs[k];
        ch.gn[k] = 1;
        ch.woff[k] = 0;
        if (fs[k].type == ASC_FILTER_MEDIAN && fs[k].enabled)
          return fail(ASC_ERR_UNSUPPORTED, "median filter is not implemented by the reference");
        if (fs[k].type == ASC_FILTER_GAUSSIAN || fs[k].type == ASC_FILTER_SHARPEN) {
          std::vector<float> wts;
          const int n = gauss_weights(fs[k].p[0], wts);
          ch.gn[k] = n;
          ch.woff[k] = (int)pool.size();
          pool.insert(pool.end(), wts.begin(), wts.end());
        }
        ch.radius[k] = dev::filter_radius(fs[k], ch.gn[k]);
        ch.halo += ch.radius[k];
      }
      // run_chain's copies: a disabled stage, or a Gaussian / Sharpen of size <= 1
      int n_work = 0, n_sobel = 0;
      for (int k = 0; k < ch.n; k++) {
        const bool copy = !fs[k].enabled ||
                          ((fs[k].type == ASC_FILTER_GAUSSIAN || fs[k].type == ASC_FILTER_SHARPEN) && ch.gn[k] <= 1);
        if (!copy) {
          ++n_work;
          n_sobel += fs[k].type == ASC_FILTER_SOBEL;
        }
      }
      ch.shape = n_work == 0 ? kChainCopies : (n_work == 1 && n_sobel == 1) ? kChainSobel : kChainGeneral;
      if (ch.halo > kMaxHalo) return fail(ASC_ERR_UNSUPPORTED, "filter chain halo exceeds 12 pixels");
      max_halo = std::max(max_halo, ch.halo);
    }
  }
  const int small = max_halo <= kSmallHalo ? 1 : 0;
  if (small != s->post_small) {
    s->post_small = small;
    s->grid_post = 0;
    destroy_graph(s);
  }
  std::vector<VolParams> vols(std::max<size_t>(s->vols.size(), 1));
  for (size_t i = 0; i < s->vols.size(); i++) {
    SceneVolume& sv = s->vols[i];
    std::memset(&vols[i], 0, sizeof(VolParams));
    if (!sv.in_use) continue;
    sv.generation_seen = sv.vol->generation;
    for (int a = 0; a < 3; a++) vols[i].v[a] = sv.vol->odev[a];
    vols[i].step = sv.step;
    vols[i].intensity = sv.intensity;
    vols[i].cutoff = sv.cutoff;
    vols[i].step_vscale = sv.step * sv.vol->dev.vscale;
  }
  if (pool.empty()) pool.push_back(0.0f);
  const void* before[3] = {s->d_cams.p, s->d_vols.p, s->d_weights.p};
  int rc;
  if ((rc = s->d_cams.ensure(cams.size()))) return rc;
  if ((rc = s->d_vols.ensure(vols.size()))) return rc;
  if ((rc = s->d_weights.ensure(pool.size()))) return rc;
  const void* after[3] = {s->d_cams.p, s->d_vols.p, s->d_weights.p};
  if (std::memcmp(before, after, sizeof before) != 0) destroy_graph(s);
  ASC_HIP(hipStreamSynchronize(s->ctx->stream));
  ASC_HIP(hipMemcpy(s->d_cams.p, cams.data(), sizeof(CamDev) * cams.size(), hipMemcpyHostToDevice));
  ASC_HIP(hipMemcpy(s->d_vols.p, vols.data(), sizeof(VolParams) * vols.size(), hipMemcpyHostToDevice));
  ASC_HIP(hipMemcpy(s->d_weights.p, pool.data(), sizeof(float) * pool.size(), hipMemcpyHostToDevice));
  s->dirty = false;
  return ASC_OK;
}

void destroy_graph(asc_scene s)
{
  if (s->graph_exec) hipGraphExecDestroy(s->graph_exec);
  if (s->graph) hipGraphDestroy(s->graph);
  s->graph_exec = nullptr;
  s->graph = nullptr;
  s->graph_items = -1;
}

// Issue the batch (upload, three kernels, download) on the context stream,
// with grid_drr and grid_post blocks for the two persistent kernels (any
// grid of at least one block drains the whole tile list).
int issue_batch(asc_scene s, int slot, size_t batch_bytes, int grid_drr, int grid_post)
{
  hipStream_t st = s->ctx->stream;
  const int n_items = s->n_items;
  BatchArgs a;
  a.hdr = (const BatchHeader*)s->d_batch.p;
  a.items = (const ItemDesc*)(s->d_batch.p + sizeof(BatchHeader));
  a.cams = s->d_cams.p;
  a.vols = s->d_vols.p;
  a.weights = s->d_weights.p;
  a.drr = s->d_drr.p;
  a.partials = s->d_partials.p;
  a.results = s->h_results[slot].p;
  a.samples = s->d_samples.p;
  a.samples_host = s->h_samples[slot].p;
  a.counters = s->d_counters.p;
  const size_t plane = s->drr_total;
  a.dbg_drr = s->debug ? s->d_dbg.p : nullptr;
  a.dbg_rad = s->debug ? s->d_dbg.p + plane : nullptr;
  a.dbg_mask = s->debug ? s->d_dbg.p + 2 * plane : nullptr;
  a.wave_shift = s->wave_shift;
  // batch_bytes: a multiple of 16 (both buffers hold that much)
  const int n16 = (int)(batch_bytes / 16);
  stage_batch<<<std::min((std::max(n16, n_items) + 255) / 256, 64), 256, 0, st>>>(
    (const uint4*)s->h_batch[slot].p, (uint4*)s->d_batch.p, n16, s->d_samples.p, n_items);
  // the timing events are the scene's: a batch queued behind another does
  // not record them
  const bool timed = s->timing && s->ev[0] && !s->fly[0].busy && !s->fly[1].busy;
  if (timed) ASC_HIP(hipEventRecord(s->ev[0], st));
  if (s->gather_probe)
    drr_tiles<4, true><<<grid_drr, kBlock, 0, st>>>(a);
  else
    switch (s->unroll) {
      case 1: drr_tiles<1><<<grid_drr, kBlock, 0, st>>>(a); break;
      case 4: drr_tiles<4><<<grid_drr, kBlock, 0, st>>>(a); break;
      default: drr_tiles<2><<<grid_drr, kBlock, 0, st>>>(a); break;
    }
  if (timed) ASC_HIP(hipEventRecord(s->ev[1], st));
  if (s->post_small)
    post_tiles<kSmallHalo><<<grid_post, kBlock, 0, st>>>(a);
  else
    post_tiles<kMaxHalo><<<grid_post, kBlock, 0, st>>>(a);
  if (timed) ASC_HIP(hipEventRecord(s->ev[2], st));
  finalize_items<<<std::min(n_items, 4096), kBlock, 0, st>>>(a);
  if (timed) ASC_HIP(hipEventRecord(s->ev[3], st));
  ASC_HIP(hipGetLastError());
  // finalize_items writes the results and the per-item sample counts into
  // the slot's pinned buffers: no copy follows
  return ASC_OK;
}

// One request of a batch: a pose of a volume seen by a camera set (-1: all
// cameras in use) at a frame (-1: each camera's own radiograph).
struct Request {
  int vol, set, frame;
  const double* pose;
};

// Geometry for every (request, camera) item, then the fused launches of
// the batch in a free submission slot, queued on the scene's stream behind
// any batch already submitted; the batch's host side moves to the slot
// (finish_requests takes it back). `graph`: the synchronous form, which
// may replay a recorded hipGraph (slot 0 only: the graph holds its
// buffers).
int submit_requests(asc_scene s, const std::vector<Request>& reqs, int stride, bool graph, int* slot_out)
{
  // the calling thread may not have selected the scene's device yet
  ASC_HIP(hipSetDevice(s->ctx->device));
  const bool other_busy = s->fly[0].busy || s->fly[1].busy;
  ASC_CHECK(!(s->fly[0].busy && s->fly[1].busy), ASC_ERR_INVALID, "two batches are already submitted");
  ASC_CHECK(!(graph && s->use_graph && other_busy), ASC_ERR_INVALID,
            "a recorded graph runs with no other batch submitted");
  ASC_CHECK(graph || !s->use_graph, ASC_ERR_UNSUPPORTED, "asynchronous batches do not replay recorded graphs");
  const int slot = s->fly[0].busy ? 1 : 0;
  hipStream_t st = s->ctx->stream;
  // device buffers that grow are reallocated: let a submitted batch that
  // uses them finish first
  bool dirty = s->dirty;
  for (const SceneVolume& sv : s->vols)
    if (sv.in_use && sv.generation_seen != sv.vol->generation) dirty = true;
  if (dirty && other_busy) ASC_HIP(hipStreamSynchronize(st));
  int rc = scene_upload_config(s);
  if (rc) return rc;
  const int n = (int)reqs.size();
  const std::vector<int> all = all_cameras(s);
  // items in request order
  s->req_base.assign(n, 0);
  s->req_ncam.assign(n, 0);
  int n_items = 0;
  for (int i = 0; i < n; i++) {
    const Request& q = reqs[i];
    ASC_CHECK(q.vol >= 0 && q.vol < (int)s->vols.size() && s->vols[q.vol].in_use, ASC_ERR_INVALID,
              "volume index out of range");
    int nc;
    if (q.set < 0) {
      nc = (int)all.size();
    } else {
      ASC_CHECK(q.set < (int)s->sets.size() && s->sets[q.set].in_use, ASC_ERR_INVALID, "bad camera set");
      nc = (int)s->sets[q.set].cams.size();
    }
    ASC_CHECK(nc > 0, ASC_ERR_INVALID, "scene has no cameras");
    ASC_CHECK(nc <= stride, ASC_ERR_INVALID, "more cameras than the per-camera output holds");
    if (q.frame >= 0)
      ASC_CHECK(q.frame < (int)s->frames.size() && s->frames[q.frame].in_use && s->frames[q.frame].set == q.set,
                ASC_ERR_INVALID, "frame does not belong to the request's camera set");
    s->req_base[i] = n_items;
    s->req_ncam[i] = nc;
    n_items += nc;
  }
  s->n_items = n_items;
  s->items.resize(n_items);
  s->in_view.assign(n_items, 0);
  // each camera's inverse frame and film-plane bounds once per batch
  std::vector<CameraView>& views = s->cam_views;
  std::vector<char>& have_view = s->cam_view_ok;
  views.resize(s->cams.size());
  have_view.assign(s->cams.size(), 0);
  // and each volume's box corners
  std::vector<BoxCorners>& boxes = s->vol_boxes;
  std::vector<char>& have_box = s->vol_box_ok;
  boxes.resize(s->vols.size());
  have_box.assign(s->vols.size(), 0);
  for (int i = 0; i < n; i++) {
    const Request& q = reqs[i];
    const asc_volume_desc& d = s->vols[q.vol].vol->desc;
    const Frame pose = frame_from_xyzypr(q.pose);
    const std::vector<int>& cl = q.set < 0 ? all : s->sets[q.set].cams;
    const int RW = q.set < 0 ? s->RW : s->sets[q.set].RW, RH = q.set < 0 ? s->RH : s->sets[q.set].RH;
    for (int k = 0; k < s->req_ncam[i]; k++) {
      const int ii = s->req_base[i] + k;
      const int c = cl[k];
      ItemDesc& it = s->items[ii];
      std::memset(&it, 0, sizeof it);
      asc_image rad = q.frame >= 0 ? s->frames[q.frame].rads[k] : s->cams[c].rad;
      ASC_CHECK(rad != nullptr, ASC_ERR_INVALID, "camera has no radiograph");
      it.rad = rad->dev;
      double vp[4];
      unsigned sz[2] = {0, 0};
      if (!have_view[c]) {
        views[c] = camera_view(s->cams[c].cam);
        have_view[c] = 1;
      }
      if (!have_box[q.vol]) {
        boxes[q.vol] = box_corners(d.inv_trans, d.inv_scale);
        have_box[q.vol] = 1;
      }
      const int inview =
        view_geometry(s->cams[c].cam, views[c], pose, boxes[q.vol], d.inv_scale, d.inv_trans, RW, RH, it.imv, vp, sz);
      it.vol = q.vol;
      it.cam = c;
      it.axis = s->vols[q.vol].vol->dev.layout == ASC_LAYOUT_BRICK ? ray_axis(s->vols[q.vol].vol->dev, it.imv) : 0;
      if (inview) {
        if (sz[0] > (unsigned)RW || sz[1] > (unsigned)RH)
          return fail(ASC_ERR_TOO_LARGE, "Rendered image is larger than the viewport buffer");
        for (int m = 0; m < 4; m++) it.vp[m] = (float)vp[m];
        it.w = (int)sz[0];
        it.h = (int)sz[1];
        it.tiles_x = (it.w + kTile - 1) / kTile;
        it.tiles_y = (it.h + kTile - 1) / kTile;
        it.ptiles_x = (it.w + kPTile - 1) / kPTile;
        it.ptiles_y = (it.h + kPTile - 1) / kPTile;
        s->in_view[ii] = 1;
      }
    }
  }
  // camera-major kernel order; compact DRR scratch and partial offsets
  std::vector<int> acams;
  {
    std::vector<int> count(s->cams.size(), 0);
    for (int ii = 0; ii < n_items; ii++) count[s->items[ii].cam]++;
    for (size_t c = 0; c < count.size(); c++)
      if (count[c]) acams.push_back((int)c);
  }
  ASC_CHECK(acams.size() <= (size_t)kMaxCams, ASC_ERR_UNSUPPORTED, "more than 64 cameras in one batch");
  std::vector<int> cam_slot(s->cams.size(), -1);
  for (size_t a = 0; a < acams.size(); a++) cam_slot[acams[a]] = (int)a;
  BatchHeader hdr;
  std::memset(&hdr, 0, sizeof hdr);
  hdr.n_items = n_items;
  hdr.debug = s->debug;
  hdr.n_acams = (int)acams.size();
  for (int ii = 0; ii < n_items; ii++) hdr.cam_items[cam_slot[s->items[ii].cam]]++;
  for (int a = 1; a < hdr.n_acams; a++) hdr.item_base[a] = hdr.item_base[a - 1] + hdr.cam_items[a - 1];
  s->korder.assign(n_items, 0);
  {
    std::vector<int> fill(hdr.n_acams, 0);
    for (int ii = 0; ii < n_items; ii++) {
      const int a = cam_slot[s->items[ii].cam];
      s->korder[ii] = hdr.item_base[a] + fill[a]++;
    }
  }
  std::vector<ItemDesc> kitems(n_items);
  size_t drr_total = 0, part_total = 0;
  for (int ii = 0; ii < n_items; ii++) {
    ItemDesc it = s->items[ii];
    it.drr_off = (long long)drr_total;
    it.part_off = (int)part_total;
    drr_total += (size_t)it.w * it.h;
    part_total += (size_t)it.ptiles_x * it.ptiles_y;
    it.tile_base = 0;
    s->items[ii].drr_off = it.drr_off;
    kitems[s->korder[ii]] = it;
  }
  ASC_CHECK(part_total < (size_t)1 << 31, ASC_ERR_TOO_LARGE, "batch has too many tiles");
  s->drr_total = std::max<size_t>(drr_total, 1);
  for (int which = 0; which < 2; which++) {
    TileEnum& en = which ? hdr.post : hdr.drr;
    long long base = 0;
    for (int a = 0; a < hdr.n_acams; a++) {
      int tx = 0, ty = 0;
      for (int e = 0; e < hdr.cam_items[a]; e++) {
        const ItemDesc& it = kitems[hdr.item_base[a] + e];
        tx = std::max(tx, which ? it.ptiles_x : it.tiles_x);
        ty = std::max(ty, which ? it.ptiles_y : it.tiles_y);
      }
      en.cam_base[a] = (int)base;
      en.tiles_x[a] = std::max(tx, 1);
      en.tiles_y[a] = ty;
      base += (long long)tx * ty * hdr.cam_items[a];
    }
    ASC_CHECK(base < (1LL << 31) - (1LL << 20), ASC_ERR_TOO_LARGE, "batch has too many tiles");
    en.total = (int)base;
  }
  // buffers
  // (rounded up to whole 16-byte words for stage_batch)
  const size_t batch_bytes = (sizeof(BatchHeader) + sizeof(ItemDesc) * (size_t)n_items + 15) & ~(size_t)15;
  // a recorded graph names these buffers (slot 0's host buffers included)
  const void* before[8] = {s->d_batch.p, s->d_drr.p, s->d_partials.p, s->h_batch[0].p, s->h_results[0].p,
                           s->h_samples[0].p, s->d_samples.p, s->d_dbg.p};
  if (other_busy && (batch_bytes > s->d_batch.cap || s->drr_total > s->d_drr.cap ||
                     std::max<size_t>(part_total, 1) * 9 > s->d_partials.cap || (size_t)n_items > s->d_samples.cap ||
                     (s->debug && s->drr_total * 3 > s->d_dbg.cap)))
    ASC_HIP(hipStreamSynchronize(st));
  if ((rc = s->d_batch.ensure(batch_bytes))) return rc;
  if ((rc = s->h_batch[slot].ensure(batch_bytes))) return rc;
  if ((rc = s->d_drr.ensure(s->drr_total))) return rc;
  if ((rc = s->d_partials.ensure(std::max<size_t>(part_total, 1) * 9))) return rc;
  if ((rc = s->h_results[slot].ensure((size_t)n_items * 10))) return rc;
  if ((rc = s->d_samples.ensure((size_t)n_items))) return rc;
  if ((rc = s->h_samples[slot].ensure((size_t)n_items))) return rc;
  if (!s->d_counters.p) {
    if ((rc = s->d_counters.ensure(1))) return rc;
    ASC_HIP(hipMemset(s->d_counters.p, 0, sizeof(Counters)));
  }
  if (s->debug && (rc = s->d_dbg.ensure(s->drr_total * 3))) return rc;
  const void* after[8] = {s->d_batch.p, s->d_drr.p, s->d_partials.p, s->h_batch[0].p, s->h_results[0].p,
                          s->h_samples[0].p, s->d_samples.p, s->d_dbg.p};
  if (std::memcmp(before, after, sizeof before) != 0) destroy_graph(s);
  if (s->grid_persistent == 0 || s->grid_post == 0) {
    hipDeviceProp_t prop;
    ASC_HIP(hipGetDeviceProperties(&prop, s->ctx->device));
    if (s->grid_persistent == 0) {
      int per_cu = 0;
      if (s->unroll == 4)
        ASC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, drr_tiles<4>, kBlock, 0));
      else if (s->unroll == 2)
        ASC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, drr_tiles<2>, kBlock, 0));
      else
        ASC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, drr_tiles<1>, kBlock, 0));
      s->grid_persistent = prop.multiProcessorCount * std::max(per_cu, 1);
    }
    if (s->grid_post == 0) {
      int per_cu = 0;
      if (s->post_small)
        ASC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, post_tiles<kSmallHalo>, kBlock, 0));
      else
        ASC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, post_tiles<kMaxHalo>, kBlock, 0));
      s->grid_post = prop.multiProcessorCount * std::max(per_cu, 1);
    }
  }
  std::memcpy(s->h_batch[slot].p, &hdr, sizeof hdr);
  std::memcpy(s->h_batch[slot].p + sizeof hdr, kitems.data(), sizeof(ItemDesc) * n_items);

  if (graph && s->use_graph) {
    if (!s->graph_exec || s->graph_items != n_items) {
      destroy_graph(s);
      ASC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      // a recorded graph replays batches of other tile counts: full grids
      rc = issue_batch(s, slot, batch_bytes, s->grid_persistent, s->grid_post);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) return rc;
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      s->graph = g;
      ASC_HIP(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
      s->graph_items = n_items;
    }
    ASC_HIP(hipGraphLaunch(s->graph_exec, st));
  } else {
    // a small batch (a simplex step) launches only as many blocks as it has
    // tiles, rounded up to the 8 partitions of the tile list
    const int gd = (int)std::min<long long>(s->grid_persistent, ((long long)hdr.drr.total + 7) / 8 * 8);
    const int gp = (int)std::min<long long>(s->grid_post, ((long long)hdr.post.total + 7) / 8 * 8);
    if ((rc = issue_batch(s, slot, batch_bytes, std::max(gd, 8), std::max(gp, 8)))) return rc;
  }
  auto& f = s->fly[slot];
  if (!f.done) ASC_HIP(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
  ASC_HIP(hipEventRecord(f.done, st));
  f.busy = true;
  f.n = (int)reqs.size();
  f.stride = stride;
  f.timed = !other_busy;  // issue_batch recorded the timing events
  f.seq = ++s->submitted_seq;
  std::swap(f.items, s->items);
  std::swap(f.in_view, s->in_view);
  std::swap(f.korder, s->korder);
  std::swap(f.req_base, s->req_base);
  std::swap(f.req_ncam, s->req_ncam);
  f.n_items = s->n_items;
  f.drr_total = s->drr_total;
  *slot_out = slot;
  return ASC_OK;
}

// Wait for the batch of `slot`, take its host side back as the scene's last
// batch and turn its moments into costs.
int finish_requests(asc_scene s, int slot, double* costs, double* per_camera)
{
  ASC_CHECK(slot == 0 || slot == 1, ASC_ERR_INVALID, "bad batch ticket");
  auto& f = s->fly[slot];
  ASC_CHECK(f.busy, ASC_ERR_INVALID, "no batch submitted under this ticket");
  ASC_HIP(hipSetDevice(s->ctx->device));
  f.busy = false;
  ASC_HIP(hipEventSynchronize(f.done));
  std::swap(f.items, s->items);
  std::swap(f.in_view, s->in_view);
  std::swap(f.korder, s->korder);
  std::swap(f.req_base, s->req_base);
  std::swap(f.req_ncam, s->req_ncam);
  s->n_items = f.n_items;
  s->drr_total = f.drr_total;
  s->last_slot = slot;
  s->finished_seq = f.seq;
  if (s->timing && s->ev[0] && !s->use_graph && f.timed)
    for (int k = 0; k < 3; k++) ASC_HIP(hipEventElapsedTime(&s->last_ms[k], s->ev[k], s->ev[k + 1]));
  const int n = f.n, stride = f.stride;
  const double* res = s->h_results[slot].p;
  // Tracker.cpp:614-618 (9999 when not in view), :769 (1 - NCC),
  // :814-826 (sum over cameras; negative totals become 9999).
  for (int i = 0; i < n; i++) {
    double total = 0.0;
    for (int k = 0; k < s->req_ncam[i]; k++) {
      const int ii = s->req_base[i] + k;
      const int ki = s->korder[ii];
      const double corr = s->in_view[ii] ? 1.0 - res[(size_t)ki * 10 + 9] : 9999.0;
      if (per_camera) per_camera[(size_t)i * stride + k] = corr;
      total = k == 0 ? corr : total + corr;
    }
    if (total < 0) total = 9999;
    if (costs) costs[i] = total;
  }
  return ASC_OK;
}

// Geometry, fused launches and costs of a batch, waiting for it.
// per_camera: n x stride values (camera k of request i at i*stride + k).
int eval_requests(asc_scene s, const std::vector<Request>& reqs, double* costs, double* per_camera, int stride)
{
  int slot = 0;
  int rc = submit_requests(s, reqs, stride, true, &slot);
  if (rc) return rc;
  return finish_requests(s, slot, costs, per_camera);
}

// The all-camera form used by eval_poses / eval_manips.
int eval_batch(asc_scene s, const int* volumes, const double* poses, int n, double* costs, double* per_camera)
{
  std::vector<Request> reqs(n);
  for (int i = 0; i < n; i++) reqs[i] = Request{volumes ? volumes[i] : 0, -1, -1, poses + 6 * i};
  const int C = (int)all_cameras(s).size();
  ASC_CHECK(C > 0, ASC_ERR_INVALID, "scene has no cameras");
  return eval_requests(s, reqs, costs, per_camera, C);
}

// First free slot of a scene table (reused after a remove), or a new one.
template <class T>
int take_slot(std::vector<T>& v)
{
  for (size_t i = 0; i < v.size(); i++)
    if (!v[i].in_use) {
      v[i] = T();
      return (int)i;
    }
  v.emplace_back();
  return (int)v.size() - 1;
}

}  // namespace

extern "C" {

ASC_API int asc_scene_create(asc_context ctx, int render_width, int render_height, asc_scene* out)
{
  ASC_CHECK(ctx && out, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(render_width > 0 && render_height > 0, ASC_ERR_INVALID, "bad render size");
  asc_scene s = new asc_scene_s();
  s->ctx = ctx;
  s->RW = render_width;
  s->RH = render_height;
  *out = s;
  return ASC_OK;
}

ASC_API int asc_scene_destroy(asc_scene s)
{
  if (!s) return ASC_OK;
  hipStreamSynchronize(s->ctx->stream);
  destroy_graph(s);
  for (hipEvent_t& e : s->ev)
    if (e) hipEventDestroy(e);
  s->d_cams.release();
  s->d_vols.release();
  s->d_weights.release();
  s->d_batch.release();
  s->d_drr.release();
  s->d_partials.release();
  s->d_samples.release();
  s->d_counters.release();
  s->d_dbg.release();
  s->p_bufs.release();
  for (int k = 0; k < 2; k++) {
    s->h_batch[k].release();
    s->h_results[k].release();
    s->h_samples[k].release();
    if (s->fly[k].done) hipEventDestroy(s->fly[k].done);
  }
  delete s;
  return ASC_OK;
}

ASC_API int asc_scene_add_volume(asc_scene s, asc_volume vol, float step, float intensity, float cutoff,
                                 const double pivot[12], int* index)
{
  ASC_CHECK(s && vol, ASC_ERR_INVALID, "bad argument");
  const int i = take_slot(s->vols);
  SceneVolume& sv = s->vols[i];
  sv.in_use = true;
  sv.vol = vol;
  sv.step = step;
  sv.intensity = intensity;
  sv.cutoff = cutoff;
  sv.pivot = pivot ? frame_load(pivot) : frame_identity();
  if (index) *index = i;
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_remove_volume(asc_scene s, int volume)
{
  ASC_CHECK(s && volume >= 0 && volume < (int)s->vols.size() && s->vols[volume].in_use, ASC_ERR_INVALID,
            "bad volume index");
  s->vols[volume] = SceneVolume();
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_volume_pivot(asc_scene s, int volume, double pivot[12])
{
  ASC_CHECK(s && pivot && volume >= 0 && volume < (int)s->vols.size() && s->vols[volume].in_use, ASC_ERR_INVALID,
            "bad volume index");
  frame_store(s->vols[volume].pivot, pivot);
  return ASC_OK;
}

ASC_API int asc_scene_set_drr_params(asc_scene s, int volume, float step, float intensity, float cutoff)
{
  ASC_CHECK(s && volume >= 0 && volume < (int)s->vols.size(), ASC_ERR_INVALID, "bad volume index");
  s->vols[volume].step = step;
  s->vols[volume].intensity = intensity;
  s->vols[volume].cutoff = cutoff;
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_add_camera(asc_scene s, const asc_camera* cam, asc_image rad, int* index)
{
  ASC_CHECK(s && cam, ASC_ERR_INVALID, "bad argument");
  const int i = take_slot(s->cams);
  ASC_CHECK(i < kMaxCams, ASC_ERR_UNSUPPORTED, "at most 64 cameras per scene");
  SceneCamera& sc = s->cams[i];
  sc.in_use = true;
  sc.cam = *cam;
  sc.rad = rad;
  if (index) *index = i;
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_remove_camera(asc_scene s, int camera)
{
  ASC_CHECK(s && camera >= 0 && camera < (int)s->cams.size() && s->cams[camera].in_use, ASC_ERR_INVALID,
            "bad camera index");
  for (const CameraSet& cs : s->sets)
    if (cs.in_use)
      for (int c : cs.cams) ASC_CHECK(c != camera, ASC_ERR_INVALID, "camera is used by a camera set");
  s->cams[camera] = SceneCamera();
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_add_camera_set(asc_scene s, const int* cameras, int count, int render_width,
                                     int render_height, int* set)
{
  ASC_CHECK(s && cameras && set, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(count > 0 && count <= kMaxSetCams, ASC_ERR_INVALID, "a camera set has 1 to 8 cameras");
  ASC_CHECK(render_width > 0 && render_height > 0, ASC_ERR_INVALID, "bad render size");
  for (int k = 0; k < count; k++)
    ASC_CHECK(cameras[k] >= 0 && cameras[k] < (int)s->cams.size() && s->cams[cameras[k]].in_use, ASC_ERR_INVALID,
              "bad camera index");
  const int i = take_slot(s->sets);
  CameraSet& cs = s->sets[i];
  cs.in_use = true;
  cs.cams.assign(cameras, cameras + count);
  cs.RW = render_width;
  cs.RH = render_height;
  *set = i;
  return ASC_OK;
}

ASC_API int asc_scene_remove_camera_set(asc_scene s, int set)
{
  ASC_CHECK(s && set >= 0 && set < (int)s->sets.size() && s->sets[set].in_use, ASC_ERR_INVALID, "bad camera set");
  for (SceneFrame& f : s->frames)
    if (f.in_use && f.set == set) f = SceneFrame();
  s->sets[set] = CameraSet();
  return ASC_OK;
}

ASC_API int asc_scene_add_frame(asc_scene s, int set, const asc_image* rads, int* frame)
{
  ASC_CHECK(s && rads && frame, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(set >= 0 && set < (int)s->sets.size() && s->sets[set].in_use, ASC_ERR_INVALID, "bad camera set");
  const size_t nc = s->sets[set].cams.size();
  for (size_t k = 0; k < nc; k++) ASC_CHECK(rads[k] != nullptr, ASC_ERR_INVALID, "null radiograph");
  const int i = take_slot(s->frames);
  SceneFrame& f = s->frames[i];
  f.in_use = true;
  f.set = set;
  f.rads.assign(rads, rads + nc);
  *frame = i;
  return ASC_OK;
}

ASC_API int asc_scene_remove_frame(asc_scene s, int frame)
{
  ASC_CHECK(s && frame >= 0 && frame < (int)s->frames.size() && s->frames[frame].in_use, ASC_ERR_INVALID,
            "bad frame");
  s->frames[frame] = SceneFrame();
  return ASC_OK;
}

ASC_API int asc_scene_set_rad(asc_scene s, int camera, asc_image rad)
{
  ASC_CHECK(s && camera >= 0 && camera < (int)s->cams.size() && s->cams[camera].in_use && rad, ASC_ERR_INVALID,
            "bad argument");
  s->cams[camera].rad = rad;
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_set_background(asc_scene s, int camera, asc_image bg, float threshold)
{
  ASC_CHECK(s && camera >= 0 && camera < (int)s->cams.size() && s->cams[camera].in_use, ASC_ERR_INVALID,
            "bad argument");
  ASC_CHECK(!bg || bg->src_bps == 32, ASC_ERR_UNSUPPORTED, "background image must be float32 (CL_FLOAT)");
  s->cams[camera].bg = bg;
  s->cams[camera].bg_threshold = threshold;
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_set_filters(asc_scene s, int camera, int which, const asc_filter* filters, int count)
{
  ASC_CHECK(s && camera >= 0 && camera < (int)s->cams.size() && s->cams[camera].in_use && (which == 0 || which == 1),
            ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(count >= 0 && (count == 0 || filters), ASC_ERR_INVALID, "bad filter list");
  s->cams[camera].chain[which].assign(filters, filters + count);
  s->dirty = true;
  return ASC_OK;
}

ASC_API int asc_scene_eval_poses(asc_scene s, const int* volumes, const double* xyzypr, int n, double* costs,
                                 double* per_camera)
{
  ASC_CHECK(s && xyzypr && n > 0, ASC_ERR_INVALID, "bad argument");
  return eval_batch(s, volumes, xyzypr, n, costs, per_camera);
}

ASC_API int asc_scene_autotune(asc_scene s, int n, const int* volumes, const double* poses, int reps,
                               asc_tune_result* result)
{
  ASC_CHECK(s && poses && n > 0 && reps > 0, ASC_ERR_INVALID, "bad argument");
  // the candidates' HIP-event timings need the stream to themselves
  ASC_CHECK(!s->fly[0].busy && !s->fly[1].busy, ASC_ERR_INVALID, "finish the submitted batches first");
  const bool timing0 = s->timing, graph0 = s->use_graph;
  if (!s->ev[0])
    for (hipEvent_t& e : s->ev) ASC_HIP(hipEventCreate(&e));
  s->timing = true;
  s->use_graph = false;
  destroy_graph(s);
  struct Cand {
    int unroll, shift, gmul;
    int grid = 0;  // persistent grid of this variant
    double ms = 0.0;
    std::vector<double> t;
  };
  std::vector<Cand> cands;
  for (int unroll : {4, 2})
    for (int shift : {3, 4, 2})
      for (int gmul : {1, 2}) cands.push_back(Cand{unroll, shift, gmul});
  int rc = ASC_OK;
  auto use = [&](const Cand& c) {
    s->unroll = c.unroll;
    s->wave_shift = c.shift;
    s->grid_user = c.gmul > 1;
    s->grid_persistent = c.grid;  // 0: this variant's occupancy grid, derived by the next batch
  };
  // a first round warms caches and clocks and derives each variant's grid;
  // the timed rounds then visit every variant once per round, so a drift in
  // clocks or cache state during tuning falls on all of them alike
  for (Cand& c : cands) {
    use(c);
    if ((rc = eval_batch(s, volumes, poses, n, nullptr, nullptr))) break;
    c.grid = s->grid_persistent * c.gmul;
  }
  for (int r = 0; r < reps && rc == ASC_OK; r++)
    for (Cand& c : cands) {
      use(c);
      if ((rc = eval_batch(s, volumes, poses, n, nullptr, nullptr))) break;
      c.t.push_back(s->last_ms[0]);
    }
  if (rc == ASC_OK)
    for (Cand& c : cands) {
      std::sort(c.t.begin(), c.t.end());
      c.ms = c.t[c.t.size() / 2];
    }
  s->timing = timing0;
  s->use_graph = graph0;
  if (rc) return rc;
  const Cand* best = &cands[0];
  for (const Cand& c : cands)
    if (c.ms < best->ms) best = &c;
  use(*best);
  destroy_graph(s);
  if (result) {
    result->unroll = best->unroll;
    result->wave_patch = 1 << best->shift;
    result->grid = best->gmul > 1 ? best->grid : 0;
    result->best_ms = best->ms;
    result->default_ms = cands[0].ms;  // unroll 4, 8 x 8 patch, occupancy grid
    result->candidates = (int)cands.size();
  }
  return ASC_OK;
}

ASC_API int asc_scene_eval_requests(asc_scene s, int n, const int* volumes, const int* sets, const int* frames,
                                    const double* xyzypr, double* costs, double* per_camera)
{
  ASC_CHECK(s && volumes && sets && xyzypr && n > 0, ASC_ERR_INVALID, "bad argument");
  std::vector<Request> reqs(n);
  for (int i = 0; i < n; i++) reqs[i] = Request{volumes[i], sets[i], frames ? frames[i] : -1, xyzypr + 6 * i};
  return eval_requests(s, reqs, costs, per_camera, kMaxSetCams);
}

ASC_API int asc_scene_submit_requests(asc_scene s, int n, const int* volumes, const int* sets, const int* frames,
                                      const double* xyzypr, int* ticket)
{
  ASC_CHECK(s && volumes && sets && xyzypr && ticket && n > 0, ASC_ERR_INVALID, "bad argument");
  std::vector<Request> reqs(n);
  for (int i = 0; i < n; i++) reqs[i] = Request{volumes[i], sets[i], frames ? frames[i] : -1, xyzypr + 6 * i};
  return submit_requests(s, reqs, kMaxSetCams, false, ticket);
}

ASC_API int asc_scene_wait_requests(asc_scene s, int ticket)
{
  ASC_CHECK(s && (ticket == 0 || ticket == 1), ASC_ERR_INVALID, "bad argument");
  // reads only the slot's event, which submit created before handing out
  // the ticket and which no call replaces while the slot is submitted
  ASC_CHECK(s->fly[ticket].busy, ASC_ERR_INVALID, "no batch submitted under this ticket");
  ASC_HIP(hipSetDevice(s->ctx->device));
  ASC_HIP(hipEventSynchronize(s->fly[ticket].done));
  return ASC_OK;
}

ASC_API int asc_scene_finish_requests(asc_scene s, int ticket, double* costs, double* per_camera)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  return finish_requests(s, ticket, costs, per_camera);
}

ASC_API int asc_scene_scratch_bytes(asc_scene s, size_t* bytes)
{
  ASC_CHECK(s && bytes, ASC_ERR_INVALID, "bad argument");
  *bytes = s->d_batch.cap + s->d_drr.cap * sizeof(float) + s->d_partials.cap * sizeof(double) +
           s->d_samples.cap * sizeof(unsigned long long) +
           s->d_dbg.cap * sizeof(float) + s->p_bufs.cap * sizeof(float);
  return ASC_OK;
}

ASC_API int asc_scene_eval_manips(asc_scene s, const int* volumes, const float* stores, const double* manips, int n,
                                  double* costs, double* per_camera)
{
  ASC_CHECK(s && stores && manips && n > 0, ASC_ERR_INVALID, "bad argument");
  std::vector<double> poses((size_t)n * 6);
  for (int i = 0; i < n; i++) {
    const int vi = volumes ? volumes[i] : 0;
    ASC_CHECK(vi >= 0 && vi < (int)s->vols.size() && s->vols[vi].in_use, ASC_ERR_INVALID,
              "volume index out of range");
    PoseStore ps;
    std::memcpy(&ps, stores + 7 * i, sizeof ps);
    compose_pose(ps, s->vols[vi].pivot, manips + 6 * i, poses.data() + 6 * i);
  }
  return eval_batch(s, volumes, poses.data(), n, costs, per_camera);
}

ASC_API int asc_scene_use_graph(asc_scene s, int enable)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  s->use_graph = enable != 0;
  if (!s->use_graph) destroy_graph(s);
  return ASC_OK;
}

ASC_API int asc_scene_set_debug(asc_scene s, int enable)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  s->debug = enable != 0;
  destroy_graph(s);
  return ASC_OK;
}

ASC_API int asc_scene_set_tuning(asc_scene s, int key, int value)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  switch (key) {
    case ASC_TUNE_UNROLL:
      ASC_CHECK(value == 1 || value == 2 || value == 4, ASC_ERR_INVALID, "unroll must be 1, 2 or 4");
      s->unroll = value;
      if (!s->grid_user) s->grid_persistent = 0;  // re-derived from this variant's occupancy
      break;
    case ASC_TUNE_GRID:
      ASC_CHECK(value >= 0, ASC_ERR_INVALID, "grid must be >= 0 (0 = one wave of blocks per CU slot)");
      s->grid_persistent = value - value % 8;
      s->grid_user = s->grid_persistent > 0;
      break;
    case ASC_TUNE_WAVE_PATCH:
      ASC_CHECK(value == 4 || value == 8 || value == 16, ASC_ERR_INVALID, "wave patch width must be 4, 8 or 16");
      s->wave_shift = value == 4 ? 2 : value == 8 ? 3 : 4;
      break;
    case ASC_TUNE_GATHER_PROBE:
      ASC_CHECK(value == 0 || value == 1, ASC_ERR_INVALID, "gather probe is 0 or 1");
      s->gather_probe = value;
      break;
    default: return fail(ASC_ERR_INVALID, "unknown tuning key");
  }
  destroy_graph(s);
  return ASC_OK;
}

ASC_API int asc_scene_set_timing(asc_scene s, int enable)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  if (enable && !s->ev[0])
    for (hipEvent_t& e : s->ev) ASC_HIP(hipEventCreate(&e));
  s->timing = enable != 0;
  destroy_graph(s);
  return ASC_OK;
}

ASC_API int asc_scene_last_timing(asc_scene s, float ms[3])
{
  ASC_CHECK(s && ms, ASC_ERR_INVALID, "bad argument");
  for (int k = 0; k < 3; k++) ms[k] = s->last_ms[k];
  return ASC_OK;
}

ASC_API int asc_scene_last_info(asc_scene s, int* info, int capacity, int64_t* samples)
{
  ASC_CHECK(s, ASC_ERR_INVALID, "null scene");
  ASC_CHECK(!s->fly[0].busy && !s->fly[1].busy, ASC_ERR_INVALID, "finish the submitted batches first");
  // the last finished batch's counts and moments, which finalize_items
  // wrote into its slot's pinned buffers. A batch that failed before its
  // buffers were sized leaves nothing to read.
  const int sl = s->last_slot;
  ASC_CHECK((size_t)s->n_items <= s->h_samples[sl].cap && (size_t)s->n_items * 10 <= s->h_results[sl].cap,
            ASC_ERR_INVALID, "no batch results to report");
  const unsigned long long* cnt = s->h_samples[sl].p;
  int64_t total = 0;
  for (int ii = 0; ii < s->n_items; ii++) {
    const int ki = s->korder[ii];
    total += (int64_t)cnt[ki];
    if (info && (ii + 1) * 8 <= capacity) {
      int* o = info + 8 * ii;
      const ItemDesc& it = s->items[ii];
      o[0] = s->in_view[ii];
      o[1] = it.w;
      o[2] = it.h;
      o[3] = (int)s->h_results[sl].p[(size_t)ki * 10 + 3];
      o[4] = it.tiles_x * it.tiles_y;
      o[5] = (int)cnt[ki];
      o[6] = o[7] = 0;
    }
  }
  if (samples) *samples = total;
  return ASC_OK;
}

ASC_API int asc_scene_last_moments(asc_scene s, double* moments, int capacity)
{
  ASC_CHECK(s && moments, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(!s->fly[0].busy && !s->fly[1].busy, ASC_ERR_INVALID, "finish the submitted batches first");
  ASC_CHECK((size_t)s->n_items * 10 <= s->h_results[s->last_slot].cap, ASC_ERR_INVALID, "no batch results to report");
  for (int ii = 0; ii < s->n_items && (ii + 1) * 10 <= capacity; ii++)
    for (int k = 0; k < 10; k++) moments[ii * 10 + k] = s->h_results[s->last_slot].p[(size_t)s->korder[ii] * 10 + k];
  return ASC_OK;
}

ASC_API int asc_scene_last_images(asc_scene s, int pose, int camera, float* drr, float* rad, float* mask)
{
  ASC_CHECK(s && s->debug, ASC_ERR_INVALID, "enable debug images with asc_scene_set_debug first");
  ASC_CHECK(!s->fly[0].busy && !s->fly[1].busy, ASC_ERR_INVALID, "finish the submitted batches first");
  ASC_CHECK(s->finished_seq == s->submitted_seq, ASC_ERR_INVALID,
            "the debug images are those of a later batch than the last one finished");
  ASC_CHECK(pose >= 0 && pose < (int)s->req_base.size() && camera >= 0 && camera < s->req_ncam[pose], ASC_ERR_INVALID,
            "bad item");
  const int ii = s->req_base[pose] + camera;
  const ItemDesc& it = s->items[ii];
  const size_t n = (size_t)it.w * it.h;
  const size_t plane = s->drr_total;
  if (n == 0) return ASC_OK;
  float* bases[3] = {drr, rad, mask};
  for (int k = 0; k < 3; k++)
    if (bases[k])
      ASC_HIP(hipMemcpy(bases[k], s->d_dbg.p + k * plane + it.drr_off, n * sizeof(float), hipMemcpyDeviceToHost));
  return ASC_OK;
}

// Tracker::trackFrame one pose at a time with one launch per operator and a
// blocking NCC per camera (Tracker.cpp:603-771, View.cpp:147-256).
ASC_API int asc_scene_track_frame_plain(asc_scene s, int volume, const double xyzypr[6], double* per_camera,
                                        double* cost)
{
  ASC_CHECK(s && xyzypr && volume >= 0 && volume < (int)s->vols.size(), ASC_ERR_INVALID, "bad argument");
  asc_context ctx = s->ctx;
  const size_t plane = (size_t)s->RW * s->RH;
  int rc;
  // rendered_drr, rendered_rad, drr_mask, background_mask, filter scratch
  if ((rc = s->p_bufs.ensure(plane * 5))) return rc;
  float* drr_raw = s->p_bufs.p;
  float* rad_buf = drr_raw + plane;
  float* mask = rad_buf + plane;
  float* bgm = mask + plane;
  float* scratch = bgm + plane;
  const SceneVolume& sv = s->vols[volume];
  ASC_CHECK(sv.in_use, ASC_ERR_INVALID, "bad volume index");
  const std::vector<int> cl = all_cameras(s);
  double total = 0.0;
  for (size_t c = 0; c < cl.size(); c++) {
    const SceneCamera& sc = s->cams[cl[c]];
    ASC_CHECK(sc.rad != nullptr, ASC_ERR_INVALID, "camera has no radiograph");
    float imv[12];
    double vp[4];
    unsigned sz[2];
    double corr;
    if (!view_geometry(sc.cam, xyzypr, sv.vol->desc.inv_scale, sv.vol->desc.inv_trans, s->RW, s->RH, imv, vp,
                       sz)) {
      corr = 9999.0;
    } else {
      if (sz[0] > (unsigned)s->RW || sz[1] > (unsigned)s->RH)
        return fail(ASC_ERR_TOO_LARGE, "Rendered image is larger than the viewport buffer");
      const float vpf[4] = {(float)vp[0], (float)vp[1], (float)vp[2], (float)vp[3]};
      const float plane4[4] = {(float)sc.cam.viewport[0], (float)sc.cam.viewport[1], (float)sc.cam.viewport[2],
                               (float)sc.cam.viewport[3]};
      const int w = (int)sz[0], h = (int)sz[1];
      // renderDrrSingle: DRR then the DRR chain, into `mask` as the filtered image holder
      float* drr_f = mask;
      if ((rc = launch_drr(ctx, sv.vol->ray_view(imv), imv, vpf, sv.step, sv.intensity, sv.cutoff, drr_raw, w, h)))
        return rc;
      if ((rc = asc_filter_chain(ctx, sc.chain[0].data(), (int)sc.chain[0].size(), drr_raw, drr_f, scratch, w, h)))
        return rc;
      // renderRad: resample then the rad chain
      if ((rc = launch_rad(ctx, sc.rad->dev, plane4, vpf, drr_raw, w, h, false, 0.0f))) return rc;
      if ((rc = asc_filter_chain(ctx, sc.chain[1].data(), (int)sc.chain[1].size(), drr_raw, rad_buf, scratch, w,
                                 h)))
        return rc;
      const float* bgp = nullptr;
      if (sc.bg && sc.bg_threshold >= 0.0f) {
        if ((rc = launch_rad(ctx, sc.bg->dev, plane4, vpf, bgm, w, h, true, sc.bg_threshold))) return rc;
        bgp = bgm;
      }
      // drr mask + 3 multiplies, then NCC(drr, rad, mask)
      if ((rc = launch_mask_multiply(ctx, drr_f, rad_buf, bgp, scratch, (size_t)w * h))) return rc;
      double ncc;
      if ((rc = ncc_moments(ctx, drr_f, rad_buf, scratch, (unsigned)(w * h), nullptr, &ncc))) return rc;
      corr = 1.0 - ncc;
    }
    if (per_camera) per_camera[c] = corr;
    total = c == 0 ? corr : total + corr;
  }
  if (total < 0) total = 9999;
  if (cost) *cost = total;
  return ASC_OK;
}

// View::render (View.cpp:355-394) over the camera's whole image plane, the
// viewport Tracker::getFullDRR uses (Tracker.cpp:1056-1064): each volume's
// DRR merged into one (Merger, min(a + b, 1)), the DRR filters, the
// radiograph and its filters, the background mask (1 without one), the DRR
// mask, then the Compositor's RGB (R = DRR, G = (DRR + rad) / 2 where the
// background mask is >= 0.5, B = rad).
ASC_API int asc_scene_render_view(asc_scene s, int camera, const double* poses, int n_poses, float* rgb,
                                  size_t rgb_capacity, unsigned* width, unsigned* height)
{
  ASC_CHECK(s && (poses || n_poses == 0) && n_poses >= 0, ASC_ERR_INVALID, "bad argument");
  for (size_t v = (size_t)n_poses; v < s->vols.size(); v++)
    ASC_CHECK(!s->vols[v].in_use, ASC_ERR_INVALID, "a pose is needed for every volume slot in use");
  ASC_CHECK(camera >= 0 && camera < (int)s->cams.size() && s->cams[camera].in_use, ASC_ERR_INVALID,
            "bad camera index");
  const SceneCamera& sc = s->cams[camera];
  ASC_CHECK(sc.rad != nullptr, ASC_ERR_INVALID, "camera has no radiograph");
  const asc_camera& cam = sc.cam;
  const double vpd[4] = {-cam.viewport[2] / 2, -cam.viewport[3] / 2, cam.viewport[2], cam.viewport[3]};
  const unsigned w = (unsigned)(vpd[2] * s->RW / cam.viewport[2]);
  const unsigned h = (unsigned)(vpd[3] * s->RH / cam.viewport[3]);
  if (width) *width = w;
  if (height) *height = h;
  if (!rgb) return ASC_OK;
  const size_t n = (size_t)w * h;
  ASC_CHECK(rgb_capacity >= 3 * n, ASC_ERR_TOO_LARGE, "rgb buffer smaller than 3 * width * height floats");
  asc_context ctx = s->ctx;
  int rc;
  // drr, merged, drr filtered, rad, rad filtered, background, drr mask, scratch
  if ((rc = s->p_bufs.ensure(n * 8))) return rc;
  float* drr = s->p_bufs.p;
  float* merged = drr + n;
  float* drr_f = merged + n;
  float* rad = drr_f + n;
  float* rad_f = rad + n;
  float* bgm = rad_f + n;
  float* mask = bgm + n;
  float* scratch = mask + n;
  const float vp[4] = {(float)vpd[0], (float)vpd[1], (float)vpd[2], (float)vpd[3]};
  const float plane[4] = {(float)cam.viewport[0], (float)cam.viewport[1], (float)cam.viewport[2],
                          (float)cam.viewport[3]};
  Frame cf;
  std::memcpy(cf.r, cam.rotation, sizeof cf.r);
  std::memcpy(cf.t, cam.translation, sizeof cf.t);
  const Frame cf_inv = frame_inverse(cf);
  if ((rc = launch_fill_f32(ctx, merged, 0.0f, n))) return rc;
  for (size_t v = 0; v < s->vols.size(); v++) {
    const SceneVolume& sv = s->vols[v];
    if (!sv.in_use) continue;
    // modelview = camera^-1 * pose; the ray caster takes its inverse (a
    // volume outside the view leaves its DRR all zero: the rays miss it)
    const Frame mv = frame_mul(cf_inv, frame_from_xyzypr(poses + 6 * v));
    double m16[16];
    frame_row_order(frame_inverse(mv), m16);
    float imv[16];
    inv_model_view(m16, sv.vol->desc.inv_scale, sv.vol->desc.inv_trans, imv);
    if ((rc = launch_drr(ctx, sv.vol->ray_view(imv), imv, vp, sv.step, sv.intensity, sv.cutoff, drr, w, h)))
      return rc;
    if ((rc = launch_binary(ctx, BIN_MERGE, merged, drr, merged, n))) return rc;
  }
  if ((rc = asc_filter_chain(ctx, sc.chain[0].data(), (int)sc.chain[0].size(), merged, drr_f, scratch, (int)w,
                             (int)h)))
    return rc;
  if ((rc = launch_rad(ctx, sc.rad->dev, plane, vp, rad, w, h, false, 0.0f))) return rc;
  if ((rc = asc_filter_chain(ctx, sc.chain[1].data(), (int)sc.chain[1].size(), rad, rad_f, scratch, (int)w,
                             (int)h)))
    return rc;
  if (sc.bg && sc.bg_threshold >= 0.0f)
    rc = launch_rad(ctx, sc.bg->dev, plane, vp, bgm, w, h, true, sc.bg_threshold);
  else
    rc = launch_fill_f32(ctx, bgm, 1.0f, n);
  if (rc) return rc;
  if ((rc = launch_binary(ctx, BIN_DRR_MASK, drr_f, drr_f, mask, n))) return rc;
  return launch_composite(ctx, drr_f, rad_f, bgm, rgb, n);
}

}  // extern "C"
