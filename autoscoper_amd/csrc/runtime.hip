// Context, memory, volumes, images and the per-operator C entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "device.hpp"

namespace asc {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg)
{
  g_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what)
{
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return ASC_ERR_HIP;
}

namespace {

// ---------------------------------------------------------------------------
// VolumeDescription (a1), VolumeDescription.cpp:152-305: flip each axis on
// request, crop to the bounding box of non-zero voxels, record min/max.
// ---------------------------------------------------------------------------
template <class T>
struct Cropped {
  std::vector<T> data;
  int mn[3], mx[3], dim[3];
  T vmin, vmax;
};

template <class T>
int flip_crop(const T* src, int W, int H, int D, const int* flip, Cropped<T>& out)
{
  const size_t plane = (size_t)W * H;
  std::vector<T> fl(plane * D);
  // flipVolume, VolumeDescription.cpp:59-75
  for (int k = 0; k < D; k++) {
    const int z = flip[2] ? D - 1 - k : k;
    for (int i = 0; i < H; i++) {
      const int y = flip[1] ? H - 1 - i : i;
      const T* srow = src + k * plane + (size_t)i * W;
      T* drow = fl.data() + z * plane + (size_t)y * W;
      if (flip[0])
        for (int j = 0; j < W; j++) drow[W - 1 - j] = srow[j];
      else
        std::memcpy(drow, srow, sizeof(T) * W);
    }
  }
  // cropVolume, VolumeDescription.cpp:77-119
  int* mn = out.mn;
  int* mx = out.mx;
  mn[0] = W; mn[1] = H; mn[2] = D;
  mx[0] = mx[1] = mx[2] = 0;
  for (int k = 0; k < D; k++) {
    bool slice_nz = false;
    for (int i = 0; i < H; i++) {
      const T* row = fl.data() + k * plane + (size_t)i * W;
      int first = -1, last = -1;
      for (int j = 0; j < W; j++)
        if (row[j] != T(0)) {
          if (first < 0) first = j;
          last = j;
        }
      if (first >= 0) {
        mn[0] = std::min(mn[0], first);
        mx[0] = std::max(mx[0], last);
        mn[1] = std::min(mn[1], i);
        mx[1] = std::max(mx[1], i);
        slice_nz = true;
      }
    }
    if (slice_nz) {
      mn[2] = std::min(mn[2], k);
      mx[2] = std::max(mx[2], k);
    }
  }
  if (mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]) return fail(ASC_ERR_EMPTY_VOLUME, "empty volume");
  for (int a = 0; a < 3; a++) out.dim[a] = mx[a] - mn[a] + 1;
  // copyVolume, VolumeDescription.cpp:121-148
  out.data.resize((size_t)out.dim[0] * out.dim[1] * out.dim[2]);
  T lo = std::numeric_limits<T>::max(), hi = std::numeric_limits<T>::min();
  T* d = out.data.data();
  for (int k = mn[2]; k <= mx[2]; k++)
    for (int i = mn[1]; i <= mx[1]; i++) {
      const T* row = fl.data() + k * plane + (size_t)i * W;
      for (int j = mn[0]; j <= mx[0]; j++) {
        const T v = row[j];
        lo = std::min(lo, v);
        hi = std::max(hi, v);
        *d++ = v;
      }
    }
  out.vmin = lo;
  out.vmax = hi;
  return ASC_OK;
}

uint16_t to_bf16(float f)
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t rounding = 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)((u + rounding) >> 16);
}

// Pack the cropped volume (as normalised UNORM values u/full) into the
// padded device layout of the requested storage.
template <class T>
int upload_volume(asc_context ctx, const Cropped<T>& c, float full, int storage, asc_volume v)
{
  const int W = c.dim[0], H = c.dim[1], D = c.dim[2];
  const size_t pW = (size_t)W + 2, pH = (size_t)H + 2, pD = (size_t)D + 2;
  const size_t n = pW * pH * pD;
  size_t esize = storage == ASC_STORAGE_F32 ? 4 : 2;
  std::vector<uint8_t> host(n * esize, 0);
  int u16_max = 0;
  for (int k = 0; k < D; k++)
    for (int j = 0; j < H; j++)
      for (int i = 0; i < W; i++) {
        const T raw = c.data[((size_t)k * H + j) * W + i];
        const size_t di = ((size_t)(k + 1) * pH + (j + 1)) * pW + (i + 1);
        // UNORM: value = raw / (2^bits - 1), correctly rounded
        const float norm = (float)raw / full;
        switch (storage) {
          case ASC_STORAGE_U16: {
            // 8-bit volumes are widened exactly: v/255 == 257v/65535
            const uint16_t u = sizeof(T) == 1 ? (uint16_t)(raw * 257) : (uint16_t)raw;
            std::memcpy(&host[di * 2], &u, 2);
            u16_max = std::max<int>(u16_max, u);
            break;
          }
          case ASC_STORAGE_F16: {
            const __half hv = __float2half_rn(norm);
            std::memcpy(&host[di * 2], &hv, 2);
            break;
          }
          case ASC_STORAGE_BF16: {
            const uint16_t b = to_bf16(norm);
            std::memcpy(&host[di * 2], &b, 2);
            break;
          }
          default: std::memcpy(&host[di * 4], &norm, 4); break;
        }
      }
  ASC_HIP(hipMalloc(&v->dptr, host.size()));
  ASC_HIP(hipMemcpy(v->dptr, host.data(), host.size(), hipMemcpyHostToDevice));
  v->bytes = host.size();
  v->dev.data = v->dptr;
  v->dev.W = W;
  v->dev.H = H;
  v->dev.D = D;
  v->dev.pW = (int)pW;
  v->dev.pS = (int)(pW * pH);
  v->dev.storage = storage;
  v->dev.vscale = storage == ASC_STORAGE_U16 ? 1.0f / 65535.0f : 1.0f;
  v->dev.layout = ASC_LAYOUT_LINEAR;
  v->dev.esz = (int)esize;
  v->dev.nbytes = (int)std::min<size_t>(host.size(), 0x7fffffff);
  v->dev.pair_diff = 0;
  for (int a = 0; a < 3; a++) v->dev.blog[a] = v->dev.bin[a] = v->dev.bstr[a] = 0;
  for (VolDev& o : v->odev) o = v->dev;
  v->u16_max = u16_max;
  return ASC_OK;
}

bool valid_ctx(asc_context c) { return c != nullptr; }

}  // namespace
}  // namespace asc

using namespace asc;

extern "C" {

ASC_API const char* asc_last_error(void) { return g_error.c_str(); }

ASC_API int asc_version(void) { return 1; }

ASC_API int asc_device_count(int* count)
{
  ASC_CHECK(count, ASC_ERR_INVALID, "count is null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return ASC_OK;
}

ASC_API int asc_context_create(int device, asc_context* out)
{
  ASC_CHECK(out, ASC_ERR_INVALID, "out is null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(ASC_ERR_NO_DEVICE, "no HIP device");
  ASC_CHECK(device >= 0 && device < n, ASC_ERR_INVALID, "device index out of range");
  ASC_HIP(hipSetDevice(device));
  // the no-allocation entry point's scratch: its address on this device
  if (int rc = noalloc_scratch_init(device)) return rc;
  asc_context c = new asc_context_s();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  c->stream = c->own_stream;
  *out = c;
  return ASC_OK;
}

ASC_API int asc_context_destroy(asc_context ctx)
{
  if (!ctx) return ASC_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->ncc_partials) hipFree(ctx->ncc_partials);
  if (ctx->ncc_result) hipFree(ctx->ncc_result);
  if (ctx->ncc_host) hipHostFree(ctx->ncc_host);
  if (ctx->conv_wts) hipFree(ctx->conv_wts);
  if (ctx->tree_buf) hipFree(ctx->tree_buf);
  hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return ASC_OK;
}

ASC_API int asc_context_set_stream(asc_context ctx, void* s)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  return ASC_OK;
}

ASC_API void* asc_context_stream(asc_context ctx) { return ctx ? (void*)ctx->stream : nullptr; }

ASC_API int asc_context_device(asc_context ctx, int* device)
{
  ASC_CHECK(valid_ctx(ctx) && device, ASC_ERR_INVALID, "bad argument");
  *device = ctx->device;
  return ASC_OK;
}

ASC_API int asc_context_set_ncc_form(asc_context ctx, int form)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  ASC_CHECK(form >= ASC_NCC_MOMENTS && form <= ASC_NCC_TREE_OPENCL, ASC_ERR_INVALID, "bad NCC form");
  ctx->ncc_form = form;
  return ASC_OK;
}

ASC_API int asc_context_set_blocking(asc_context ctx, int enable)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  ctx->blocking = enable != 0;
  return ASC_OK;
}

ASC_API int asc_synchronize(asc_context ctx)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  return ASC_OK;
}

ASC_API int asc_malloc(asc_context ctx, size_t bytes, void** dptr)
{
  ASC_CHECK(valid_ctx(ctx) && dptr, ASC_ERR_INVALID, "bad argument");
  *dptr = nullptr;
  if (bytes == 0) return ASC_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) return fail(ASC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return ASC_OK;
}

ASC_API int asc_free(asc_context ctx, void* dptr)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  if (!dptr) return ASC_OK;
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  ASC_HIP(hipFree(dptr));
  return ASC_OK;
}

ASC_API int asc_memcpy_h2d(asc_context ctx, void* dst, const void* src, size_t bytes)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  if (bytes == 0) return ASC_OK;
  ASC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  return ASC_OK;
}

ASC_API int asc_memcpy_d2h(asc_context ctx, void* dst, const void* src, size_t bytes)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  if (bytes == 0) return ASC_OK;
  ASC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  return ASC_OK;
}

ASC_API int asc_memcpy_d2d(asc_context ctx, void* dst, const void* src, size_t bytes)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  return launch_copy(ctx, dst, src, bytes);
}

ASC_API int asc_fill_f32(asc_context ctx, float* dst, float value, size_t count)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  return launch_fill_f32(ctx, dst, value, count);
}

ASC_API int asc_fill_u8(asc_context ctx, void* dst, unsigned char value, size_t bytes)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  if (bytes == 0) return ASC_OK;
  ASC_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  return ASC_OK;
}

ASC_API int asc_volume_set_layout(asc_volume vol, int layout);

// The ray march addresses an element with signed 24-bit multiplies of the
// row and slice byte strides and a 32-bit buffer offset: a layout is usable
// when both strides are below 2^23 bytes and the whole array below 2^31.
// BRICK copies are addressed through LDS tables instead (no stride limit).
struct LayoutShape {
  int64_t esz;    // bytes per element
  int64_t pW;     // elements per row step (bytes = pW * esz)
  int64_t pS;     // elements per slice step
  int64_t bytes;  // whole array
};

static LayoutShape layout_shape(const VolDev& lin, int layout)
{
  const int64_t tex = lin.storage == ASC_STORAGE_F32 ? 4 : 2;
  const int64_t per = layout == ASC_LAYOUT_LINEAR ? 1 : layout == ASC_LAYOUT_QUAD ? 4 : 8;
  LayoutShape s;
  s.esz = per * tex;
  s.pW = lin.W + 2;
  s.pS = (int64_t)(lin.W + 2) * (lin.H + 2);
  s.bytes = s.pS * (lin.D + 2) * s.esz;
  return s;
}

// The BRICK copy for rays along `axis`: bricks of 1 cell along the axis, 2
// along the next (x -> y -> z -> x) and 4 along the third, ordered with
// the ray axis fastest. Cells -1 .. n-1 of each axis are sampled; brick b
// of an axis of extent 2^l holds cells (b - 1) 2^l .. b 2^l - 1, so
// floor((n - 1) / 2^l) + 2 bricks cover them. In-brick cells are ordered
// with the extent-2 axis fastest. Fills blog, bin, bstr, esz and nbytes
// (bytes in `bytes`).
static VolDev brick_view(const VolDev& lin, int axis, int64_t* bytes)
{
  VolDev b = lin;
  const int64_t E = 8 * (lin.storage == ASC_STORAGE_F32 ? 4 : 2);  // bytes per cell
  const int n[3] = {lin.W, lin.H, lin.D};
  const int u = (axis + 1) % 3, w = (axis + 2) % 3;
  b.blog[axis] = 0;
  b.blog[u] = 1;
  b.blog[w] = 2;
  b.bin[axis] = 0;
  b.bin[u] = (int)E;
  b.bin[w] = (int)(2 * E);
  int64_t nb[3];
  for (int a = 0; a < 3; a++) nb[a] = ((int64_t)(n[a] - 1) >> b.blog[a]) + 2;
  const int64_t sa = 8 * E, su = nb[axis] * sa, sw = nb[u] * su;
  *bytes = nb[w] * sw;
  b.bstr[axis] = (int)sa;
  b.bstr[u] = (int)std::min<int64_t>(su, INT32_MAX);
  b.bstr[w] = (int)std::min<int64_t>(sw, INT32_MAX);
  b.layout = ASC_LAYOUT_BRICK;
  b.esz = (int)E;
  b.pW = 0;
  b.pS = 0;
  b.nbytes = (int)std::min<int64_t>(*bytes, INT32_MAX);
  return b;
}

static bool layout_fits(const VolDev& lin, int layout)
{
  if (layout == ASC_LAYOUT_BRICK) {
    // the march's LDS offset tables, and 32-bit buffer offsets
    if (lin.W + lin.H + lin.D + 3 > kBrickTab) return false;
    for (int a = 0; a < 3; a++) {
      int64_t bytes = 0;
      brick_view(lin, a, &bytes);
      if (bytes >= (int64_t(1) << 31)) return false;
    }
    return true;
  }
  const LayoutShape s = layout_shape(lin, layout);
  return s.pS * s.esz < (int64_t(1) << 23) && s.bytes < (int64_t(1) << 31);
}


ASC_API int asc_volume_create(asc_context ctx, const void* data, int width, int height, int depth, int bps,
                              const int flip[3], const double scale[3], int storage, asc_volume* out,
                              asc_volume_desc* desc)
{
  ASC_CHECK(valid_ctx(ctx) && data && out && scale, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(width > 0 && height > 0 && depth > 0, ASC_ERR_INVALID, "bad volume shape");
  ASC_CHECK(bps == 8 || bps == 16, ASC_ERR_UNSUPPORTED, "unsupported volume bit depth (8 or 16)");
  ASC_CHECK(storage >= ASC_STORAGE_U16 && storage <= ASC_STORAGE_F32, ASC_ERR_INVALID, "bad storage");
  const int noflip[3] = {0, 0, 0};
  const int* fl = flip ? flip : noflip;
  ASC_HIP(hipSetDevice(ctx->device));
  asc_volume v = new asc_volume_s();
  v->ctx = ctx;
  v->dptr = nullptr;
  int mn[3], mx[3], dim[3];
  float vmin, vmax;
  int rc;
  if (bps == 16) {
    Cropped<uint16_t> c;
    rc = flip_crop((const uint16_t*)data, width, height, depth, fl, c);
    if (rc == ASC_OK) rc = upload_volume(ctx, c, 65535.0f, storage, v);
    if (rc == ASC_OK) {
      std::memcpy(mn, c.mn, sizeof mn);
      std::memcpy(mx, c.mx, sizeof mx);
      std::memcpy(dim, c.dim, sizeof dim);
      vmin = c.vmin / (float)65535;
      vmax = c.vmax / (float)65535;
    }
  } else {
    Cropped<uint8_t> c;
    rc = flip_crop((const uint8_t*)data, width, height, depth, fl, c);
    if (rc == ASC_OK) rc = upload_volume(ctx, c, 255.0f, storage, v);
    if (rc == ASC_OK) {
      std::memcpy(mn, c.mn, sizeof mn);
      std::memcpy(mx, c.mx, sizeof mx);
      std::memcpy(dim, c.dim, sizeof dim);
      vmin = c.vmin / (float)255;
      vmax = c.vmax / (float)255;
    }
  }
  if (rc != ASC_OK) {
    if (v->dptr) hipFree(v->dptr);
    delete v;
    return rc;
  }
  asc_volume_desc& d = v->desc;
  std::memset(&d, 0, sizeof d);
  for (int a = 0; a < 3; a++) {
    d.dims[a] = dim[a];
    d.crop_min[a] = mn[a];
    d.crop_max[a] = mx[a];
  }
  d.min_value = vmin;
  d.max_value = vmax;
  // VolumeDescription.cpp:255-267
  d.trans_center[0] = -scale[0] * 0.5 * (mn[0] + mx[0]);
  d.trans_center[1] = -scale[1] * 0.5 * (mn[1] + mx[1]);
  d.trans_center[2] = scale[2] * 0.5 * (mn[2] + mx[2]);
  for (int a = 0; a < 3; a++) d.inv_scale[a] = 1.0f / (float)(scale[a] * dim[a]);
  d.inv_trans[0] = -mn[0] / (float)dim[0];
  d.inv_trans[1] = -((height - mx[1] - 1) / (float)dim[1]);
  d.inv_trans[2] = mn[2] / (float)dim[2];
  d.storage = storage;
  // corner-replicated layout by default: one aligned vector load per sample
  // the densest layout whose addressing fits (OCTET: one load per sample)
  int layout = ASC_LAYOUT_BRICK;
  while (layout > ASC_LAYOUT_LINEAR && !layout_fits(v->dev, layout)) --layout;
  rc = asc_volume_set_layout(v, layout);
  if (rc != ASC_OK) {
    asc_volume_destroy(v);
    return rc;
  }
  if (desc) *desc = d;
  *out = v;
  return ASC_OK;
}

ASC_API int asc_volume_describe(asc_volume vol, asc_volume_desc* desc)
{
  ASC_CHECK(vol && desc, ASC_ERR_INVALID, "bad argument");
  *desc = vol->desc;
  return ASC_OK;
}

ASC_API int asc_volume_bytes(asc_volume vol, size_t* bytes)
{
  ASC_CHECK(vol && bytes, ASC_ERR_INVALID, "bad argument");
  *bytes = vol->bytes + vol->packed_bytes;
  return ASC_OK;
}

ASC_API int asc_volume_set_layout(asc_volume vol, int layout)
{
  ASC_CHECK(vol, ASC_ERR_INVALID, "null volume");
  ASC_CHECK(layout >= ASC_LAYOUT_LINEAR && layout <= ASC_LAYOUT_BRICK, ASC_ERR_INVALID, "bad layout");
  ASC_CHECK(layout_fits(vol->dev, layout), ASC_ERR_TOO_LARGE,
            "volume too large for this layout (slice >= 8 MiB, array >= 2 GiB, or W + H + D > 4093 for BRICK)");
  if (layout == vol->dev.layout && vol->dev.data) return ASC_OK;
  asc_context ctx = vol->ctx;
  vol->generation++;
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  for (void*& p : vol->packed) {
    if (p) ASC_HIP(hipFree(p));
    p = nullptr;
  }
  vol->packed_bytes = 0;
  VolDev lin = vol->dev;
  const LayoutShape lin_shape = layout_shape(lin, ASC_LAYOUT_LINEAR);
  lin.data = vol->dptr;
  lin.layout = ASC_LAYOUT_LINEAR;
  lin.esz = (int)lin_shape.esz;
  lin.pW = (int)lin_shape.pW;
  lin.pS = (int)lin_shape.pS;
  lin.nbytes = (int)lin_shape.bytes;
  lin.pair_diff = 0;
  vol->dev = lin;
  for (VolDev& o : vol->odev) o = lin;
  if (layout == ASC_LAYOUT_LINEAR) return ASC_OK;
  // the differences of u16 texels fit signed 16-bit integers: x1 - x0
  // when every texel (and the zero border) lies in [0, 32767], and the row
  // difference (t11 - t01) - (t10 - t00) too when they lie in [0, 16383]
  const int pair_diff =
    lin.storage != ASC_STORAGE_U16 ? 0 : vol->u16_max <= 16383 ? 2 : vol->u16_max <= 32767 ? 1 : 0;
  const int copies = layout == ASC_LAYOUT_BRICK ? 3 : 1;
  for (int a = 0; a < copies; a++) {
    VolDev v;
    int64_t bytes = 0;
    if (layout == ASC_LAYOUT_BRICK) {
      v = brick_view(lin, a, &bytes);
    } else {
      const LayoutShape shape = layout_shape(lin, layout);
      v = lin;
      v.layout = layout;
      v.esz = (int)shape.esz;
      v.pW = (int)shape.pW;
      v.pS = (int)shape.pS;
      v.nbytes = (int)shape.bytes;
      bytes = shape.bytes;
    }
    hipError_t e = hipMalloc(&vol->packed[a], (size_t)bytes);
    if (e != hipSuccess) {
      vol->packed[a] = nullptr;
      return fail(ASC_ERR_NOMEM, "packed volume allocation failed");
    }
    v.data = vol->packed[a];
    v.pair_diff = pair_diff;
    const int rc = layout == ASC_LAYOUT_BRICK ? launch_pack_brick(ctx, lin, v, vol->packed[a], pair_diff)
                                              : launch_pack_volume(ctx, lin, vol->packed[a], layout, pair_diff);
    if (rc) return rc;
    vol->packed_bytes += (size_t)bytes;
    vol->odev[a] = v;
  }
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  if (copies == 1) vol->odev[1] = vol->odev[2] = vol->odev[0];
  // `dev` describes the volume (size, storage, layout); the march reads the
  // view ray_view() picks
  vol->dev = vol->odev[0];
  return ASC_OK;
}

ASC_API int asc_volume_destroy(asc_volume vol)
{
  if (!vol) return ASC_OK;
  hipSetDevice(vol->ctx->device);
  hipStreamSynchronize(vol->ctx->stream);
  if (vol->dptr) hipFree(vol->dptr);
  for (void* p : vol->packed)
    if (p) hipFree(p);
  delete vol;
  return ASC_OK;
}

// Texels as the sampler reads them (UNORM8/UNORM16 -> raw / (2^bits - 1),
// correctly rounded, as the device division would give; float as is),
// converted once here instead of at every sample.
static std::vector<float> normalised_texels(const void* data, size_t n, int bps)
{
  std::vector<float> out(n);
  if (bps == 8) {
    const uint8_t* p = (const uint8_t*)data;
    for (size_t i = 0; i < n; i++) out[i] = (float)p[i] / 255.0f;
  } else if (bps == 16) {
    const uint16_t* p = (const uint16_t*)data;
    for (size_t i = 0; i < n; i++) out[i] = (float)p[i] / 65535.0f;
  } else {
    std::memcpy(out.data(), data, n * sizeof(float));
  }
  return out;
}

ASC_API int asc_image_create(asc_context ctx, const void* data, int width, int height, int bps, asc_image* out)
{
  ASC_CHECK(valid_ctx(ctx) && out, ASC_ERR_INVALID, "bad argument");
  ASC_CHECK(width > 0 && height > 0, ASC_ERR_INVALID, "bad image shape");
  // the samplers address texels with 32-bit byte offsets, and mark texels
  // outside the image with offsets of 2^30 and more
  ASC_CHECK((int64_t)width * height * 4 <= ((int64_t)1 << 30), ASC_ERR_TOO_LARGE, "image larger than 2^28 texels");
  ASC_CHECK(bps == 8 || bps == 16 || bps == 32, ASC_ERR_UNSUPPORTED, "unsupported image bit depth (8, 16, 32)");
  asc_image im = new asc_image_s();
  im->ctx = ctx;
  im->src_bps = bps;
  const size_t n = (size_t)width * height;
  im->bytes = n * sizeof(float);
  hipError_t e = hipMalloc(&im->dptr, im->bytes);
  if (e != hipSuccess) {
    delete im;
    return fail(ASC_ERR_NOMEM, "image allocation failed");
  }
  im->dev.data = im->dptr;
  im->dev.W = width;
  im->dev.H = height;
  im->dev.bps = 32;
  if (data) {
    const std::vector<float> host = normalised_texels(data, n, bps);
    e = hipMemcpy(im->dptr, host.data(), im->bytes, hipMemcpyHostToDevice);
  } else {
    e = hipMemset(im->dptr, 0, im->bytes);
  }
  if (e != hipSuccess) {
    hipFree(im->dptr);
    delete im;
    return hip_fail(e, "image upload");
  }
  *out = im;
  return ASC_OK;
}

ASC_API int asc_image_update(asc_image img, const void* data)
{
  ASC_CHECK(img && data, ASC_ERR_INVALID, "bad argument");
  const std::vector<float> host = normalised_texels(data, img->bytes / sizeof(float), img->src_bps);
  ASC_HIP(hipStreamSynchronize(img->ctx->stream));
  ASC_HIP(hipMemcpy(img->dptr, host.data(), img->bytes, hipMemcpyHostToDevice));
  return ASC_OK;
}

ASC_API int asc_image_destroy(asc_image img)
{
  if (!img) return ASC_OK;
  hipStreamSynchronize(img->ctx->stream);
  hipFree(img->dptr);
  delete img;
  return ASC_OK;
}

// ---- geometry -------------------------------------------------------------

ASC_API int asc_frame_from_xyzypr(const double xyzypr[6], double frame[12])
{
  ASC_CHECK(xyzypr && frame, ASC_ERR_INVALID, "null pointer");
  frame_store(frame_from_xyzypr(xyzypr), frame);
  return ASC_OK;
}

ASC_API int asc_frame_from_xyz_axis_angle(const double p[6], double frame[12])
{
  ASC_CHECK(p && frame, ASC_ERR_INVALID, "null pointer");
  frame_store(frame_from_xyz_axis_angle(p), frame);
  return ASC_OK;
}

ASC_API int asc_frame_to_xyzypr(const double frame[12], double xyzypr[6])
{
  ASC_CHECK(xyzypr && frame, ASC_ERR_INVALID, "null pointer");
  frame_to_xyzypr(frame_load(frame), xyzypr);
  return ASC_OK;
}

ASC_API int asc_frame_mul(const double a[12], const double b[12], double out[12])
{
  ASC_CHECK(a && b && out, ASC_ERR_INVALID, "null pointer");
  frame_store(frame_mul(frame_load(a), frame_load(b)), out);
  return ASC_OK;
}

ASC_API int asc_frame_inverse(const double a[12], double out[12])
{
  ASC_CHECK(a && out, ASC_ERR_INVALID, "null pointer");
  frame_store(frame_inverse(frame_load(a)), out);
  return ASC_OK;
}

ASC_API int asc_frame_linear_extrap(const double a[12], const double b[12], double out[12])
{
  ASC_CHECK(a && b && out, ASC_ERR_INVALID, "null pointer");
  frame_store(frame_linear_extrap(frame_load(a), frame_load(b)), out);
  return ASC_OK;
}

// Camera::loadMayaCam2 after parsing (Camera.cpp:300-362): undo the file's
// y-axis convention, invert the extrinsics, then calculateViewport and
// calculateImagePlane (Camera.cpp:424-484).
ASC_API int asc_camera_from_mayacam2(const double Kf[9], const double Rf[9], const double tf[3],
                                     const double size[2], asc_camera* out)
{
  ASC_CHECK(Kf && Rf && tf && size && out, ASC_ERR_INVALID, "null pointer");
  double K[3][3], rot[3][3], tr[3] = {tf[0], tf[1], tf[2]};
  for (int r = 0; r < 3; r++)
    for (int j = 0; j < 3; j++) {
      K[j][r] = Kf[3 * r + j];  // file line 4+r, value j -> K[j][r]
      rot[j][r] = Rf[3 * r + j];
    }
  asc_camera c;
  std::memset(&c, 0, sizeof c);
  c.size[0] = size[0];
  c.size[1] = size[1];
  tr[0] = -tr[0];
  tr[2] = -tr[2];
  for (int i = 0; i < 3; i++) {
    rot[i][0] = -rot[i][0];
    rot[i][2] = -rot[i][2];
  }
  K[2][1] = (c.size[1] - 1) - K[2][1];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < i; j++) std::swap(rot[i][j], rot[j][i]);
  double tn[3];
  for (int i = 0; i < 3; i++) tn[i] = -(tr[0] * rot[0][i] + tr[1] * rot[1][i] + tr[2] * rot[2][i]);
  for (int i = 0; i < 3; i++) {
    rot[0][i] = -rot[0][i];
    rot[1][i] = -rot[1][i];
  }
  std::memcpy(c.rotation, &rot[0][0], sizeof c.rotation);
  std::memcpy(c.translation, tn, sizeof c.translation);
  const double fx = K[0][0], fy = K[1][1];
  // average focal length, negated as MayaCam 1.0's z (Camera.cpp:358-361)
  const int rc = camera_finish(c, K[2][0], K[2][1], fx, fy, -0.5 * (fx + fy));
  if (rc) return rc;
  *out = c;
  return ASC_OK;
}

ASC_API int asc_inv_model_view(const double imv16[16], const float inv_scale[3], const float inv_trans[3],
                               float out16[16])
{
  ASC_CHECK(imv16 && inv_scale && inv_trans && out16, ASC_ERR_INVALID, "null pointer");
  inv_model_view(imv16, inv_scale, inv_trans, out16);
  return ASC_OK;
}

ASC_API int asc_calculate_viewport(const double modelview[12], const asc_camera* cam, const asc_volume_desc* desc,
                                   double vp_out[4])
{
  ASC_CHECK(modelview && cam && desc && vp_out, ASC_ERR_INVALID, "null pointer");
  return calc_viewport(frame_load(modelview), *cam, desc->inv_trans, desc->inv_scale, vp_out) ? 1 : 0;
}

ASC_API int asc_view_geometry(const asc_camera* cam, const double xyzypr[6], const asc_volume_desc* desc,
                              int render_w, int render_h, float imv_out[12], double vp_out[4],
                              unsigned size_out[2])
{
  ASC_CHECK(cam && xyzypr && desc && imv_out && vp_out && size_out, ASC_ERR_INVALID, "null pointer");
  return view_geometry(*cam, xyzypr, desc->inv_scale, desc->inv_trans, render_w, render_h, imv_out, vp_out,
                       size_out);
}

// ---- per-operator entry points --------------------------------------------

// A blocking context waits for each launch before the call returns.
static int finish(asc_context ctx, int rc)
{
  if (rc != ASC_OK || !ctx->blocking) return rc;
  ASC_HIP(hipStreamSynchronize(ctx->stream));
  return ASC_OK;
}

ASC_API int asc_drr_render(asc_context ctx, asc_volume vol, const float imv[12], const float viewport[4],
                           float step, float intensity, float cutoff, float* out, unsigned width, unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && vol && imv && viewport && out, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_drr(ctx, vol->ray_view(imv), imv, viewport, step, intensity, cutoff, out, width, height));
}

ASC_API int asc_rad_render(asc_context ctx, asc_image img, const float image_plane[4], const float viewport[4],
                           float* out, unsigned width, unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && img && image_plane && viewport && out, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_rad(ctx, img->dev, image_plane, viewport, out, width, height, false, 0.0f));
}

ASC_API int asc_background_render(asc_context ctx, asc_image img, const float image_plane[4],
                                  const float viewport[4], float threshold, float* out, unsigned width,
                                  unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && img && image_plane && viewport && out, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_rad(ctx, img->dev, image_plane, viewport, out, width, height, true, threshold));
}

ASC_API int asc_sobel(asc_context ctx, const float* in, float* out, int width, int height, float scale,
                      float blend)
{
  ASC_CHECK(valid_ctx(ctx) && in && out, ASC_ERR_INVALID, "bad argument");
  asc_filter f{ASC_FILTER_SOBEL, 1, {scale, blend, 0, 0}};
  return finish(ctx, launch_filter(ctx, f, in, out, width, height));
}

ASC_API int asc_contrast(asc_context ctx, const float* in, float* out, int width, int height, float alpha,
                         float beta, int size)
{
  ASC_CHECK(valid_ctx(ctx) && in && out, ASC_ERR_INVALID, "bad argument");
  asc_filter f{ASC_FILTER_CONTRAST, 1, {alpha, beta, (float)size, 0}};
  return finish(ctx, launch_filter(ctx, f, in, out, width, height));
}

ASC_API int asc_gaussian_weights(float radius, float* out, int capacity, int* size)
{
  ASC_CHECK(size, ASC_ERR_INVALID, "null size");
  std::vector<float> w;
  const int n = gauss_weights(radius, w);
  *size = n;
  if (out && capacity >= n * n) std::memcpy(out, w.data(), sizeof(float) * n * n);
  return ASC_OK;
}

ASC_API int asc_gaussian(asc_context ctx, const float* in, float* out, int width, int height, float radius)
{
  ASC_CHECK(valid_ctx(ctx) && in && out, ASC_ERR_INVALID, "bad argument");
  asc_filter f{ASC_FILTER_GAUSSIAN, 1, {radius, 0, 0, 0}};
  return finish(ctx, launch_filter(ctx, f, in, out, width, height));
}

ASC_API int asc_sharpen(asc_context ctx, const float* in, float* out, int width, int height, float radius,
                        float contrast, float threshold)
{
  ASC_CHECK(valid_ctx(ctx) && in && out, ASC_ERR_INVALID, "bad argument");
  asc_filter f{ASC_FILTER_SHARPEN, 1, {radius, contrast, threshold, 0}};
  return finish(ctx, launch_filter(ctx, f, in, out, width, height));
}

ASC_API int asc_filter_apply(asc_context ctx, const asc_filter* f, const float* in, float* out, int width,
                             int height)
{
  ASC_CHECK(valid_ctx(ctx) && f && in && out, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_filter(ctx, *f, in, out, width, height));
}

// View::filter, View.cpp:520-554: with an odd number of filters the first
// stage writes `out`, otherwise `scratch`; stages then alternate so the last
// one lands in `out`.
ASC_API int asc_filter_chain(asc_context ctx, const asc_filter* filters, int count, const float* in, float* out,
                             float* scratch, int width, int height)
{
  ASC_CHECK(valid_ctx(ctx) && in && out && (count == 0 || filters), ASC_ERR_INVALID, "bad argument");
  const size_t bytes = sizeof(float) * (size_t)width * height;
  if (count == 0) return finish(ctx, launch_copy(ctx, out, in, bytes));
  ASC_CHECK(count == 1 || scratch, ASC_ERR_INVALID, "scratch buffer required for chains longer than 1");
  float* b1 = (count % 2) ? out : scratch;
  float* b2 = (count % 2) ? scratch : out;
  int rc = finish(ctx, launch_filter(ctx, filters[0], in, b1, width, height));
  if (rc) return rc;
  for (int k = 1; k < count; k++) {
    rc = finish(ctx, launch_filter(ctx, filters[k], b1, b2, width, height));
    if (rc) return rc;
    std::swap(b1, b2);
  }
  return ASC_OK;
}

ASC_API int asc_drr_background(asc_context ctx, const float* src, float* dest, unsigned width, unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && src && dest, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_binary(ctx, BIN_DRR_MASK, src, src, dest, (size_t)width * height));
}

ASC_API int asc_multiply(asc_context ctx, const float* a, const float* b, float* dest, unsigned width,
                         unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && a && b && dest, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_binary(ctx, BIN_MULTIPLY, a, b, dest, (size_t)width * height));
}

ASC_API int asc_mask_multiply(asc_context ctx, float* drr, float* rad, const float* bg, float* mask,
                              unsigned width, unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && drr && rad, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_mask_multiply(ctx, drr, rad, bg, mask, (size_t)width * height));
}

ASC_API int asc_ncc(asc_context ctx, const float* f, const float* g, const float* mask, unsigned n, double* ncc,
                    double moments[9])
{
  ASC_CHECK(valid_ctx(ctx) && f && g && mask, ASC_ERR_INVALID, "bad argument");
  if (ctx->ncc_form != ASC_NCC_MOMENTS) {
    float sums[6], v;
    const int rc = ncc_tree(ctx, f, g, mask, n, ctx->ncc_form == ASC_NCC_TREE_OPENCL, sums, &v);
    if (rc) return rc;
    if (ncc) *ncc = v;
    if (moments) {
      for (int k = 0; k < 6; k++) moments[k] = sums[k];
      for (int k = 6; k < 9; k++) moments[k] = std::numeric_limits<double>::quiet_NaN();
    }
    return ASC_OK;
  }
  return ncc_moments(ctx, f, g, mask, n, moments, ncc);  // waits for its result already
}

ASC_API int asc_ncc_noalloc(asc_context ctx, const float* f, const float* g, const float* mask, unsigned n,
                            double* out10)
{
  ASC_CHECK(valid_ctx(ctx) && f && g && mask && out10, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, ncc_noalloc(ctx, f, g, mask, n, out10));
}

ASC_API int asc_ncc_init(asc_context ctx, unsigned max_n)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  // partial buffers grow on demand; nothing else to reserve
  (void)max_n;
  return ASC_OK;
}

ASC_API int asc_ncc_deinit(asc_context ctx)
{
  ASC_CHECK(valid_ctx(ctx), ASC_ERR_INVALID, "null context");
  return ASC_OK;
}

ASC_API int asc_merge(asc_context ctx, const float* a, const float* b, float* dest, unsigned width,
                      unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && a && b && dest, ASC_ERR_INVALID, "bad argument");
  return finish(ctx, launch_binary(ctx, BIN_MERGE, a, b, dest, (size_t)width * height));
}

ASC_API int asc_composite(asc_context ctx, const float* src1, const float* src2, const float* src3,
                          const float* src4, float* dest, unsigned width, unsigned height)
{
  ASC_CHECK(valid_ctx(ctx) && src1 && src2 && src3 && dest, ASC_ERR_INVALID, "bad argument");
  (void)src4;  // unused by the reference kernel as well (Compositor.cl:5)
  return finish(ctx, launch_composite(ctx, src1, src2, src3, dest, (size_t)width * height));
}

ASC_API int asc_pose_store_set(float store[7], const double xyzypr[6])
{
  ASC_CHECK(store && xyzypr, ASC_ERR_INVALID, "null pointer");
  PoseStore s;
  pose_store_set(s, xyzypr);
  std::memcpy(store, &s, sizeof s);
  return ASC_OK;
}

ASC_API int asc_pose_store_get(const float store[7], double xyzypr[6])
{
  ASC_CHECK(store && xyzypr, ASC_ERR_INVALID, "null pointer");
  PoseStore s;
  std::memcpy(&s, store, sizeof s);
  pose_store_get(s, xyzypr);
  return ASC_OK;
}

ASC_API int asc_compose_pose(const float store[7], const double pivot[12], const double manip[6],
                             double xyzypr[6])
{
  ASC_CHECK(store && manip && xyzypr, ASC_ERR_INVALID, "null pointer");
  PoseStore s;
  std::memcpy(&s, store, sizeof s);
  const Frame pv = pivot ? frame_load(pivot) : frame_identity();
  compose_pose(s, pv, manip, xyzypr);
  return ASC_OK;
}

}  // extern "C"
