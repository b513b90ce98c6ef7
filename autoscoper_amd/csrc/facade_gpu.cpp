// xromm::gpu rendering backend over the C ABI (include/autoscoper_xromm_gpu.hpp).
// Every device operation goes through an asc_* entry point; this file holds
// no device code. It compiles standalone and, with
// AUTOSCOPER_XROMM_HOST_TYPES, against libautoscoper's own Volume.hpp and
// Filter.hpp.
#include "autoscoper_xromm_gpu.hpp"

#include <cstring>
#include <iostream>
#include <mutex>
#include <sstream>
#include <utility>

namespace xromm {

namespace {

void check(int rc)
{
  if (rc < 0) throw gpu::Error(rc, asc_last_error());
}

}  // namespace

#ifndef AUTOSCOPER_XROMM_HOST_TYPES
Volume::Volume(const void* data, size_t width, size_t height, size_t depth, size_t bps, float scaleX, float scaleY,
               float scaleZ, bool flipX, bool flipY, bool flipZ)
  : w_(width), h_(height), d_(depth), bps_(bps)
{
  if (bps != 8 && bps != 16) throw gpu::Error(ASC_ERR_UNSUPPORTED, "Volume: unsupported bit depth");
  const size_t n = width * height * depth * (bps / 8);
  data_.assign((const unsigned char*)data, (const unsigned char*)data + n);
  s_[0] = scaleX;
  s_[1] = scaleY;
  s_[2] = scaleZ;
  f_[0] = flipX;
  f_[1] = flipY;
  f_[2] = flipZ;
}
#endif

namespace gpu {

namespace {
std::mutex g_ctx_mutex;
asc_context g_ctx = nullptr;
int g_device = 0;
int g_ray_casters = 0, g_rad_renderers = 0, g_sobel = 0, g_contrast = 0, g_gauss = 0, g_sharpen = 0;

std::string numbered(const char* base, int& counter)
{
  std::ostringstream s;
  s << base << (++counter);
  return s.str();
}
}  // namespace

asc_context context()
{
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  if (!g_ctx) check(asc_context_create(g_device, &g_ctx));
  return g_ctx;
}

void set_device(int device)
{
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  if (g_ctx) throw Error(ASC_ERR_INVALID, "set_device must precede the first GPU call");
  g_device = device;
}

// OpenCL.cpp lists platforms with their GPU devices; here there is one
// platform whose devices are the HIP devices.
std::vector<std::vector<std::string>> get_platforms()
{
  int n = 0;
  check(asc_device_count(&n));
  std::vector<std::string> devs;
  for (int i = 0; i < n; i++) {
    std::ostringstream s;
    s << "HIP device " << i;
    devs.push_back(s.str());
  }
  std::vector<std::vector<std::string>> out;
  out.push_back(devs);
  return out;
}

void setUsedPlatform(int platform_idx) { setUsedPlatform(platform_idx, 0); }

void setUsedPlatform(int platform, int device)
{
  if (platform != 0) throw Error(ASC_ERR_INVALID, "setUsedPlatform: the only platform is 0 (HIP)");
  set_device(device);
}

std::pair<int, int> getUsedPlatform()
{
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  return std::make_pair(0, g_device);
}

// ---- Buffer (OpenCL.cpp:1027-1095) ----------------------------------------
Buffer::Buffer(size_t size, int) : ptr_(nullptr), size_(size) { check(asc_malloc(context(), size, &ptr_)); }

Buffer::~Buffer()
{
  if (ptr_) asc_free(context(), ptr_);
}

void Buffer::read(const void* buf, size_t size) const
{
  check(asc_memcpy_h2d(context(), ptr_, buf, size ? size : size_));
}

void Buffer::write(void* buf, size_t size) const { check(asc_memcpy_d2h(context(), buf, ptr_, size ? size : size_)); }

void Buffer::copy(const Buffer* dst, size_t size) const
{
  check(asc_memcpy_d2d(context(), dst->ptr_, ptr_, size ? size : size_));
}

void Buffer::fill(const char c) const { check(asc_fill_u8(context(), ptr_, (unsigned char)c, size_)); }

void Buffer::fill(const float v) const { check(asc_fill_f32(context(), (float*)ptr_, v, size_ / sizeof(float))); }

// ---- Image, VolumeDescription (VolumeDescription.cpp:59-305) ----------------
Image::~Image()
{
  if (vol_) asc_volume_destroy(vol_);
}

namespace {
// flip, crop, normalise and upload (asc_volume_create), then the
// description's fields as the reference's constructor sets them
Image* make_image(const Volume& v, int storage, float& minV, float& maxV, float* invScale, float* invTrans,
                  double* transCenter)
{
  const int flip[3] = {v.flipX(), v.flipY(), v.flipZ()};
  const double scale[3] = {v.scaleX(), v.scaleY(), v.scaleZ()};
  asc_volume vol = nullptr;
  asc_volume_desc d;
  check(asc_volume_create(context(), v.data(), (int)v.width(), (int)v.height(), (int)v.depth(), v.bps(), flip, scale,
                          storage, &vol, &d));
  minV = d.min_value;
  maxV = d.max_value;
  for (int i = 0; i < 3; i++) {
    invScale[i] = d.inv_scale[i];
    invTrans[i] = d.inv_trans[i];
    transCenter[i] = d.trans_center[i];
  }
  return new Image(vol, d);
}
}  // namespace

#ifdef AUTOSCOPER_XROMM_HOST_TYPES
// the library's class (VolumeDescription.hpp:62-87): u16 storage, the
// reference's UNORM16
VolumeDescription::VolumeDescription(const Volume& v)
  : image_(make_image(v, ASC_STORAGE_U16, minValue_, maxValue_, invScale_, invTrans_, transCenter_))
{
}
#else
VolumeDescription::VolumeDescription(const Volume& v, int storage)
  : image_(make_image(v, storage, minValue_, maxValue_, invScale_, invTrans_, transCenter_))
{
}
#endif

VolumeDescription::~VolumeDescription() { delete image_; }

// ---- RayCaster (RayCaster.cpp:60-193) ---------------------------------------
RayCaster::RayCaster()
  : volumeDescription_(nullptr), sampleDistance_(0.5f), rayIntensity_(10.f), cutoff_(0.0f), visible_(true)
{
  std::memset(imv_, 0, sizeof imv_);
  viewport_[0] = -1.0f;
  viewport_[1] = -1.0f;
  viewport_[2] = 2.0f;
  viewport_[3] = 2.0f;
  name_ = numbered("DrrRenderer", g_ray_casters);
}

void RayCaster::setVolume(VolumeDescription& volume) { volumeDescription_ = &volume; }

void RayCaster::setInvModelView(const double* invModelView)
{
  if (!volumeDescription_) throw Error(ASC_ERR_INVALID, "RayCaster: ERROR: Unable to calculate matrix.");
  check(asc_inv_model_view(invModelView, volumeDescription_->invScale(), volumeDescription_->invTrans(), imv_));
}

void RayCaster::setViewport(float x, float y, float width, float height)
{
  viewport_[0] = x;
  viewport_[1] = y;
  viewport_[2] = width;
  viewport_[3] = height;
}

void RayCaster::render(const Buffer* buffer, unsigned width, unsigned height)
{
  if (!volumeDescription_) {
    std::cerr << "RayCaster: WARNING: No volume loaded." << std::endl;
    return;
  }
  if (!visible_) {
    buffer->fill((char)0x00);
    return;
  }
  check(asc_drr_render(context(), volumeDescription_->image()->volume(), imv_, viewport_, sampleDistance_,
                       rayIntensity_, cutoff_, buffer->get(), width, height));
}

float RayCaster::getMinCutoff() const { return volumeDescription_ ? volumeDescription_->minValue() : 0.0f; }
float RayCaster::getMaxCutoff() const { return volumeDescription_ ? volumeDescription_->maxValue() : 1.0f; }

// ---- RadRenderer / BackgroundRenderer -----------------------------------------
RadRenderer::RadRenderer() : image_(nullptr)
{
  const float d[4] = {-1.0f, -1.0f, 2.0f, 2.0f};
  std::memcpy(image_plane_, d, sizeof d);
  std::memcpy(viewport_, d, sizeof d);
  name_ = numbered("RadRenderer", g_rad_renderers);
}

RadRenderer::~RadRenderer()
{
  if (image_) asc_image_destroy(image_);
}

void RadRenderer::set_rad(const void* data, size_t width, size_t height, size_t bps)
{
  if (bps != 8 && bps != 16 && bps != 32) throw Error(ASC_ERR_UNSUPPORTED, "RadRenderer::rad(): Unsupported bit depth");
  if (image_) asc_image_destroy(image_);
  image_ = nullptr;
  check(asc_image_create(context(), data, (int)width, (int)height, (int)bps, &image_));
}

void RadRenderer::set_image_plane(float x, float y, float width, float height)
{
  image_plane_[0] = x;
  image_plane_[1] = y;
  image_plane_[2] = width;
  image_plane_[3] = height;
}

void RadRenderer::set_viewport(float x, float y, float width, float height)
{
  viewport_[0] = x;
  viewport_[1] = y;
  viewport_[2] = width;
  viewport_[3] = height;
}

void RadRenderer::render(const Buffer* buffer, unsigned width, unsigned height) const
{
  if (!image_) throw Error(ASC_ERR_INVALID, "RadRenderer: no radiograph set");
  check(asc_rad_render(context(), image_, image_plane_, viewport_, buffer->get(), width, height));
}

BackgroundRenderer::BackgroundRenderer() : image_(nullptr)
{
  const float d[4] = {-1.0f, -1.0f, 2.0f, 2.0f};
  std::memcpy(image_plane_, d, sizeof d);
  std::memcpy(viewport_, d, sizeof d);
}

BackgroundRenderer::~BackgroundRenderer()
{
  if (image_) asc_image_destroy(image_);
}

void BackgroundRenderer::set_back(const void* data, size_t width, size_t height)
{
  if (image_) asc_image_destroy(image_);
  image_ = nullptr;
  check(asc_image_create(context(), data, (int)width, (int)height, 32, &image_));
}

void BackgroundRenderer::set_image_plane(float x, float y, float width, float height)
{
  image_plane_[0] = x;
  image_plane_[1] = y;
  image_plane_[2] = width;
  image_plane_[3] = height;
}

void BackgroundRenderer::set_viewport(float x, float y, float width, float height)
{
  viewport_[0] = x;
  viewport_[1] = y;
  viewport_[2] = width;
  viewport_[3] = height;
}

void BackgroundRenderer::render(const Buffer* buffer, unsigned width, unsigned height, float threshold) const
{
  if (!image_) throw Error(ASC_ERR_INVALID, "BackgroundRenderer: no background set");
  check(asc_background_render(context(), image_, image_plane_, viewport_, threshold, buffer->get(), width, height));
}

// ---- filters ---------------------------------------------------------------
SobelFilter::SobelFilter() : Filter(XROMM_GPU_SOBEL_FILTER, ""), scale_(1.0f), blend_(0.5f)
{
  name_ = numbered("SobelFilter", g_sobel);
}

void SobelFilter::apply(const Buffer* in, Buffer* out, int w, int h)
{
  check(asc_sobel(context(), in->get(), out->get(), w, h, scale_, blend_));
}

ContrastFilter::ContrastFilter() : Filter(XROMM_GPU_CONTRAST_FILTER, ""), alpha_(1.0f), beta_(1.0f), size_(3)
{
  name_ = numbered("ContrastFilter", g_contrast);
}

void ContrastFilter::apply(const Buffer* in, Buffer* out, int w, int h)
{
  check(asc_contrast(context(), in->get(), out->get(), w, h, alpha_, beta_, size_));
}

GaussianFilter::GaussianFilter() : Filter(XROMM_GPU_GAUSSIAN_FILTER, ""), radius_(1.0f)
{
  name_ = numbered("GaussianFilter", g_gauss);
}

void GaussianFilter::apply(const Buffer* in, Buffer* out, int w, int h)
{
  check(asc_gaussian(context(), in->get(), out->get(), w, h, radius_));
}

SharpenFilter::SharpenFilter()
  : Filter(XROMM_GPU_SHARPEN_FILTER, ""), radius_(1.0f), contrast_(1.0f), threshold_(0.0f)
{
  name_ = numbered("SharpenFilter", g_sharpen);
}

void SharpenFilter::apply(const Buffer* in, Buffer* out, int w, int h)
{
  check(asc_sharpen(context(), in->get(), out->get(), w, h, radius_, contrast_, threshold_));
}

asc_filter filter_record(const Filter& f)
{
  const int en = f.enabled() ? 1 : 0;
  switch (f.type()) {
    case Filter::XROMM_GPU_SOBEL_FILTER: {
      const SobelFilter& s = static_cast<const SobelFilter&>(f);
      return asc_filter{ASC_FILTER_SOBEL, en, {s.scale(), s.blend(), 0, 0}};
    }
    case Filter::XROMM_GPU_CONTRAST_FILTER: {
      const ContrastFilter& c = static_cast<const ContrastFilter&>(f);
      return asc_filter{ASC_FILTER_CONTRAST, en, {c.alpha(), c.beta(), (float)c.size(), 0}};
    }
    case Filter::XROMM_GPU_GAUSSIAN_FILTER: {
      const GaussianFilter& g = static_cast<const GaussianFilter&>(f);
      return asc_filter{ASC_FILTER_GAUSSIAN, en, {g.radius(), 0, 0, 0}};
    }
    case Filter::XROMM_GPU_SHARPEN_FILTER: {
      const SharpenFilter& s = static_cast<const SharpenFilter&>(f);
      return asc_filter{ASC_FILTER_SHARPEN, en, {s.radius(), s.contrast(), s.threshold(), 0}};
    }
    default:
      throw Error(ASC_ERR_UNSUPPORTED, "filter type not implemented by this backend");
  }
}

// ---- free functions ----------------------------------------------------------
void ncc_init(unsigned max_n) { check(asc_ncc_init(context(), max_n)); }
void ncc_deinit() { check(asc_ncc_deinit(context())); }

float ncc(Buffer* f, Buffer* g, Buffer* mask, unsigned n)
{
  double v = 0.0;
  check(asc_ncc(context(), f->get(), g->get(), mask->get(), n, &v, nullptr));
  return (float)v;
}

void multiply(const Buffer* a, const Buffer* b, const Buffer* dest, unsigned w, unsigned h)
{
  check(asc_multiply(context(), a->get(), b->get(), dest->get(), w, h));
}

void drr_background(const Buffer* src, const Buffer* dest, unsigned w, unsigned h)
{
  check(asc_drr_background(context(), src->get(), dest->get(), w, h));
}

void merge(const Buffer* a, const Buffer* b, const Buffer* dest, unsigned w, unsigned h)
{
  check(asc_merge(context(), a->get(), b->get(), dest->get(), w, h));
}

void composite(const Buffer* s1, const Buffer* s2, const Buffer* s3, const Buffer* s4, const Buffer* dest, unsigned w,
               unsigned h)
{
  check(asc_composite(context(), s1->get(), s2->get(), s3->get(), s4 ? s4->get() : nullptr, dest->get(), w, h));
}

}  // namespace gpu
}  // namespace xromm
