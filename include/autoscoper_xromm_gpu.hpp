// autoscoper_xromm_gpu.hpp -- the xromm::gpu rendering backend (what
// libautoscoper/src/gpu/opencl/*.hpp declare) on the MI355X C ABI of
// autoscoper_hip.h.
//
// Two ways to compile against it:
//   * inside libautoscoper, as a third backend next to gpu/opencl and
//     gpu/cuda: define AUTOSCOPER_XROMM_HOST_TYPES and put the library's
//     src/ and math/ on the include path. xromm::Volume and the
//     xromm::gpu::Filter base then come from the library's own Volume.hpp
//     and Filter.hpp, unchanged;
//   * standalone (autoscoper_xromm.hpp): this header defines an in-memory
//     Volume and the Filter base with the same members.
//
// Signatures follow the OpenCL flavour of the reference (RayCaster.hpp:53-96,
// RadRenderer.hpp:52-73, BackgroundRenderer.hpp, SobelFilter.hpp,
// ContrastFilter.hpp, GaussianFilter.hpp, SharpenFilter.hpp,
// VolumeDescription.hpp:62-87, OpenCL.hpp:61-66,124-141, Ncc.hpp, Mult.hpp,
// DRRBackground.hpp, Merger.hpp, Compositor.hpp), including its quirks:
//   Buffer::read() copies host -> device and Buffer::write() device -> host
//     (OpenCL.cpp:1043-1057);
//   ncc() returns 1e5 when the denominator is < 1e-5 (Ncc.cpp:215-217).
// Errors that the reference handles with exit() throw xromm::gpu::Error (a
// std::runtime_error) here.
#pragma once

#include "autoscoper_xromm_gpu_types.hpp"

#ifdef AUTOSCOPER_XROMM_HOST_TYPES
// libautoscoper's own Volume (TIFF stack), Filter base and VolumeDescription
// (whose constructor and destructor this backend implements, as the
// reference's gpu/opencl/VolumeDescription.cpp does); they include
// autoscoper_xromm_gpu_types.hpp at their backend include sites.
#include "Filter.hpp"
#include "Volume.hpp"
#include "VolumeDescription.hpp"
#else
namespace xromm {

// Volume as the reference's Volume exposes it (Volume.hpp): here an
// in-memory stack, x fastest.
class XROMM_API Volume {
 public:
  Volume(const void* data, size_t width, size_t height, size_t depth, size_t bps, float scaleX, float scaleY,
         float scaleZ, bool flipX = false, bool flipY = false, bool flipZ = false);
  const void* data() const { return data_.data(); }
  size_t width() const { return w_; }
  size_t height() const { return h_; }
  size_t depth() const { return d_; }
  size_t bps() const { return bps_; }
  float scaleX() const { return s_[0]; }
  float scaleY() const { return s_[1]; }
  float scaleZ() const { return s_[2]; }
  bool flipX() const { return f_[0]; }
  bool flipY() const { return f_[1]; }
  bool flipZ() const { return f_[2]; }

 private:
  std::vector<unsigned char> data_;
  size_t w_, h_, d_, bps_;
  float s_[3];
  bool f_[3];
};

namespace gpu {

// Filter interface (Filter.hpp:56-93).
class XROMM_API Filter {
 public:
  enum {
    XROMM_GPU_CONTRAST_FILTER,
    XROMM_GPU_SOBEL_FILTER,
    XROMM_GPU_MEDIAN_FILTER,
    XROMM_GPU_GAUSSIAN_FILTER,
    XROMM_GPU_SHARPEN_FILTER
  };
  Filter(int type, const std::string& name) : type_(type), name_(name), enabled_(true) {}
  virtual ~Filter() {}
  virtual void apply(const Buffer* input, Buffer* output, int width, int height) = 0;
  int type() const { return type_; }
  const std::string& name() const { return name_; }
  void set_name(const std::string& n) { name_ = n; }
  bool enabled() const { return enabled_; }
  void set_enabled(bool e) { enabled_ = e; }

 protected:
  int type_;
  std::string name_;
  bool enabled_;
};

// VolumeDescription (VolumeDescription.hpp:62-87), with the library's
// members; `storage` selects the device copy (ASC_STORAGE_U16 is the
// reference's UNORM16, exact).
class XROMM_API VolumeDescription {
 public:
  explicit VolumeDescription(const Volume& volume, int storage = ASC_STORAGE_U16);
  ~VolumeDescription();
  VolumeDescription(const VolumeDescription&) = delete;
  VolumeDescription& operator=(const VolumeDescription&) = delete;
  const float* invScale() const { return invScale_; }
  const float* invTrans() const { return invTrans_; }
  const double* transCenter() const { return transCenter_; }
  float minValue() const { return minValue_; }
  float maxValue() const { return maxValue_; }
  const Image* image() const { return image_; }

 private:
  float minValue_;
  float maxValue_;
  float invScale_[3];
  float invTrans_[3];
  double transCenter_[3];
  Image* image_;
};

}  // namespace gpu
}  // namespace xromm
#endif

namespace xromm {
namespace gpu {


// RayCaster (RayCaster.hpp:53-96).
class XROMM_API RayCaster {
 public:
  RayCaster();
  void setVolume(VolumeDescription& volume);
  void setInvModelView(const double* invModelView);
  const float* getInvModelView() const { return imv_; }
  void setViewport(float x, float y, float width, float height);
  void render(const Buffer* buffer, unsigned width, unsigned height);
  void setVisible(bool v) { visible_ = v; }
  bool getVisible() const { return visible_; }
  float getSampleDistance() const { return sampleDistance_; }
  void setSampleDistance(float d) { sampleDistance_ = d; }
  float getRayIntensity() const { return rayIntensity_; }
  void setRayIntensity(float i) { rayIntensity_ = i; }
  float getCutoff() const { return cutoff_; }
  void setCutoff(float c) { cutoff_ = c; }
  float getMinCutoff() const;
  float getMaxCutoff() const;
  const std::string& getName() const { return name_; }
  void setName(const std::string& n) { name_ = n; }

 private:
  VolumeDescription* volumeDescription_;
  float imv_[16];
  float viewport_[4];
  float sampleDistance_, rayIntensity_, cutoff_;
  bool visible_;
  std::string name_;
};

// RadRenderer (RadRenderer.hpp:52-73).
class XROMM_API RadRenderer {
 public:
  RadRenderer();
  ~RadRenderer();
  RadRenderer(const RadRenderer&) = delete;
  RadRenderer& operator=(const RadRenderer&) = delete;
  void set_rad(const void* data, size_t width, size_t height, size_t bps);
  void set_image_plane(float x, float y, float width, float height);
  void set_viewport(float x, float y, float width, float height);
  void render(const Buffer* buffer, unsigned width, unsigned height) const;
  const std::string& getName() const { return name_; }
  void setName(const std::string& n) { name_ = n; }
  asc_image image() const { return image_; }

 private:
  asc_image image_;
  float image_plane_[4];
  float viewport_[4];
  std::string name_;
};

// BackgroundRenderer (BackgroundRenderer.cpp:76-129).
class XROMM_API BackgroundRenderer {
 public:
  BackgroundRenderer();
  ~BackgroundRenderer();
  BackgroundRenderer(const BackgroundRenderer&) = delete;
  BackgroundRenderer& operator=(const BackgroundRenderer&) = delete;
  void set_back(const void* data, size_t width, size_t height);
  void set_image_plane(float x, float y, float width, float height);
  void set_viewport(float x, float y, float width, float height);
  void render(const Buffer* buffer, unsigned width, unsigned height, float threshold) const;
  asc_image image() const { return image_; }

 private:
  asc_image image_;
  float image_plane_[4];
  float viewport_[4];
};

// The four filters (SobelFilter.hpp, ContrastFilter.hpp, GaussianFilter.hpp,
// SharpenFilter.hpp); defaults as SobelFilter.cpp:56-83,
// ContrastFilter.cpp:57-86, GaussianFilter.cpp:75-150, SharpenFilter.cpp:80-173.
class XROMM_API SobelFilter : public Filter {
 public:
  SobelFilter();
  void apply(const Buffer* input, Buffer* output, int width, int height) override;
  float scale() const { return scale_; }
  void set_scale(float s) { scale_ = s; }
  float blend() const { return blend_; }
  void set_blend(float b) { blend_ = b; }

 private:
  float scale_, blend_;
};

class XROMM_API ContrastFilter : public Filter {
 public:
  ContrastFilter();
  void apply(const Buffer* input, Buffer* output, int width, int height) override;
  float alpha() const { return alpha_; }
  void set_alpha(float a) { alpha_ = a; }
  float beta() const { return beta_; }
  void set_beta(float b) { beta_ = b; }
  int size() const { return size_; }

 private:
  float alpha_, beta_;
  int size_;
};

class XROMM_API GaussianFilter : public Filter {
 public:
  GaussianFilter();
  void apply(const Buffer* input, Buffer* output, int width, int height) override;
  float radius() const { return radius_; }
  void set_radius(float r) { radius_ = r < 0 ? 0 : r; }

 private:
  float radius_;
};

class XROMM_API SharpenFilter : public Filter {
 public:
  SharpenFilter();
  void apply(const Buffer* input, Buffer* output, int width, int height) override;
  float radius() const { return radius_; }
  void set_radius(float r) { radius_ = r < 0 ? 0 : r; }
  float contrast() const { return contrast_; }
  void set_contrast(float c) { contrast_ = c < 1 ? 1 : c; }
  float threshold() const { return threshold_; }
  void set_threshold(float t) { threshold_ = t; }

 private:
  float radius_, contrast_, threshold_;
};

// The filter's parameters as a record for the fused batch path
// (asc_scene_set_filters); throws for a filter type this backend has not
// implemented (the reference never implemented the median filter).
XROMM_API asc_filter filter_record(const Filter& f);

// Free functions (Ncc.hpp, Mult.hpp, DRRBackground.hpp, Merger.hpp,
// Compositor.hpp).
XROMM_API void ncc_init(unsigned max_n);
XROMM_API void ncc_deinit();
XROMM_API float ncc(Buffer* f, Buffer* g, Buffer* mask, unsigned n);
XROMM_API void multiply(const Buffer* src1, const Buffer* src2, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void drr_background(const Buffer* src1, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void merge(const Buffer* src1, const Buffer* src2, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void composite(const Buffer* src1, const Buffer* src2, const Buffer* src3, const Buffer* src4,
                         const Buffer* dest, unsigned width, unsigned height);

}  // namespace gpu
}  // namespace xromm
