// autoscoper_xromm.hpp -- C++ facade with the reference's xromm / xromm::gpu
// class surface, implemented on the C ABI of autoscoper_hip.h.
//
// It replaces the compile-time OpenCL/CUDA backend plugin
// (libautoscoper/src/gpu/opencl/*.hpp, selected in Tracker.hpp:50-59,
// View.hpp:47-52, Filter.hpp:47-51) with an MI355X backend. Names, argument
// meaning and the odd corners of the reference are kept:
//   Buffer::read() copies host -> device and Buffer::write() device -> host
//     (OpenCL.cpp:1043-1057);
//   View::filter() ping-pongs through one scratch buffer (View.cpp:475-555);
//   ncc() returns 1e5 when the denominator is < 1e-5 (Ncc.cpp:215-217).
// Errors that the reference handles with exit() or a thrown
// std::runtime_error throw xromm::gpu::Error (a std::runtime_error) here.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "autoscoper_hip.h"

#define XROMM_API __attribute__((visibility("default")))

namespace xromm {

// Rigid transform, CoordFrame.hpp (rotation columns are the images of the
// axes; translation in mm).
class XROMM_API CoordFrame {
 public:
  CoordFrame();
  CoordFrame(const double* rotation, const double* translation);
  static CoordFrame from_xyzypr(const double* xyzypr);
  static CoordFrame from_xyzAxis_angle(const double* xyzijk);
  void to_xyzypr(double* xyzypr) const;
  void to_matrix_row_order(double* m16) const;
  CoordFrame inverse() const;
  CoordFrame operator*(const CoordFrame& other) const;
  void point_to_world_space(const double* p, double* q) const;
  void translate(const double* v);
  const double* rotation() const { return f_; }
  const double* translation() const { return f_ + 9; }

 private:
  double f_[12];
};

// Volume as the reference's Volume (TIFF stack) exposes it (Volume.hpp):
// here an in-memory stack, x fastest.
class XROMM_API Volume {
 public:
  Volume(const void* data, size_t width, size_t height, size_t depth, int bps, double scaleX, double scaleY,
         double scaleZ, bool flipX = false, bool flipY = false, bool flipZ = false);
  const void* data() const { return data_.data(); }
  size_t width() const { return w_; }
  size_t height() const { return h_; }
  size_t depth() const { return d_; }
  int bps() const { return bps_; }
  double scaleX() const { return s_[0]; }
  double scaleY() const { return s_[1]; }
  double scaleZ() const { return s_[2]; }
  bool flipX() const { return f_[0]; }
  bool flipY() const { return f_[1]; }
  bool flipZ() const { return f_[2]; }

 private:
  std::vector<unsigned char> data_;
  size_t w_, h_, d_;
  int bps_;
  double s_[3];
  bool f_[3];
};

// Camera (Camera.hpp): coord_frame, image_plane[12], viewport[4], size[2].
class XROMM_API Camera {
 public:
  explicit Camera(const asc_camera& c);
  // Camera::loadMayaCam2 arithmetic on the parsed K, R, t (Camera.cpp:229-362)
  static Camera fromMayaCam2(const double K[9], const double R[9], const double t[3], const double size[2]);
  const CoordFrame& coord_frame() const { return frame_; }
  const double* image_plane() const { return cam_.image_plane; }
  const double* viewport() const { return cam_.viewport; }
  const double* size() const { return cam_.size; }
  const asc_camera& record() const { return cam_; }

 private:
  asc_camera cam_;
  CoordFrame frame_;
};

namespace gpu {

struct XROMM_API Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// The process-wide context of the facade (the reference keeps one global
// OpenCL context and queue, OpenCL.cpp:83-102). Created on first use on
// device 0 unless set_device() chose another.
XROMM_API asc_context context();
XROMM_API void set_device(int device);

// Device buffer (OpenCL.hpp:124-141 Buffer).
class XROMM_API Buffer {
 public:
  explicit Buffer(size_t size, int access = 0);
  ~Buffer();
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  void read(const void* buf, size_t size = 0) const;  // host -> device
  void write(void* buf, size_t size = 0) const;       // device -> host
  void copy(const Buffer* dst, size_t size = 0) const;
  void fill(const char c) const;
  void fill(const float val) const;
  float* get() const { return (float*)ptr_; }
  size_t size() const { return size_; }

 private:
  void* ptr_;
  size_t size_;
};

// VolumeDescription (VolumeDescription.hpp:62-87).
class XROMM_API VolumeDescription {
 public:
  explicit VolumeDescription(const Volume& volume, int storage = ASC_STORAGE_U16);
  ~VolumeDescription();
  VolumeDescription(const VolumeDescription&) = delete;
  VolumeDescription& operator=(const VolumeDescription&) = delete;
  const float* invScale() const { return desc_.inv_scale; }
  const float* invTrans() const { return desc_.inv_trans; }
  const double* transCenter() const { return desc_.trans_center; }
  float minValue() const { return desc_.min_value; }
  float maxValue() const { return desc_.max_value; }
  asc_volume image() const { return vol_; }
  const asc_volume_desc& desc() const { return desc_; }

 private:
  asc_volume vol_;
  asc_volume_desc desc_;
};

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
  void set_rad(const void* data, size_t width, size_t height, size_t bps);
  void set_image_plane(float x, float y, float width, float height);
  void set_viewport(float x, float y, float width, float height);
  void render(const Buffer* buffer, unsigned width, unsigned height) const;
  const std::string& getName() const { return name_; }
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
  // record for the fused path
  virtual asc_filter record() const = 0;

 protected:
  int type_;
  std::string name_;
  bool enabled_;
};

class XROMM_API SobelFilter : public Filter {
 public:
  SobelFilter();
  void apply(const Buffer* input, Buffer* output, int width, int height) override;
  float scale() const { return scale_; }
  void set_scale(float s) { scale_ = s; }
  float blend() const { return blend_; }
  void set_blend(float b) { blend_ = b; }
  asc_filter record() const override;

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
  asc_filter record() const override;

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
  asc_filter record() const override;

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
  asc_filter record() const override;

 private:
  float radius_, contrast_, threshold_;
};

// Free functions (Ncc.hpp:49-53, Mult.hpp, DRRBackground.hpp, Merger.hpp,
// Compositor.hpp).
XROMM_API void ncc_init(unsigned max_n);
XROMM_API void ncc_deinit();
XROMM_API float ncc(Buffer* f, Buffer* g, Buffer* mask, unsigned n);
XROMM_API void multiply(const Buffer* src1, const Buffer* src2, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void drr_background(const Buffer* src1, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void merge(const Buffer* src1, const Buffer* src2, const Buffer* dest, unsigned width, unsigned height);
XROMM_API void composite(const Buffer* src1, const Buffer* src2, const Buffer* src3, const Buffer* src4, const Buffer* dest,
               unsigned width, unsigned height);

// View (View.hpp:68-154); the GL display targets are replaced by a device
// RGB buffer (3 floats per pixel).
class XROMM_API View {
 public:
  explicit View(Camera& camera);
  ~View();
  View(const View&) = delete;
  View& operator=(const View&) = delete;
  Camera* camera() { return camera_; }
  RayCaster* drrRenderer(int idx) { return drrRenderer_[idx].get(); }
  RadRenderer* radRenderer() { return radRenderer_.get(); }
  BackgroundRenderer* backgroundRenderer() { return backgroundRenderer_.get(); }
  std::vector<Filter*>& drrFilters() { return drrFilters_; }
  std::vector<Filter*>& radFilters() { return radFilters_; }
  void addDrrRenderer();
  unsigned nbDrrRenderer() const { return (unsigned)drrRenderer_.size(); }
  void renderRad(Buffer* buffer, unsigned width, unsigned height);
  void renderBackground(Buffer* buffer, unsigned width, unsigned height);
  void renderDRRMask(Buffer* in_buffer, Buffer* out_buffer, unsigned width, unsigned height);
  void renderDrr(Buffer* buffer, unsigned width, unsigned height);
  void renderDrrSingle(int volume, Buffer* buffer, unsigned width, unsigned height);
  void render(Buffer* rgb, unsigned width, unsigned height);
  void updateBackground(const float* buffer, unsigned width, unsigned height);
  void setBackgroundThreshold(float t) { backgroundThreshold_ = t; }
  float backgroundThreshold() const { return backgroundThreshold_; }
  bool drr_enabled;
  bool rad_enabled;

 private:
  void init(unsigned width, unsigned height);
  void filter(const std::vector<Filter*>& filters, const Buffer* input, Buffer* output, unsigned width,
              unsigned height);
  Camera* camera_;
  std::vector<std::unique_ptr<RayCaster>> drrRenderer_;
  std::unique_ptr<RadRenderer> radRenderer_;
  std::unique_ptr<BackgroundRenderer> backgroundRenderer_;
  std::vector<Filter*> drrFilters_, radFilters_;
  size_t maxWidth_, maxHeight_;
  std::vector<std::unique_ptr<Buffer>> drrBuffer_;
  std::unique_ptr<Buffer> drrBufferMerged_, drrFilterBuffer_, radBuffer_, radFilterBuffer_, filterBuffer_,
    backgroundmask_, drr_mask_;
  float backgroundThreshold_;
  bool inited_;
};

}  // namespace gpu

// Tracker: the cost function and optimiser glue (Tracker.hpp:127-199) over
// the facade (one pose at a time, as the reference) and over the fused
// batch (many poses at once).
class XROMM_API Tracker {
 public:
  struct VolumeSpec {
    const Volume* volume;
    CoordFrame pivot;  // Trial::getVolumeMatrix before the transCenter shift
  };
  Tracker();
  ~Tracker();
  // Tracker::load (Tracker.cpp:244-335) from in-memory inputs: one
  // radiograph per camera (bps 8/16), render resolution.
  void load(const std::vector<VolumeSpec>& volumes, std::vector<Camera>& cameras,
            const std::vector<const void*>& rads, const std::vector<std::pair<int, int>>& rad_sizes, int rad_bps,
            int render_width, int render_height);
  gpu::View* view(int i) { return views_[i].get(); }
  size_t numViews() const { return views_.size(); }
  void setDrrParams(int volume, float step, float intensity, float cutoff);
  // Tracker::trackFrame (Tracker.cpp:583-773): per-camera 1 - NCC or 9999.
  std::vector<double> trackFrame(unsigned volumeID, const double* xyzypr) const;
  // Tracker::calculate_viewport (Tracker.cpp:923-1032)
  bool calculate_viewport(const CoordFrame& modelview, const Camera& camera, double* viewport) const;
  // Tracker::minimizationFunc on a stored pose (Tracker.cpp:775-827)
  double minimizationFunc(const float store[7], const double* values) const;
  // the same for n manipulator offsets in one fused batch
  std::vector<double> minimizationFuncBatch(const float store[7], const std::vector<double>& values) const;
  // Tracker::optimize for one frame (Tracker.cpp:337-580)
  asc_optimize_stats optimize(float store[7], const asc_optimize_options& opt);
  int currentVolume() const { return current_volume_; }
  void setCurrentVolume(int v) { current_volume_ = v; }
  asc_scene scene() const { return scene_; }
  CoordFrame volumeMatrix(int i) const { return pivots_[i]; }

 private:
  void sync_scene_filters();
  std::vector<std::unique_ptr<gpu::VolumeDescription>> volumeDescription_;
  std::vector<std::unique_ptr<gpu::View>> views_;
  std::vector<Camera> cameras_;
  std::vector<CoordFrame> pivots_;
  std::unique_ptr<gpu::Buffer> rendered_drr_, rendered_rad_, drr_mask_, background_mask_;
  std::vector<asc_image> scene_rads_;
  asc_scene scene_;
  int render_width_, render_height_;
  int current_volume_;
};

}  // namespace xromm
