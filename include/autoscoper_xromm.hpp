// autoscoper_xromm.hpp -- standalone C++ facade with the reference's xromm /
// xromm::gpu class surface, implemented on the C ABI of autoscoper_hip.h.
//
// The rendering backend proper (what libautoscoper/src/gpu/opencl/*.hpp
// declare: Buffer, VolumeDescription, RayCaster, RadRenderer,
// BackgroundRenderer, the filters, ncc and the image operators) is
// autoscoper_xromm_gpu.hpp, which also compiles inside libautoscoper against
// the library's own Volume and Filter. This header adds, for use without
// libautoscoper, the host classes the backend's callers use: CoordFrame,
// Camera, View (View.cpp) and Tracker (the cost function and optimiser glue
// of Tracker.cpp), with the reference's odd corners kept:
//   View::filter() ping-pongs through one scratch buffer (View.cpp:475-555);
//   trackFrame returns 9999 for a camera the volume does not project into.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "autoscoper_xromm_gpu.hpp"

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
    // device copy of the volume (an extension: the reference always keeps
    // the file's texel type): ASC_STORAGE_* and ASC_LAYOUT_* (-1 = densest)
    int storage = ASC_STORAGE_U16;
    int layout = -1;
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
