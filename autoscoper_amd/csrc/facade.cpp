// Standalone host classes of the facade (include/autoscoper_xromm.hpp):
// CoordFrame, Camera, View and Tracker over the C ABI and the xromm::gpu
// backend of facade_gpu.cpp. No device code.
#include "autoscoper_xromm.hpp"

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

// ---------------------------------------------------------------------------
// CoordFrame (CoordFrame.cpp) through the C ABI's double-precision helpers.
// ---------------------------------------------------------------------------
CoordFrame::CoordFrame()
{
  std::memset(f_, 0, sizeof f_);
  f_[0] = f_[4] = f_[8] = 1.0;
}

CoordFrame::CoordFrame(const double* rotation, const double* translation)
{
  std::memcpy(f_, rotation, 9 * sizeof(double));
  std::memcpy(f_ + 9, translation, 3 * sizeof(double));
}

CoordFrame CoordFrame::from_xyzypr(const double* xyzypr)
{
  CoordFrame c;
  check(asc_frame_from_xyzypr(xyzypr, c.f_));
  return c;
}

CoordFrame CoordFrame::from_xyzAxis_angle(const double* xyzijk)
{
  CoordFrame c;
  check(asc_frame_from_xyz_axis_angle(xyzijk, c.f_));
  return c;
}

void CoordFrame::to_xyzypr(double* xyzypr) const { check(asc_frame_to_xyzypr(f_, xyzypr)); }

void CoordFrame::to_matrix_row_order(double* m) const
{
  // CoordFrame.cpp:223-241
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) m[4 * row + col] = f_[3 * col + row];
    m[4 * row + 3] = f_[9 + row];
  }
  m[12] = m[13] = m[14] = 0.0;
  m[15] = 1.0;
}

CoordFrame CoordFrame::inverse() const
{
  CoordFrame c;
  check(asc_frame_inverse(f_, c.f_));
  return c;
}

CoordFrame CoordFrame::operator*(const CoordFrame& o) const
{
  CoordFrame c;
  check(asc_frame_mul(f_, o.f_, c.f_));
  return c;
}

void CoordFrame::point_to_world_space(const double* p, double* q) const
{
  // CoordFrame.cpp:355-370
  for (int i = 0; i < 3; i++) q[i] = f_[i] * p[0] + f_[3 + i] * p[1] + f_[6 + i] * p[2] + f_[9 + i];
}

void CoordFrame::translate(const double* v)
{
  for (int i = 0; i < 3; i++) f_[9 + i] += v[i];
}

Camera::Camera(const asc_camera& c) : cam_(c), frame_(c.rotation, c.translation) {}

Camera Camera::fromMayaCam2(const double K[9], const double R[9], const double t[3], const double size[2])
{
  asc_camera c;
  check(asc_camera_from_mayacam2(K, R, t, size, &c));
  return Camera(c);
}

namespace gpu {

// ---- View (View.cpp) -----------------------------------------------------------
View::View(Camera& camera)
  : drr_enabled(true),
    rad_enabled(true),
    camera_(&camera),
    radRenderer_(new RadRenderer()),
    backgroundRenderer_(new BackgroundRenderer()),
    maxWidth_(3000),
    maxHeight_(3000),
    backgroundThreshold_(-1.0f),
    inited_(false)
{
}

View::~View()
{
  for (Filter* f : drrFilters_) delete f;
  for (Filter* f : radFilters_) delete f;
}

void View::addDrrRenderer()
{
  drrRenderer_.emplace_back(new RayCaster());
  drrBuffer_.emplace_back(nullptr);
}

// View.cpp:446-472 (OpenCL flavour): 3000 x 3000 scratch images, masks = 1
void View::init(unsigned width, unsigned height)
{
  if ((size_t)width * height > maxWidth_ * maxHeight_)
    throw Error(ASC_ERR_TOO_LARGE, "GPU Buffers too small.");
  if (inited_) return;
  const size_t bytes = maxWidth_ * maxHeight_ * sizeof(float);
  filterBuffer_.reset(new Buffer(bytes));
  for (auto& b : drrBuffer_) b.reset(new Buffer(bytes));
  drrBufferMerged_.reset(new Buffer(bytes));
  drrFilterBuffer_.reset(new Buffer(bytes));
  radBuffer_.reset(new Buffer(bytes));
  radFilterBuffer_.reset(new Buffer(bytes));
  backgroundmask_.reset(new Buffer(bytes));
  drr_mask_.reset(new Buffer(bytes));
  backgroundmask_->fill(1.0f);
  drr_mask_->fill(1.0f);
  inited_ = true;
}

void View::renderRad(Buffer* buffer, unsigned width, unsigned height)
{
  init(width, height);
  radRenderer_->render(radBuffer_.get(), width, height);
  filter(radFilters_, radBuffer_.get(), buffer, width, height);
}

void View::renderBackground(Buffer* buffer, unsigned width, unsigned height)
{
  if (backgroundThreshold_ < 0.0) return;
  backgroundRenderer_->render(buffer, width, height, backgroundThreshold_);
}

void View::renderDRRMask(Buffer* in_buffer, Buffer* out_buffer, unsigned width, unsigned height)
{
  drr_background(in_buffer, out_buffer, width, height);
}

void View::renderDrr(Buffer* buffer, unsigned width, unsigned height)
{
  init(width, height);
  drrBufferMerged_->fill((char)0x00);
  for (size_t i = 0; i < drrRenderer_.size(); i++) {
    drrRenderer_[i]->render(drrBuffer_[i].get(), width, height);
    merge(drrBufferMerged_.get(), drrBuffer_[i].get(), drrBufferMerged_.get(), width, height);
  }
  filter(drrFilters_, drrBufferMerged_.get(), buffer, width, height);
}

void View::renderDrrSingle(int volume, Buffer* buffer, unsigned width, unsigned height)
{
  init(width, height);
  drrRenderer_[volume]->render(drrBuffer_[volume].get(), width, height);
  filter(drrFilters_, drrBuffer_[volume].get(), buffer, width, height);
}

void View::render(Buffer* rgb, unsigned width, unsigned height)
{
  init(width, height);
  if (drr_enabled)
    renderDrr(drrFilterBuffer_.get(), width, height);
  else
    drrFilterBuffer_->fill((char)0x00);
  if (rad_enabled)
    renderRad(radFilterBuffer_.get(), width, height);
  else
    radFilterBuffer_->fill((char)0x00);
  renderBackground(backgroundmask_.get(), width, height);
  renderDRRMask(drrFilterBuffer_.get(), drr_mask_.get(), width, height);
  composite(drrFilterBuffer_.get(), radFilterBuffer_.get(), backgroundmask_.get(), drr_mask_.get(), rgb, width, height);
}

void View::updateBackground(const float* buffer, unsigned width, unsigned height)
{
  backgroundRenderer_->set_back(buffer, width, height);
}

void View::filter(const std::vector<Filter*>& filters, const Buffer* input, Buffer* output, unsigned width,
                  unsigned height)
{
  const size_t bytes = (size_t)width * height * sizeof(float);
  if (filters.empty()) {
    input->copy(output, bytes);
    return;
  }
  Buffer* b1 = (filters.size() % 2) ? output : filterBuffer_.get();
  Buffer* b2 = (filters.size() % 2) ? filterBuffer_.get() : output;
  if (filters[0]->enabled())
    filters[0]->apply(input, b1, (int)width, (int)height);
  else
    input->copy(b1, bytes);
  for (size_t k = 1; k < filters.size(); k++) {
    if (filters[k]->enabled())
      filters[k]->apply(b1, b2, (int)width, (int)height);
    else
      b1->copy(b2, bytes);
    std::swap(b1, b2);
  }
}

}  // namespace gpu

// ---------------------------------------------------------------------------
// Tracker
// ---------------------------------------------------------------------------
Tracker::Tracker() : scene_(nullptr), render_width_(0), render_height_(0), current_volume_(0) {}

Tracker::~Tracker()
{
  if (scene_) asc_scene_destroy(scene_);
  for (asc_image im : scene_rads_) asc_image_destroy(im);
}

void Tracker::load(const std::vector<VolumeSpec>& volumes, std::vector<Camera>& cameras,
                   const std::vector<const void*>& rads, const std::vector<std::pair<int, int>>& rad_sizes,
                   int rad_bps, int render_width, int render_height)
{
  if (cameras.size() != rads.size() || rads.size() != rad_sizes.size())
    throw gpu::Error(ASC_ERR_INVALID, "Tracker::load: one radiograph per camera");
  views_.clear();
  volumeDescription_.clear();
  pivots_.clear();
  cameras_ = cameras;
  render_width_ = render_width;
  render_height_ = render_height;
  // Tracker.cpp:281-290: describe each volume and centre its pivot
  for (const VolumeSpec& vs : volumes) {
    volumeDescription_.emplace_back(new gpu::VolumeDescription(*vs.volume, vs.storage));
    if (vs.layout >= 0) check(asc_volume_set_layout(volumeDescription_.back()->image()->volume(), vs.layout));
    CoordFrame pivot = vs.pivot;
    pivot.translate(volumeDescription_.back()->transCenter());
    pivots_.push_back(pivot);
  }
  const size_t bytes = (size_t)render_width * render_height * sizeof(float);
  rendered_drr_.reset(new gpu::Buffer(bytes));
  rendered_rad_.reset(new gpu::Buffer(bytes));
  drr_mask_.reset(new gpu::Buffer(bytes));
  background_mask_.reset(new gpu::Buffer(bytes));
  drr_mask_->fill(1.0f);
  background_mask_->fill(1.0f);
  gpu::ncc_init((unsigned)(render_width * render_height));
  // Tracker.cpp:310-334: one View per camera
  for (size_t i = 0; i < cameras_.size(); i++) {
    std::unique_ptr<gpu::View> view(new gpu::View(cameras_[i]));
    for (size_t v = 0; v < volumeDescription_.size(); v++) {
      view->addDrrRenderer();
      view->drrRenderer((int)v)->setVolume(*volumeDescription_[v]);
    }
    const double* vp = cameras_[i].viewport();
    view->radRenderer()->set_image_plane((float)vp[0], (float)vp[1], (float)vp[2], (float)vp[3]);
    view->radRenderer()->set_rad(rads[i], rad_sizes[i].first, rad_sizes[i].second, rad_bps);
    view->backgroundRenderer()->set_image_plane((float)vp[0], (float)vp[1], (float)vp[2], (float)vp[3]);
    views_.push_back(std::move(view));
  }
  // the fused batch path sees the same device volumes and radiographs
  if (scene_) asc_scene_destroy(scene_);
  for (asc_image im : scene_rads_) asc_image_destroy(im);
  scene_rads_.clear();
  check(asc_scene_create(gpu::context(), render_width, render_height, &scene_));
  for (size_t v = 0; v < volumeDescription_.size(); v++) {
    double pv[12];
    std::memcpy(pv, pivots_[v].rotation(), 9 * sizeof(double));
    std::memcpy(pv + 9, pivots_[v].translation(), 3 * sizeof(double));
    check(asc_scene_add_volume(scene_, volumeDescription_[v]->image()->volume(), 0.5f, 10.0f, 0.0f, pv, nullptr));
  }
  for (size_t i = 0; i < cameras_.size(); i++) {
    asc_image im = nullptr;
    check(asc_image_create(gpu::context(), rads[i], rad_sizes[i].first, rad_sizes[i].second, rad_bps, &im));
    scene_rads_.push_back(im);
    check(asc_scene_add_camera(scene_, &cameras_[i].record(), im, nullptr));
  }
}

void Tracker::setDrrParams(int volume, float step, float intensity, float cutoff)
{
  for (auto& v : views_) {
    v->drrRenderer(volume)->setSampleDistance(step);
    v->drrRenderer(volume)->setRayIntensity(intensity);
    v->drrRenderer(volume)->setCutoff(cutoff);
  }
}

bool Tracker::calculate_viewport(const CoordFrame& modelview, const Camera& camera, double* viewport) const
{
  double mv[12];
  std::memcpy(mv, modelview.rotation(), 9 * sizeof(double));
  std::memcpy(mv + 9, modelview.translation(), 3 * sizeof(double));
  const int rc =
    asc_calculate_viewport(mv, &camera.record(), &volumeDescription_[current_volume_]->image()->desc(), viewport);
  check(rc);
  return rc == 1;
}

std::vector<double> Tracker::trackFrame(unsigned volumeID, const double* xyzypr) const
{
  std::vector<double> correlations;
  const CoordFrame xcframe = CoordFrame::from_xyzypr(xyzypr);
  for (size_t i = 0; i < views_.size(); ++i) {
    gpu::View* view = views_[i].get();
    const CoordFrame modelview = view->camera()->coord_frame().inverse() * xcframe;
    double imv[16];
    modelview.inverse().to_matrix_row_order(imv);
    view->drrRenderer(volumeID)->setInvModelView(imv);
    double viewport[4];
    double mv[12];
    std::memcpy(mv, modelview.rotation(), 9 * sizeof(double));
    std::memcpy(mv + 9, modelview.translation(), 3 * sizeof(double));
    const int inview = asc_calculate_viewport(mv, &view->camera()->record(),
                                              &volumeDescription_[volumeID]->image()->desc(), viewport);
    check(inview);
    if (!inview) {
      correlations.push_back(9999.0);
      continue;
    }
    const unsigned rw = (unsigned)(viewport[2] * render_width_ / view->camera()->viewport()[2]);
    const unsigned rh = (unsigned)(viewport[3] * render_height_ / view->camera()->viewport()[3]);
    if (rw > (unsigned)render_width_ || rh > (unsigned)render_height_)
      throw gpu::Error(ASC_ERR_TOO_LARGE, "Tracker::trackFrame(): Rendered image is larger than the viewport buffer!");
    view->drrRenderer(volumeID)->setViewport((float)viewport[0], (float)viewport[1], (float)viewport[2],
                                             (float)viewport[3]);
    view->radRenderer()->set_viewport((float)viewport[0], (float)viewport[1], (float)viewport[2], (float)viewport[3]);
    view->renderDrrSingle((int)volumeID, rendered_drr_.get(), rw, rh);
    view->renderRad(rendered_rad_.get(), rw, rh);
    view->backgroundRenderer()->set_viewport((float)viewport[0], (float)viewport[1], (float)viewport[2],
                                             (float)viewport[3]);
    view->renderBackground(background_mask_.get(), rw, rh);
    view->renderDRRMask(rendered_drr_.get(), drr_mask_.get(), rw, rh);
    gpu::multiply(background_mask_.get(), drr_mask_.get(), drr_mask_.get(), rw, rh);
    gpu::multiply(rendered_rad_.get(), drr_mask_.get(), rendered_rad_.get(), rw, rh);
    gpu::multiply(rendered_drr_.get(), drr_mask_.get(), rendered_drr_.get(), rw, rh);
    correlations.push_back(1.0 - gpu::ncc(rendered_drr_.get(), rendered_rad_.get(), drr_mask_.get(), rw * rh));
  }
  return correlations;
}

double Tracker::minimizationFunc(const float store[7], const double* values) const
{
  double pv[12];
  std::memcpy(pv, pivots_[current_volume_].rotation(), 9 * sizeof(double));
  std::memcpy(pv + 9, pivots_[current_volume_].translation(), 3 * sizeof(double));
  double xyzypr[6];
  check(asc_compose_pose(store, pv, values, xyzypr));
  const std::vector<double> c = trackFrame((unsigned)current_volume_, xyzypr);
  double total = c[0];
  for (size_t i = 1; i < c.size(); i++) total += c[i];
  if (total < 0) total = 9999;
  return total;
}

void Tracker::sync_scene_filters()
{
  for (size_t i = 0; i < views_.size(); i++) {
    for (int which = 0; which < 2; which++) {
      const std::vector<gpu::Filter*>& fs = which ? views_[i]->radFilters() : views_[i]->drrFilters();
      std::vector<asc_filter> recs;
      for (const gpu::Filter* f : fs) recs.push_back(gpu::filter_record(*f));
      check(asc_scene_set_filters(scene_, (int)i, which, recs.data(), (int)recs.size()));
    }
    gpu::BackgroundRenderer* bg = views_[i]->backgroundRenderer();
    check(asc_scene_set_background(scene_, (int)i, bg->image(), views_[i]->backgroundThreshold()));
  }
  for (size_t v = 0; v < volumeDescription_.size() && !views_.empty(); v++) {
    gpu::RayCaster* rc = views_[0]->drrRenderer((int)v);
    check(asc_scene_set_drr_params(scene_, (int)v, rc->getSampleDistance(), rc->getRayIntensity(), rc->getCutoff()));
  }
}

std::vector<double> Tracker::minimizationFuncBatch(const float store[7], const std::vector<double>& values) const
{
  const int n = (int)(values.size() / 6);
  const_cast<Tracker*>(this)->sync_scene_filters();
  std::vector<float> stores((size_t)n * 7);
  for (int i = 0; i < n; i++) std::memcpy(&stores[(size_t)i * 7], store, 7 * sizeof(float));
  std::vector<int> vols(n, current_volume_);
  std::vector<double> costs(n);
  check(asc_scene_eval_manips(scene_, vols.data(), stores.data(), values.data(), n, costs.data(), nullptr));
  return costs;
}

asc_optimize_stats Tracker::optimize(float store[7], const asc_optimize_options& opt)
{
  sync_scene_filters();
  asc_optimize_stats st;
  check(asc_optimize_frame(scene_, store, &opt, &st));
  return st;
}

}  // namespace xromm
