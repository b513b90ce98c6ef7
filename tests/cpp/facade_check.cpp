// Drives the xromm::gpu facade the way the reference Tracker does, on inputs
// written by tests/test_gpu_facade.py, and prints the per-camera costs of
// the one-pose-at-a-time path (View/RayCaster/filters/ncc) and of the fused
// batch path as JSON lines.
//   facade_check <input.bin>
// With --time N it instead times N trackFrame calls (cycling over the input
// poses, after two untimed ones) and prints the rate; --blocking makes every
// per-operator call wait for its kernel, as the reference's OpenCL backend
// does after each launch.
// --storage and --layout choose the device copy of the volume (default u16,
// the densest layout); --ncc-tree opencl|cuda computes gpu::ncc the
// reference's way (float tree sums, a blocking read after each).
//   facade_check <input.bin> --time N [--blocking] [--storage u16|f16|bf16|f32]
//                [--layout linear|quad|octet|brick] [--ncc-tree opencl|cuda]
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <vector>

#include "../../include/autoscoper_xromm.hpp"

template <class T>
static void rd(std::ifstream& f, T* p, size_t n)
{
  f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
}

int main(int argc, char** argv)
{
  if (argc < 2) return 2;
  int time_n = 0;
  bool blocking = false;
  int storage = ASC_STORAGE_U16, layout = -1, ncc_form = ASC_NCC_MOMENTS;
  static const char* const kStorages[] = {"u16", "f16", "bf16", "f32"};
  static const char* const kLayouts[] = {"linear", "quad", "octet", "brick"};
  auto pick = [](const char* s, const char* const* names, int n) {
    for (int i = 0; i < n; i++)
      if (!std::strcmp(s, names[i])) return i;
    return -2;
  };
  for (int a = 2; a < argc; a++) {
    if (!std::strcmp(argv[a], "--time") && a + 1 < argc)
      time_n = std::atoi(argv[++a]);
    else if (!std::strcmp(argv[a], "--blocking"))
      blocking = true;
    else if (!std::strcmp(argv[a], "--storage") && a + 1 < argc)
      storage = pick(argv[++a], kStorages, 4);
    else if (!std::strcmp(argv[a], "--layout") && a + 1 < argc)
      layout = pick(argv[++a], kLayouts, 4);
    else if (!std::strcmp(argv[a], "--ncc-tree") && a + 1 < argc) {
      const char* m = argv[++a];
      ncc_form = !std::strcmp(m, "opencl") ? ASC_NCC_TREE_OPENCL : !std::strcmp(m, "cuda") ? ASC_NCC_TREE_CUDA : -1;
      if (ncc_form < 0) return 2;
    }
    else
      return 2;
    if (storage < 0 || layout < -1) return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  int32_t hdr[9];
  rd(f, hdr, 9);
  const int W = hdr[0], H = hdr[1], D = hdr[2], ncam = hdr[3], RW = hdr[4], RH = hdr[5], npose = hdr[6];
  const int rw = hdr[7], rh = hdr[8];
  double vox;
  rd(f, &vox, 1);
  std::vector<uint16_t> vol((size_t)W * H * D);
  rd(f, vol.data(), vol.size());
  std::vector<double> cams((size_t)ncam * 30);
  rd(f, cams.data(), cams.size());
  std::vector<uint16_t> rads((size_t)ncam * rw * rh);
  rd(f, rads.data(), rads.size());
  std::vector<double> poses((size_t)npose * 6);
  rd(f, poses.data(), poses.size());
  float drr[3];
  rd(f, drr, 3);
  if (!f) return 3;
  try {
    xromm::Volume volume(vol.data(), W, H, D, 16, vox, vox, vox);
    std::vector<xromm::Camera> cameras;
    for (int c = 0; c < ncam; c++) {
      asc_camera rec;
      std::memcpy(&rec, &cams[(size_t)c * 30], sizeof rec);
      cameras.emplace_back(rec);
    }
    xromm::Tracker tracker;
    std::vector<const void*> rp;
    std::vector<std::pair<int, int>> rs;
    for (int c = 0; c < ncam; c++) {
      rp.push_back(&rads[(size_t)c * rw * rh]);
      rs.emplace_back(rw, rh);
    }
    xromm::Tracker::VolumeSpec vs{&volume, xromm::CoordFrame(), storage, layout};
    tracker.load({vs}, cameras, rp, rs, 16, RW, RH);
    tracker.setDrrParams(0, drr[0], drr[1], drr[2]);
    for (size_t i = 0; i < tracker.numViews(); i++) {
      auto* s = new xromm::gpu::SobelFilter();
      s->set_scale(1.7f);
      s->set_blend(0.1f);
      tracker.view((int)i)->drrFilters().push_back(s);
      auto* s2 = new xromm::gpu::SobelFilter();
      s2->set_scale(3.0f);
      s2->set_blend(0.43f);
      auto* c2 = new xromm::gpu::ContrastFilter();
      c2->set_alpha(2.5f);
      c2->set_beta(0.9f);
      tracker.view((int)i)->radFilters().push_back(s2);
      tracker.view((int)i)->radFilters().push_back(c2);
    }
    if (time_n > 0) {
      if (blocking && asc_context_set_blocking(xromm::gpu::context(), 1) != 0)
        throw xromm::gpu::Error(-1, asc_last_error());
      if (asc_context_set_ncc_form(xromm::gpu::context(), ncc_form) != 0)
        throw xromm::gpu::Error(-1, asc_last_error());
      double sink = 0;
      for (int i = 0; i < 2; i++) sink += tracker.trackFrame(0, &poses[(size_t)(i % npose) * 6])[0];
      const auto t0 = std::chrono::steady_clock::now();
      for (int i = 0; i < time_n; i++) sink += tracker.trackFrame(0, &poses[(size_t)(i % npose) * 6])[0];
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("{\"evals\": %d, \"seconds\": %.9g, \"evals_per_s\": %.9g, \"blocking\": %s, \"storage\": \"%s\", \"layout\": %d, \"ncc_form\": %d, \"sink\": %.9g}\n",
                  time_n, s, time_n / s, blocking ? "true" : "false", kStorages[storage], layout, ncc_form, sink);
      return 0;
    }
    // fused batch of all poses (absolute poses through a zero manipulator)
    std::vector<double> fused((size_t)npose * ncam);
    std::vector<int> vols(npose, 0);
    std::vector<double> tmp(npose);
    tracker.minimizationFuncBatch(std::vector<float>(7, 0.0f).data(), std::vector<double>(6, 0.0));  // sync filters
    if (asc_scene_eval_poses(tracker.scene(), vols.data(), poses.data(), npose, tmp.data(), fused.data()) != 0)
      throw xromm::gpu::Error(-1, asc_last_error());
    for (int i = 0; i < npose; i++) {
      const std::vector<double> c = tracker.trackFrame(0, &poses[(size_t)i * 6]);
      std::printf("{\"pose\": %d, \"plain\": [", i);
      for (int k = 0; k < ncam; k++) std::printf("%s%.17g", k ? ", " : "", c[k]);
      std::printf("], \"fused\": [");
      for (int k = 0; k < ncam; k++) std::printf("%s%.17g", k ? ", " : "", fused[(size_t)i * ncam + k]);
      std::printf("]}\n");
    }
    // View::render composite (display path) of the last pose
    xromm::gpu::Buffer rgb((size_t)RW * RH * 3 * sizeof(float));
    tracker.view(0)->render(&rgb, 32, 24);
    std::vector<float> host(32 * 24 * 3);
    rgb.write(host.data(), host.size() * sizeof(float));
    double sum = 0;
    for (float v : host) sum += v;
    std::printf("{\"composite_sum\": %.9g}\n", sum);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "facade_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
