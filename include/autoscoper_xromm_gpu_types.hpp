// autoscoper_xromm_gpu_types.hpp -- the device types of the xromm::gpu
// backend (what libautoscoper/src/gpu/opencl/OpenCL.hpp declares): the
// process-wide context and device selection, Buffer, GLBuffer and Image.
//
// libautoscoper's own Filter.hpp, View.hpp and VolumeDescription.hpp include
// this header at their backend include sites (they need Buffer, GLBuffer and
// Image only); Tracker.hpp, Tracker.cpp and View.cpp include
// autoscoper_xromm_gpu.hpp, which includes this one and those library
// headers in turn.
#pragma once

#include <stddef.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "autoscoper_hip.h"

#define XROMM_API __attribute__((visibility("default")))

namespace xromm {
namespace gpu {

struct XROMM_API Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// The process-wide context of the backend (the reference keeps one global
// OpenCL context and queue, OpenCL.cpp:83-102). Created on first use on the
// selected device (0 unless set_device() / setUsedPlatform() chose another).
XROMM_API asc_context context();
XROMM_API void set_device(int device);

// Device selection as the reference GUI drives it (OpenCL.hpp:63-66): one
// platform ("HIP"), whose devices are the visible HIP devices.
XROMM_API std::vector<std::vector<std::string>> get_platforms();
XROMM_API void setUsedPlatform(int platform_idx);
XROMM_API void setUsedPlatform(int platform, int device);
XROMM_API std::pair<int, int> getUsedPlatform();

// Device buffer (OpenCL.hpp:124-141). `access` is accepted for signature
// compatibility (cl_mem_flags there) and ignored: device memory is
// read-write.
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

// GLBuffer (OpenCL.hpp:143-153): in the reference, a CL view of the GL
// pixel buffer the display draws. This backend has no GL interop: a
// GLBuffer is a device buffer of the same floats (View::render composites
// into it) that the host copies into the pixel buffer.
class XROMM_API GLBuffer : public Buffer {
 public:
  using Buffer::Buffer;
};

// Image (OpenCL.hpp:155-165): the device copy of a volume, as
// VolumeDescription holds it (its image()). Owns the asc_volume.
class XROMM_API Image {
 public:
  Image(asc_volume volume, const asc_volume_desc& desc) : vol_(volume), desc_(desc) {}
  ~Image();
  Image(const Image&) = delete;
  Image& operator=(const Image&) = delete;
  asc_volume volume() const { return vol_; }
  const asc_volume_desc& desc() const { return desc_; }

 private:
  asc_volume vol_;
  asc_volume_desc desc_;
};

}  // namespace gpu
}  // namespace xromm
