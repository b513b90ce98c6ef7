// On-disk inputs and outputs of a tracking trial, host code only:
//   * TIFF images and stacks (volumes: one page per slice, Volume.cpp:51-140;
//     radiograph frames: one file per frame, Video.cpp:66-239), baseline
//     strips, uncompressed / PackBits / LZW, horizontal predictor;
//   * camera calibrations: MayaCam 1.0, MayaCam 2.0, VTKCam 1.0
//     (Camera.cpp:111-422);
//   * filter settings (.vie) as the GUI applies them
//     (FilterTreeWidget.cpp:290-329, FilterTreeWidgetItem.cpp:148-195,
//     ModelViewTreeWidgetItem.cpp:126-148,321-344);
//   * trial configurations (.cfg v1.0-1.2, Trial.cpp:74-257);
//   * tracking results (.tra, AutoscoperMainWindow.cpp:736-1092).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <sys/stat.h>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace asc {
namespace {

int io_fail(const std::string& msg) { return fail(ASC_ERR_IO, msg); }

// The entry points return a status whatever happens inside: a failed
// allocation or a library exception becomes an error, never an abort.
template <class F>
int guarded(F&& f)
{
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(ASC_ERR_NOMEM, "out of host memory");
  } catch (const std::exception& e) {
    return fail(ASC_ERR_IO, e.what());
  }
}

bool read_file(const std::string& path, std::vector<uint8_t>& out)
{
  struct stat st;
  if (::stat(path.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return false;  // not a directory
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)n);
  const size_t got = n > 0 ? std::fread(out.data(), 1, out.size(), f) : 0;
  std::fclose(f);
  return got == out.size();
}

// A line without its terminator ('\n', and a '\r' before it), as
// kwsys::SystemTools::GetLineFromStream returns it.
bool get_line(std::istream& in, std::string& line)
{
  if (!std::getline(in, line)) return false;
  if (!line.empty() && line.back() == '\r') line.pop_back();
  return true;
}

// ---------------------------------------------------------------------------
// TIFF (TIFF 6.0 baseline + LZW + PackBits + predictor 2)
// ---------------------------------------------------------------------------
struct TiffPage {
  uint32_t width = 0, height = 0;
  uint16_t bps = 1, spp = 1, compression = 1, photometric = 1, planar = 1, predictor = 1, sample_format = 1;
  uint32_t rows_per_strip = 0xFFFFFFFFu;
  std::vector<uint64_t> strip_offsets, strip_bytes;
  bool tiled = false;
};

struct TiffFile {
  std::vector<uint8_t> bytes;
  bool big_endian = false;
  std::vector<TiffPage> pages;

  uint16_t u16(size_t o) const
  {
    return big_endian ? (uint16_t)(bytes[o] << 8 | bytes[o + 1]) : (uint16_t)(bytes[o] | bytes[o + 1] << 8);
  }
  uint32_t u32(size_t o) const
  {
    return big_endian ? (uint32_t)bytes[o] << 24 | (uint32_t)bytes[o + 1] << 16 | (uint32_t)bytes[o + 2] << 8 | bytes[o + 3]
                      : (uint32_t)bytes[o] | (uint32_t)bytes[o + 1] << 8 | (uint32_t)bytes[o + 2] << 16 |
                          (uint32_t)bytes[o + 3] << 24;
  }
};

// Values of one IFD entry (SHORT or LONG, inline or at an offset).
bool tiff_values(const TiffFile& t, size_t entry, std::vector<uint64_t>& out)
{
  const uint16_t type = t.u16(entry + 2);
  const uint32_t count = t.u32(entry + 4);
  const size_t size = type == 3 ? 2 : type == 4 ? 4 : type == 1 ? 1 : 0;
  if (!size) return false;
  const uint64_t total = (uint64_t)size * count;
  size_t at = entry + 8;
  if (total > 4) at = t.u32(entry + 8);
  if (at + total > t.bytes.size()) return false;
  out.resize(count);
  for (uint32_t i = 0; i < count; i++)
    out[i] = size == 2 ? t.u16(at + 2 * i) : size == 4 ? t.u32(at + 4 * i) : t.bytes[at + i];
  return true;
}

int tiff_open(const std::string& path, TiffFile& t)
{
  if (!read_file(path, t.bytes)) return io_fail("Unable to open TIFF file: " + path);
  const auto& b = t.bytes;
  if (b.size() < 8) return io_fail("Not a TIFF file: " + path);
  if (b[0] == 'I' && b[1] == 'I')
    t.big_endian = false;
  else if (b[0] == 'M' && b[1] == 'M')
    t.big_endian = true;
  else
    return io_fail("Not a TIFF file: " + path);
  if (t.u16(2) != 42) return io_fail("Not a classic TIFF file (BigTIFF is not supported): " + path);
  size_t ifd = t.u32(4);
  while (ifd != 0) {
    if (ifd + 2 > b.size() || t.pages.size() > 1000000) return io_fail("Corrupt TIFF directory in " + path);
    const uint16_t n = t.u16(ifd);
    if (ifd + 2 + 12 * (size_t)n + 4 > b.size()) return io_fail("Corrupt TIFF directory in " + path);
    TiffPage p;
    std::vector<uint64_t> v;
    for (uint16_t e = 0; e < n; e++) {
      const size_t entry = ifd + 2 + 12 * (size_t)e;
      const uint16_t tag = t.u16(entry);
      if (tag == 322 || tag == 323 || tag == 324 || tag == 325) {
        p.tiled = true;
        continue;
      }
      if (tag != 256 && tag != 257 && tag != 258 && tag != 259 && tag != 262 && tag != 273 && tag != 277 &&
          tag != 278 && tag != 279 && tag != 284 && tag != 317 && tag != 339)
        continue;
      if (!tiff_values(t, entry, v) || v.empty()) return io_fail("Corrupt TIFF tag in " + path);
      switch (tag) {
        case 256: p.width = (uint32_t)v[0]; break;
        case 257: p.height = (uint32_t)v[0]; break;
        case 258: p.bps = (uint16_t)v[0]; break;
        case 259: p.compression = (uint16_t)v[0]; break;
        case 262: p.photometric = (uint16_t)v[0]; break;
        case 273: p.strip_offsets = v; break;
        case 277: p.spp = (uint16_t)v[0]; break;
        case 278: p.rows_per_strip = (uint32_t)v[0]; break;
        case 279: p.strip_bytes = v; break;
        case 284: p.planar = (uint16_t)v[0]; break;
        case 317: p.predictor = (uint16_t)v[0]; break;
        case 339: p.sample_format = (uint16_t)v[0]; break;
      }
    }
    t.pages.push_back(p);
    ifd = t.u32(ifd + 2 + 12 * (size_t)n);
  }
  if (t.pages.empty()) return io_fail("TIFF file without images: " + path);
  return ASC_OK;
}

// PackBits (TIFF 6.0 section 9): header n in [0, 127] copies n + 1 bytes,
// [-127, -1] repeats the next byte 1 - n times, -128 is a no-op.
bool unpack_bits(const uint8_t* src, size_t n, uint8_t* dst, size_t want)
{
  size_t i = 0, o = 0;
  while (o < want && i < n) {
    const int h = (int8_t)src[i++];
    if (h >= 0) {
      const size_t c = (size_t)h + 1;
      if (i + c > n || o + c > want) return false;
      std::memcpy(dst + o, src + i, c);
      i += c;
      o += c;
    } else if (h != -128) {
      const size_t c = (size_t)(1 - h);
      if (i >= n || o + c > want) return false;
      std::memset(dst + o, src[i++], c);
      o += c;
    }
  }
  return o == want;
}

// LZW (TIFF 6.0 section 13): MSB-first codes of 9 to 12 bits, 256 = Clear,
// 257 = EndOfInformation, and the code width grows one code early (after
// entry 2^w - 2 has been added, i.e. when the next free entry is 2^w - 1).
bool unlzw(const uint8_t* src, size_t n, uint8_t* dst, size_t want)
{
  constexpr int kClear = 256, kEoi = 257;
  std::vector<uint32_t> prefix(4096), first(4096);
  std::vector<uint16_t> length(4096);
  std::vector<uint8_t> suffix(4096);
  for (int c = 0; c < 256; c++) {
    prefix[c] = 0xFFFFFFFFu;
    suffix[c] = (uint8_t)c;
    first[c] = (uint32_t)c;
    length[c] = 1;
  }
  size_t bitpos = 0, o = 0;
  int width = 9, next = 258, old = -1;
  auto read_code = [&](int& code) {
    if (bitpos + width > 8 * n) return false;
    uint32_t v = 0;
    for (int k = 0; k < width; k++) {
      const size_t b = bitpos + k;
      v = v << 1 | ((src[b >> 3] >> (7 - (b & 7))) & 1u);
    }
    bitpos += width;
    code = (int)v;
    return true;
  };
  // writes the string of `code` at dst + o (its bytes are found by walking
  // the prefix chain from the last byte back)
  auto emit = [&](int code) {
    const size_t len = length[code];
    if (o + len > want) return false;
    int c = code;
    for (size_t k = len; k-- > 0;) {
      dst[o + k] = suffix[c];
      c = (int)prefix[c];
    }
    o += len;
    return true;
  };
  int code;
  while (o < want && read_code(code)) {
    if (code == kEoi) break;
    if (code == kClear) {
      width = 9;
      next = 258;
      if (!read_code(code)) break;
      if (code == kEoi) break;
      if (code >= 256) return false;
      if (!emit(code)) return false;
      old = code;
      continue;
    }
    if (old < 0) return false;
    if (code < next) {
      if (!emit(code)) return false;
      if (next < 4096) {
        prefix[next] = (uint32_t)old;
        suffix[next] = (uint8_t)first[code];
        first[next] = first[old];
        length[next] = (uint16_t)(length[old] + 1);
        next++;
      }
    } else if (code == next && next < 4096) {
      // KwKwK: the new entry is old's string plus old's first byte
      prefix[next] = (uint32_t)old;
      suffix[next] = (uint8_t)first[old];
      first[next] = first[old];
      length[next] = (uint16_t)(length[old] + 1);
      next++;
      if (!emit(code)) return false;
    } else {
      return false;
    }
    old = code;
    if (next + 1 >= (1 << width) && width < 12) width++;
  }
  return o == want;
}

int tiff_decode_page(const TiffFile& t, const TiffPage& p, const std::string& path, uint8_t* dst)
{
  if (p.tiled) return fail(ASC_ERR_UNSUPPORTED, "Tiled TIFF images are not supported: " + path);
  if (p.planar != 1) return fail(ASC_ERR_UNSUPPORTED, "Planar TIFF data is not supported: " + path);
  if (p.bps != 8 && p.bps != 16 && p.bps != 32)
    return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF bits per sample " + std::to_string(p.bps) + ": " + path);
  if (p.compression != 1 && p.compression != 5 && p.compression != 32773)
    return fail(ASC_ERR_UNSUPPORTED,
                "Unsupported TIFF compression " + std::to_string(p.compression) + " (1, 5 and 32773 are read): " + path);
  if (p.predictor != 1 && p.predictor != 2)
    return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF predictor " + std::to_string(p.predictor) + ": " + path);
  const size_t bpp = (size_t)p.bps / 8 * p.spp;
  const size_t row = (size_t)p.width * bpp;
  const uint32_t rps = std::min<uint32_t>(p.rows_per_strip, p.height);
  if (rps == 0 && p.height) return io_fail("Corrupt TIFF strips in " + path);
  const size_t nstrips = p.height ? (p.height + rps - 1) / rps : 0;
  if (p.strip_offsets.size() < nstrips || (p.compression != 1 && p.strip_bytes.size() < nstrips))
    return io_fail("Corrupt TIFF strips in " + path);
  for (size_t s = 0; s < nstrips; s++) {
    const uint32_t rows = std::min<uint32_t>(rps, p.height - (uint32_t)(s * rps));
    const size_t want = rows * row;
    const uint64_t off = p.strip_offsets[s];
    const uint64_t len = s < p.strip_bytes.size() ? p.strip_bytes[s] : want;
    if (off + len > t.bytes.size()) return io_fail("TIFF strip beyond the end of " + path);
    uint8_t* out = dst + s * rps * row;
    const uint8_t* in = t.bytes.data() + off;
    bool ok;
    if (p.compression == 1) {
      ok = len >= want;
      if (ok) std::memcpy(out, in, want);
    } else if (p.compression == 32773) {
      ok = unpack_bits(in, (size_t)len, out, want);
    } else {
      ok = unlzw(in, (size_t)len, out, want);
    }
    if (!ok) return io_fail("Corrupt compressed TIFF strip in " + path);
  }
  // samples to native (little-endian) order, then undo the predictor
  const size_t n = (size_t)p.height * row;
  if (t.big_endian && p.bps == 16)
    for (size_t i = 0; i + 1 < n; i += 2) std::swap(dst[i], dst[i + 1]);
  if (t.big_endian && p.bps == 32)
    for (size_t i = 0; i + 3 < n; i += 4) {
      std::swap(dst[i], dst[i + 3]);
      std::swap(dst[i + 1], dst[i + 2]);
    }
  if (p.predictor == 2) {
    for (uint32_t y = 0; y < p.height; y++) {
      uint8_t* r = dst + y * row;
      const size_t spr = (size_t)p.width * p.spp;
      if (p.bps == 8) {
        for (size_t i = p.spp; i < spr; i++) r[i] = (uint8_t)(r[i] + r[i - p.spp]);
      } else if (p.bps == 16) {
        uint16_t* q = (uint16_t*)r;
        for (size_t i = p.spp; i < spr; i++) q[i] = (uint16_t)(q[i] + q[i - p.spp]);
      } else {
        uint32_t* q = (uint32_t*)r;
        for (size_t i = p.spp; i < spr; i++) q[i] = q[i] + q[i - p.spp];
      }
    }
  }
  return ASC_OK;
}

// All pages of a TIFF file of uniform size, as one array pages x H x W.
int tiff_read_all(const std::string& path, std::vector<uint8_t>& out, int dims[3], int* bps, bool volume)
{
  TiffFile t;
  int rc = tiff_open(path, t);
  if (rc) return rc;
  const TiffPage& p0 = t.pages[0];
  if (volume && p0.spp != 1)
    return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF image format. Invalid 'SamplesPerPixel' value in " + path +
                                       ". Current value is " + std::to_string(p0.spp) + ", expected 1");
  if (volume && p0.sample_format != 1)
    return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF image format. Invalid 'SampleFormat' value in " + path +
                                       ". Current value is " + std::to_string(p0.sample_format) + ", expected 1");
  if (p0.width > (1u << 20) || p0.height > (1u << 20) || p0.spp > 4)
    return io_fail("Unsupported TIFF image size in " + path);
  const size_t page_bytes = (size_t)p0.width * p0.height * (p0.bps / 8) * p0.spp;
  const size_t npages = volume ? t.pages.size() : 1;
  // the pixels must fit in the file: uncompressed strips hold them as they
  // are, and no PackBits or LZW strip expands by more than 2048 times
  uint64_t stored = 0;
  for (size_t i = 0; i < npages; i++) {
    const TiffPage& p = t.pages[i];
    uint64_t bytes = 0;
    for (uint64_t b : p.strip_bytes) bytes += std::min<uint64_t>(b, t.bytes.size());
    stored += p.compression == 1 ? std::min<uint64_t>(bytes ? bytes : t.bytes.size(), t.bytes.size())
                                 : 2048 * std::min<uint64_t>(bytes, t.bytes.size());
  }
  if ((uint64_t)page_bytes * npages > stored) return io_fail("TIFF pixel data missing or truncated in " + path);
  for (size_t i = 0; i < npages; i++) {
    const TiffPage& p = t.pages[i];
    if (p.width != p0.width || p.height != p0.height || p.bps != p0.bps || p.spp != p0.spp)
      return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF image format. Non uniform volume slices in " + path +
                                         ". Slice " + std::to_string(i) + " has dimensions " +
                                         std::to_string(p.width) + "x" + std::to_string(p.height) + "x" +
                                         std::to_string(p.bps));
  }
  out.assign(page_bytes * npages, 0);
  for (size_t i = 0; i < npages; i++)
    if ((rc = tiff_decode_page(t, t.pages[i], path, out.data() + i * page_bytes))) return rc;
  dims[0] = (int)p0.width;
  dims[1] = (int)p0.height;
  dims[2] = (int)npages;
  *bps = p0.bps;
  return ASC_OK;
}

int tiff_info(const std::string& path, int dims[3], int* bps, bool volume)
{
  TiffFile t;
  int rc = tiff_open(path, t);
  if (rc) return rc;
  // what the reader can decode, so that the sizes reported describe it
  const TiffPage& p = t.pages[0];
  if (p.tiled) return fail(ASC_ERR_UNSUPPORTED, "Tiled TIFF images are not supported: " + path);
  if (p.bps != 8 && p.bps != 16 && p.bps != 32)
    return fail(ASC_ERR_UNSUPPORTED, "Unsupported TIFF bits per sample " + std::to_string(p.bps) + ": " + path);
  if (p.width == 0 || p.height == 0 || p.width > (1u << 20) || p.height > (1u << 20))
    return io_fail("Unsupported TIFF image size in " + path);
  dims[0] = (int)t.pages[0].width;
  dims[1] = (int)t.pages[0].height;
  dims[2] = volume ? (int)t.pages.size() : 1;
  *bps = t.pages[0].bps;
  return ASC_OK;
}

// ---------------------------------------------------------------------------
// Cameras
// ---------------------------------------------------------------------------
// The comma-separated values of one line read as the reference does:
// std::getline(.., ',') then operator>> on each piece.
int csv_doubles(const std::string& line, double* out, int n)
{
  std::istringstream ls(line);
  std::string piece;
  int k = 0;
  for (; k < n && std::getline(ls, piece, ','); ++k) {
    std::istringstream ps(piece);
    if (!(ps >> out[k])) break;
  }
  return k;
}

std::string maya_error(const char* version, int line, const std::string& path, const std::string& msg)
{
  return std::string("Invalid MayaCam ") + version + ".0 file. " + msg + " See line " + std::to_string(line) +
         " in " + path + ".";
}

// Camera::loadMayaCam1 after parsing (Camera.cpp:185-227): v = the five
// lines {x, y, z}, {rx, ry, rz}, {film plane, unused}, {u0, v0, Z},
// {scale, width, height}.
int camera_mayacam1(const double* v, asc_camera& c)
{
  std::memset(&c, 0, sizeof c);
  const double u0 = v[9] - 1, v0 = v[10] - 1, z = v[11];
  c.size[0] = v[13] == 0 ? 1024 : v[13];
  c.size[1] = v[14] == 0 ? 1024 : v[14];
  const double xyzypr[6] = {v[0], v[1], v[2], v[5], v[4], v[3]};
  const Frame f = frame_from_xyzypr(xyzypr);
  frame_store(f, c.rotation);
  return camera_finish(c, u0, v0, z < 0 ? -z : z, z < 0 ? -z : z, z);
}

// Camera::calculateLookAtMatrix (Camera.cpp:495-512).
void look_at(const double* eye, const double* centre, const double* up, double* m)
{
  auto unit = [](double* a) {
    const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int i = 0; i < 3; i++) a[i] /= l;
  };
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double fwd[3] = {centre[0] - eye[0], centre[1] - eye[1], centre[2] - eye[2]};
  unit(fwd);
  double side[3];
  cross(fwd, up, side);
  unit(side);
  double pu[3];
  cross(side, fwd, pu);
  for (int i = 0; i < 3; i++) {
    m[i] = side[i];
    m[3 + i] = pu[i];
    m[6 + i] = -fwd[i];
  }
}

// Camera::loadVTKCamera after parsing (Camera.cpp:364-422).
int camera_vtk(const double* pos, const double* focal, const double* up, double view_angle, const double* size,
               asc_camera& c)
{
  std::memset(&c, 0, sizeof c);
  c.size[0] = size[0];
  c.size[1] = size[1];
  look_at(pos, focal, up, c.rotation);
  for (int i = 0; i < 3; i++) c.translation[i] = pos[i];
  const double a = view_angle * (M_PI / 180);
  const double fx = c.size[0] / (2 * std::tan(a / 2)), fy = c.size[1] / (2 * std::tan(a / 2));
  return camera_finish(c, c.size[0] / 2.0, c.size[1] / 2.0, fx, fy, -0.5 * (fx + fy));
}

// A small JSON reader: enough for a VTKCam file (objects, arrays, numbers,
// strings, true/false/null).
struct Json {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;
};

struct JsonParser {
  const std::string& s;
  size_t i = 0;
  explicit JsonParser(const std::string& text) : s(text) {}
  void ws()
  {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
  }
  bool lit(const char* w)
  {
    const size_t n = std::strlen(w);
    if (s.compare(i, n, w) != 0) return false;
    i += n;
    return true;
  }
  bool string(std::string& out)
  {
    if (i >= s.size() || s[i] != '"') return false;
    i++;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) {
        const char e = s[++i];
        out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        i++;
      } else {
        out += s[i++];
      }
    }
    if (i >= s.size()) return false;
    i++;
    return true;
  }
  bool value(Json& v, int depth = 0)
  {
    if (depth > 64) return false;
    ws();
    if (i >= s.size()) return false;
    const char ch = s[i];
    if (ch == '{') {
      v.kind = Json::kObj;
      i++;
      ws();
      if (i < s.size() && s[i] == '}') {
        i++;
        return true;
      }
      for (;;) {
        ws();
        std::string key;
        if (!string(key)) return false;
        ws();
        if (i >= s.size() || s[i] != ':') return false;
        i++;
        if (!value(v.obj[key], depth + 1)) return false;
        ws();
        if (i < s.size() && s[i] == ',') {
          i++;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          i++;
          return true;
        }
        return false;
      }
    }
    if (ch == '[') {
      v.kind = Json::kArr;
      i++;
      ws();
      if (i < s.size() && s[i] == ']') {
        i++;
        return true;
      }
      for (;;) {
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) return false;
        ws();
        if (i < s.size() && s[i] == ',') {
          i++;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          i++;
          return true;
        }
        return false;
      }
    }
    if (ch == '"') {
      v.kind = Json::kStr;
      return string(v.str);
    }
    if (lit("true")) {
      v.kind = Json::kBool;
      v.num = 1;
      return true;
    }
    if (lit("false")) {
      v.kind = Json::kBool;
      return true;
    }
    if (lit("null")) return true;
    char* end = nullptr;
    v.num = std::strtod(s.c_str() + i, &end);
    if (end == s.c_str() + i) return false;
    v.kind = Json::kNum;
    i = (size_t)(end - s.c_str());
    return true;
  }
};

int read_vtkcam(const std::string& path, asc_camera& c)
{
  std::vector<uint8_t> b;
  if (!read_file(path, b)) return io_fail("Error opening VTKCam file: " + path);
  const std::string text(b.begin(), b.end());
  Json root;
  JsonParser jp(text);
  if (!jp.value(root) || root.kind != Json::kObj) return io_fail("Invalid VTKCam 1.0 file (not a JSON object): " + path);
  auto get = [&](const char* k) -> const Json* {
    auto it = root.obj.find(k);
    return it == root.obj.end() ? nullptr : &it->second;
  };
  const Json* schema = get("@schema");
  const char* url = "https://autoscoperm.slicer.org/vtkCamera-schema-1.0.json";
  if (!schema || schema->kind != Json::kStr || schema->str != url)
    return io_fail("Invalid VTKCam 1.0 file. Unsupported schema in " + path);
  const Json* ver = get("version");
  if (!ver || ver->kind != Json::kNum || ver->num != 1.0)
    return io_fail("Invalid VTKCam 1.0 file. Unsupported version number in " + path);
  auto num = [&](const char* k, double& out) {
    const Json* j = get(k);
    if (!j || j->kind != Json::kNum) return false;
    out = j->num;
    return true;
  };
  auto vec3 = [&](const char* k, double* out) {
    const Json* j = get(k);
    if (!j || j->kind != Json::kArr || j->arr.size() < 3) return false;
    for (int i = 0; i < 3; i++) {
      if (j->arr[i].kind != Json::kNum) return false;
      out[i] = j->arr[i].num;
    }
    return true;
  };
  double size[2], pos[3], focal[3], up[3], angle;
  if (!num("image-width", size[0]) || !num("image-height", size[1]) || !vec3("camera-position", pos) ||
      !vec3("focal-point", focal) || !vec3("view-up", up) || !num("view-angle", angle))
    return io_fail("Invalid VTKCam 1.0 file. Missing or malformed keys in " + path);
  return camera_vtk(pos, focal, up, angle, size, c);
}

int read_mayacam1(const std::string& path, asc_camera& c)
{
  std::ifstream f(path);
  double v[15] = {0};
  std::string line;
  int lines = 0;
  for (int i = 0; i < 5 && get_line(f, line); ++i) {
    if (csv_doubles(line, v + 3 * i, 3) != 3)
      return io_fail(maya_error("1", i + 1, path, "There was an error reading values."));
    lines++;
  }
  if (lines != 5) return io_fail(maya_error("1", 1, path, "There was an error reading values."));
  return camera_mayacam1(v, c);
}

int read_mayacam2(const std::string& path, asc_camera& c)
{
  std::ifstream f(path);
  double K[9] = {0}, R[9] = {0}, t[3] = {0}, size[2] = {0, 0};
  int t_count = 0;
  std::string line;
  for (int i = 0; i < 17 && get_line(f, line); ++i) {
    if (i == 1) {
      if (csv_doubles(line, size, 2) != 2)
        return io_fail(maya_error("2", i + 1, path, "There was an error reading 'image size' values."));
    } else if (i >= 4 && i <= 6) {
      if (csv_doubles(line, K + 3 * (i - 4), 3) != 3)
        return io_fail(maya_error("2", i + 1, path, "There was an error reading 'camera matrix' values."));
    } else if (i >= 9 && i <= 11) {
      if (csv_doubles(line, R + 3 * (i - 9), 3) != 3)
        return io_fail(maya_error("2", i + 1, path, "There was an error reading 'rotation' values."));
    } else if (i >= 14 && i <= 16) {
      std::istringstream ls(line);
      if (!(ls >> t[i - 14]))
        return io_fail(maya_error("2", i + 1, path, "There was an error reading 'translation' values."));
      t_count++;
    }
  }
  if (t_count != 3) return io_fail(maya_error("2", 15, path, "There was an error reading 'translation' values."));
  return asc_camera_from_mayacam2(K, R, t, size, &c);
}

int read_camera(const std::string& path, asc_camera& c)
{
  const size_t dot = path.find_last_of('.');
  if (dot == std::string::npos) return io_fail("Invalid MayaCam file: " + path);
  if (path.compare(dot + 1, std::string::npos, "json") == 0) return read_vtkcam(path, c);
  std::ifstream f(path);
  if (!f.is_open()) return io_fail("File not found: " + path);
  std::string first;
  get_line(f, first);
  f.close();
  return first == "image size" ? read_mayacam2(path, c) : read_mayacam1(path, c);
}

// ---------------------------------------------------------------------------
// Filter settings (.vie)
// ---------------------------------------------------------------------------
// QDoubleSpinBox::setValue with the default 2 decimals: the value is rounded
// to 2 decimals (Qt formats it with 'f', 2, where a remainder of half a unit
// or more rounds away from zero, and parses it back) and then bounded to
// [lo, hi].
double spinbox(double v, double lo, double hi)
{
  char buf[512];
  std::snprintf(buf, sizeof buf, "%.400f", std::fabs(v));  // the exact binary value, in decimal
  const char* dot = std::strchr(buf, '.');
  long long cents = std::atoll(buf) * 100 + (dot[1] - '0') * 10 + (dot[2] - '0');
  if (dot[3] >= '5') cents++;
  std::snprintf(buf, sizeof buf, "%s%lld.%02lld", v < 0 ? "-" : "", cents / 100, cents % 100);
  const double r = std::strtod(buf, nullptr);
  return std::min(std::max(r, lo), hi);
}

int read_filters(std::istream& f, const char* end_tag, std::vector<asc_filter>& out)
{
  std::string line;
  while (get_line(f, line) && line != "DrrFilters_end" && line != "RadFilters_end") {
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    int type = -1;
    if (key == "SobelFilter_begin")
      type = ASC_FILTER_SOBEL;
    else if (key == "ContrastFilter_begin")
      type = ASC_FILTER_CONTRAST;
    else if (key == "GaussianFilter_begin")
      type = ASC_FILTER_GAUSSIAN;
    else if (key == "SharpenFilter_begin")
      type = ASC_FILTER_SHARPEN;
    if (type < 0) continue;
    // GUI spinbox values: Sobel {Blend 0.5 in [0, 1], Scale 1 in [0, 10]},
    // Contrast {Alpha 1, Beta 1 in [0, 10]}, Gaussian {Radius 1 in [0, 10]},
    // Sharpen {Radius 1, Contrast 1 in [0, 10]} (FilterTreeWidgetItem.cpp:83-104)
    double p0 = type == ASC_FILTER_SOBEL ? 0.5 : 1.0, p1 = 1.0;
    const double hi0 = type == ASC_FILTER_SOBEL ? 1.0 : 10.0;
    std::string fl;
    while (get_line(f, fl) && fl != "SobelFilter_end" && fl != "ContrastFilter_end" && fl != "GaussianFilter_end" &&
           fl != "SharpenFilter_end") {
      std::istringstream ps(fl);
      std::string k;
      ps >> k;
      float value = 0.0f;
      // the key `Scale` sets the Sobel filter's first spinbox (blend) and
      // `Blend` the second (scale): FilterTreeWidgetItem.cpp:158-165
      if (type == ASC_FILTER_SOBEL && k == "Scale" && (ps >> value)) p0 = spinbox(value, 0.0, hi0);
      if (type == ASC_FILTER_SOBEL && k == "Blend" && (ps >> value)) p1 = spinbox(value, 0.0, 10.0);
      if (type == ASC_FILTER_CONTRAST && k == "Alpha" && (ps >> value)) p0 = spinbox(value, 0.0, hi0);
      if (type == ASC_FILTER_CONTRAST && k == "Beta" && (ps >> value)) p1 = spinbox(value, 0.0, 10.0);
      if (type == ASC_FILTER_GAUSSIAN && k == "Radius" && (ps >> value)) p0 = spinbox(value, 0.0, hi0);
      if (type == ASC_FILTER_SHARPEN && k == "Radius" && (ps >> value)) p0 = spinbox(value, 0.0, hi0);
      if (type == ASC_FILTER_SHARPEN && k == "Contrast" && (ps >> value)) p1 = spinbox(value, 0.0, 10.0);
    }
    asc_filter r;
    std::memset(&r, 0, sizeof r);
    r.type = type;
    r.enabled = 1;
    switch (type) {
      case ASC_FILTER_SOBEL:  // updateFilter: setBlend(param 0), setScale(param 1)
        r.p[0] = (float)p1;
        r.p[1] = (float)p0;
        break;
      case ASC_FILTER_CONTRAST:
        r.p[0] = (float)p0;
        r.p[1] = (float)p1;
        r.p[2] = 3.0f;  // ContrastFilter size_ (ContrastFilter.cpp:61)
        break;
      case ASC_FILTER_GAUSSIAN: r.p[0] = (float)p0; break;
      case ASC_FILTER_SHARPEN:
        r.p[0] = (float)p0;
        r.p[1] = (float)p1;
        r.p[2] = 0.0f;  // threshold (SharpenFilter.cpp:68-71)
        break;
    }
    out.push_back(r);
  }
  (void)end_tag;
  return ASC_OK;
}

// ---------------------------------------------------------------------------
// Trial configuration (.cfg)
// ---------------------------------------------------------------------------
void unix_slashes(std::string& p)
{
  std::replace(p.begin(), p.end(), '\\', '/');
  size_t pos = 0;
  while ((pos = p.find("//", pos)) != std::string::npos) p.replace(pos, 2, "/");
  if (p.size() > 1 && p.back() == '/' && !(p.size() == 3 && p[1] == ':')) p.pop_back();
}

std::string parent_dir(const std::string& p)
{
  const size_t s = p.find_last_of('/');
  if (s == std::string::npos) return "";
  if (s == 0) return "/";
  return p.substr(0, s);
}

std::string absolute(const std::string& p, const std::string& base)
{
  if (!p.empty() && p[0] == '/') return p;
  if (base.empty()) return p;
  return base == "/" ? "/" + p : base + "/" + p;
}

}  // namespace
}  // namespace asc

using namespace asc;

struct asc_trial_s {
  asc_trial_info info;
  std::vector<asc_camera> cameras;
  std::vector<std::string> volume_files;
  std::vector<std::array<double, 3>> voxel_mm;
  std::vector<std::array<int, 3>> flips;
  std::vector<std::vector<std::string>> frames;  // per camera, sorted file names
};

ASC_API int asc_tiff_info(const char* path, int* width, int* height, int* pages, int* bits_per_sample)
{
  return guarded([&]() -> int {
    if (!path || !width || !height || !pages || !bits_per_sample) return fail(ASC_ERR_INVALID, "null pointer");
    int d[3], bps;
    const int rc = tiff_info(path, d, &bps, true);
    if (rc) return rc;
    *width = d[0];
    *height = d[1];
    *pages = d[2];
    *bits_per_sample = bps;
    return ASC_OK;
  });
}

ASC_API int asc_tiff_read(const char* path, void* dst, size_t dst_bytes)
{
  return guarded([&]() -> int {
    if (!path || !dst) return fail(ASC_ERR_INVALID, "null pointer");
    std::vector<uint8_t> data;
    int d[3], bps;
    const int rc = tiff_read_all(path, data, d, &bps, true);
    if (rc) return rc;
    if (dst_bytes != data.size())
      return fail(ASC_ERR_INVALID, "asc_tiff_read: buffer of " + std::to_string(dst_bytes) + " bytes for " +
                                       std::to_string(data.size()) + " bytes of pixels");
    std::memcpy(dst, data.data(), data.size());
    return ASC_OK;
  });
}

ASC_API int asc_camera_from_mayacam1(const double values[15], asc_camera* out)
{
  return guarded([&]() -> int {
    if (!values || !out) return fail(ASC_ERR_INVALID, "null pointer");
    return camera_mayacam1(values, *out);
  });
}

ASC_API int asc_camera_from_vtkcam(const double position[3], const double focal_point[3], const double view_up[3],
                                   double view_angle_deg, const double size[2], asc_camera* out)
{
  return guarded([&]() -> int {
    if (!position || !focal_point || !view_up || !size || !out) return fail(ASC_ERR_INVALID, "null pointer");
    return camera_vtk(position, focal_point, view_up, view_angle_deg, size, *out);
  });
}

ASC_API int asc_read_camera(const char* path, asc_camera* out)
{
  return guarded([&]() -> int {
    if (!path || !out) return fail(ASC_ERR_INVALID, "null pointer");
    return read_camera(path, *out);
  });
}

ASC_API int asc_read_filter_settings(const char* path, asc_filter_settings* out)
{
  return guarded([&]() -> int {
    if (!path || !out) return fail(ASC_ERR_INVALID, "null pointer");
    std::ifstream f(path);
    if (!f.is_open()) return io_fail(std::string("File not found: ") + path);
    asc_filter_settings s;
    std::memset(&s, 0, sizeof s);
    // DRR renderer spinboxes: Sample Distance 0.62 in [0.01, 1], XRay
    // Intensity 0.49 in [0, 1], XRay Cutoff 0 in [0, 1]
    // (ModelViewTreeWidgetItem.cpp:88-90)
    double sd = 0.62, ri = 0.49, co = 0.0;
    std::vector<asc_filter> drr, rad;
    std::string line;
    while (get_line(f, line)) {
      if (line == "DrrRenderer_begin") {
        s.has_drr_renderer = 1;
        while (get_line(f, line) && line != "DrrRenderer_end") {
          std::istringstream ls(line);
          std::string k;
          float v = 0.0f;
          ls >> k;
          if (k == "SampleDistance" && (ls >> v)) sd = spinbox(v, 0.01, 1.0);
          if (k == "RayIntensity" && (ls >> v)) ri = spinbox(v, 0.0, 1.0);
          if (k == "Cutoff" && (ls >> v)) co = spinbox(v, 0.0, 1.0);
        }
      } else if (line == "DrrFilters_begin") {
        s.has_drr_filters = 1;
        drr.clear();
        read_filters(f, "DrrFilters_end", drr);
      } else if (line == "RadFilters_begin") {
        s.has_rad_filters = 1;
        rad.clear();
        read_filters(f, "RadFilters_end", rad);
      }
    }
    if ((int)drr.size() > ASC_MAX_FILTERS || (int)rad.size() > ASC_MAX_FILTERS)
      return fail(ASC_ERR_UNSUPPORTED, std::string("more than ") + std::to_string(ASC_MAX_FILTERS) +
                                           " filters in one chain: " + path);
    s.n_drr_filters = (int)drr.size();
    s.n_rad_filters = (int)rad.size();
    std::copy(drr.begin(), drr.end(), s.drr_filters);
    std::copy(rad.begin(), rad.end(), s.rad_filters);
    s.sample_distance = sd;
    s.ray_intensity = ri;
    s.cutoff_fraction = co;
    // ModelViewTreeWidgetItem::updateModelview (:330-338), through the
    // RayCaster's float setters
    s.drr_step = (float)std::exp(7 * sd - 5);
    s.drr_intensity = (float)std::exp(15 * ri - 5);
    *out = s;
    return ASC_OK;
  });
}

ASC_API int asc_drr_cutoff(double cutoff_fraction, const asc_volume_desc* desc, float* cutoff)
{
  return guarded([&]() -> int {
    if (!desc || !cutoff) return fail(ASC_ERR_INVALID, "null pointer");
    // value * (max cutoff - min cutoff) + min cutoff, the volume's normalised
    // value range (ModelViewTreeWidgetItem.cpp:336-338)
    *cutoff = (float)(cutoff_fraction * ((double)desc->max_value - (double)desc->min_value) + (double)desc->min_value);
    return ASC_OK;
  });
}

ASC_API int asc_trial_open(const char* cfg, asc_trial* out)
{
  return guarded([&]() -> int {
    if (!cfg || !out) return fail(ASC_ERR_INVALID, "null pointer");
    const std::string path(cfg);
    std::ifstream f(path);
    if (!f.is_open()) return io_fail("File not found: " + path);
    auto err = [&](const std::string& m) {
      return io_fail("Invalid trial configuration file. " + m + " See " + path + ".");
    };
    int version[2] = {0, 0};
    std::vector<std::string> cams, roots, vols, sizes, flips, res, offs, meshes;
    std::string line;
    while (get_line(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ls(line);
      std::string key, value;
      std::getline(ls, key, ' ');
      std::getline(ls, value);
      if (key == "mayaCam_csv") {
        unix_slashes(value);
        cams.push_back(value);
      } else if (key == "CameraRootDir") {
        unix_slashes(value);
        roots.push_back(value);
      } else if (key == "VolumeFile") {
        unix_slashes(value);
        vols.push_back(value);
      } else if (key == "VolumeFlip") {
        flips.push_back(value);
      } else if (key == "VoxelSize") {
        sizes.push_back(value);
      } else if (key == "RenderResolution") {
        res.push_back(value);
      } else if (key == "OptimizationOffsets") {
        offs.push_back(value);
      } else if (key == "MeshFile") {
        meshes.push_back(value);
      } else if (key == "Version") {
        std::istringstream vs(value);
        for (int k = 0; k < 2; k++) {
          std::string part;
          std::getline(vs, part, '.');
          version[k] = std::atoi(part.c_str());
        }
      }
    }
    if (version[0] == 0 && version[1] == 0) {
      version[0] = 1;
      version[1] = 0;
    }
    if (version[0] == 1 && version[1] >= 1) {
      const std::string base = parent_dir(path);
      for (auto* list : {&cams, &roots, &vols, &meshes})
        for (auto& p : *list) p = absolute(p, base);
    }
    if (cams.size() < 1) return err("There must be at least one mayacam files.");
    if (cams.size() != roots.size())
      return err("The number of cameras and videos must match. Found " + std::to_string(cams.size()) +
                 " cameras and " + std::to_string(roots.size()) + " videos.");
    if (vols.size() < 1) return err("There must be at least one volume file.");
    if (vols.size() != sizes.size())
      return err("Each volume must be associated with its corresponding voxel sizes. Found " +
                 std::to_string(vols.size()) + " volumes and " + std::to_string(sizes.size()) + " voxel sizes.");
    if (!meshes.empty() && meshes.size() != vols.size())
      return err("You must specify a mesh file for each volume or none at all. Found " + std::to_string(vols.size()) +
                 " volumes and " + std::to_string(meshes.size()) + " mesh files.");
    std::unique_ptr<asc_trial_s> t(new asc_trial_s());
    std::memset(&t->info, 0, sizeof t->info);
    t->info.version[0] = version[0];
    t->info.version[1] = version[1];
    int rc;
    for (const std::string& c : cams) {
      asc_camera cam;
      if ((rc = read_camera(c, cam))) return rc;
      t->cameras.push_back(cam);
    }
    for (size_t i = 0; i < vols.size(); i++) {
      int d[3], bps;
      if ((rc = tiff_info(vols[i], d, &bps, true))) return rc;
      int fx = 0, fy = 0, fz = 0;
      if (i < flips.size()) {
        std::stringstream fs(flips[i]);
        fs >> fx >> fy >> fz;
      }
      float sx = 0, sy = 0, sz = 0;
      std::stringstream ss(sizes[i]);
      ss >> sx >> sy >> sz;
      t->volume_files.push_back(vols[i]);
      t->voxel_mm.push_back({(double)sx, (double)sy, (double)sz});
      t->flips.push_back({fx != 0, fy != 0, fz != 0});
    }
    int max_frames = 0;
    for (const std::string& root : roots) {
      DIR* dir = opendir(root.c_str());
      if (!dir) return io_fail("Unable to open directory: " + root);
      std::vector<std::string> names;
      while (dirent* ent = readdir(dir)) {
        const std::string n(ent->d_name);
        if (n.compare(0, 1, ".") == 0) continue;
        const size_t ext = n.find_last_of('.') + 1;
        if (n.compare(ext, 3, "tif") != 0) continue;
        names.push_back(root + "/" + n);
      }
      closedir(dir);
      if (names.empty()) return io_fail("No TIFF files found in directory: " + root);
      std::sort(names.begin(), names.end());
      max_frames = std::max(max_frames, (int)names.size());
      t->frames.push_back(names);
    }
    t->info.n_cameras = (int)t->cameras.size();
    t->info.n_volumes = (int)t->volume_files.size();
    t->info.n_frames = max_frames;
    for (int k = 0; k < 6; k++) t->info.offsets[k] = 0.1;
    if (!offs.empty()) {
      std::stringstream os(offs.back());
      for (int k = 0; k < 6; k++) os >> t->info.offsets[k];
    }
    t->info.render_width = 512;
    t->info.render_height = 512;
    if (!res.empty()) {
      std::stringstream rs(res.back());
      rs >> t->info.render_width >> t->info.render_height;
    }
    *out = t.release();
    return ASC_OK;
  });
}

ASC_API int asc_trial_get_info(asc_trial t, asc_trial_info* out)
{
  return guarded([&]() -> int {
    if (!t || !out) return fail(ASC_ERR_INVALID, "null pointer");
    *out = t->info;
    return ASC_OK;
  });
}

ASC_API int asc_trial_camera(asc_trial t, int camera, asc_camera* out)
{
  return guarded([&]() -> int {
    if (!t || !out) return fail(ASC_ERR_INVALID, "null pointer");
    if (camera < 0 || camera >= t->info.n_cameras) return fail(ASC_ERR_INVALID, "camera index out of range");
    *out = t->cameras[camera];
    return ASC_OK;
  });
}

ASC_API int asc_trial_volume_info(asc_trial t, int volume, int dims[3], int* bits_per_sample, double voxel_mm[3],
                                  int flip[3])
{
  return guarded([&]() -> int {
    if (!t || !dims || !bits_per_sample || !voxel_mm || !flip) return fail(ASC_ERR_INVALID, "null pointer");
    if (volume < 0 || volume >= t->info.n_volumes) return fail(ASC_ERR_INVALID, "volume index out of range");
    const int rc = tiff_info(t->volume_files[volume], dims, bits_per_sample, true);
    if (rc) return rc;
    for (int a = 0; a < 3; a++) {
      voxel_mm[a] = t->voxel_mm[volume][a];
      flip[a] = t->flips[volume][a];
    }
    return ASC_OK;
  });
}

ASC_API int asc_trial_volume_read(asc_trial t, int volume, void* dst, size_t dst_bytes)
{
  return guarded([&]() -> int {
    if (!t || !dst) return fail(ASC_ERR_INVALID, "null pointer");
    if (volume < 0 || volume >= t->info.n_volumes) return fail(ASC_ERR_INVALID, "volume index out of range");
    return asc_tiff_read(t->volume_files[volume].c_str(), dst, dst_bytes);
  });
}

// Video::set_frame clamps the index to the last frame (Video.cpp:221-225).
static const std::string* trial_frame(asc_trial t, int camera, int frame)
{
  if (camera < 0 || camera >= t->info.n_cameras || frame < 0) return nullptr;
  const auto& names = t->frames[camera];
  return &names[std::min<size_t>((size_t)frame, names.size() - 1)];
}

ASC_API int asc_trial_frame_info(asc_trial t, int camera, int frame, int dims[2], int* bits_per_sample)
{
  return guarded([&]() -> int {
    if (!t || !dims || !bits_per_sample) return fail(ASC_ERR_INVALID, "null pointer");
    const std::string* p = trial_frame(t, camera, frame);
    if (!p) return fail(ASC_ERR_INVALID, "camera or frame index out of range");
    int d[3];
    const int rc = tiff_info(*p, d, bits_per_sample, false);
    dims[0] = d[0];
    dims[1] = d[1];
    return rc;
  });
}

ASC_API int asc_trial_frame_read(asc_trial t, int camera, int frame, void* dst, size_t dst_bytes)
{
  return guarded([&]() -> int {
    if (!t || !dst) return fail(ASC_ERR_INVALID, "null pointer");
    const std::string* p = trial_frame(t, camera, frame);
    if (!p) return fail(ASC_ERR_INVALID, "camera or frame index out of range");
    std::vector<uint8_t> data;
    int d[3], bps;
    const int rc = tiff_read_all(*p, data, d, &bps, false);
    if (rc) return rc;
    if (dst_bytes != data.size())
      return fail(ASC_ERR_INVALID, "asc_trial_frame_read: buffer of " + std::to_string(dst_bytes) + " bytes for " +
                                       std::to_string(data.size()) + " bytes of pixels");
    std::memcpy(dst, data.data(), data.size());
    return ASC_OK;
  });
}

ASC_API int asc_trial_frame_path(asc_trial t, int camera, int frame, char* dst, size_t dst_bytes)
{
  return guarded([&]() -> int {
    if (!t || !dst || dst_bytes == 0) return fail(ASC_ERR_INVALID, "null pointer");
    const std::string* p = trial_frame(t, camera, frame);
    if (!p) return fail(ASC_ERR_INVALID, "camera or frame index out of range");
    if (p->size() + 1 > dst_bytes) return fail(ASC_ERR_INVALID, "path buffer too small");
    std::memcpy(dst, p->c_str(), p->size() + 1);
    return ASC_OK;
  });
}

// Video::create_background_image (Video.cpp:147-218): per pixel, the
// largest normalised value over all frames of the camera.
ASC_API int asc_trial_background(asc_trial t, int camera, float* dst, size_t count)
{
  return guarded([&]() -> int {
    if (!t || !dst) return fail(ASC_ERR_INVALID, "null pointer");
    if (camera < 0 || camera >= t->info.n_cameras) return fail(ASC_ERR_INVALID, "camera index out of range");
    const auto& names = t->frames[camera];
    if (names.size() < 2) return fail(ASC_ERR_INVALID, "Not enough images to create background image.");
    std::fill(dst, dst + count, 0.0f);
    for (const std::string& n : names) {
      std::vector<uint8_t> data;
      int d[3], bps;
      const int rc = tiff_read_all(n, data, d, &bps, false);
      if (rc) return rc;
      const size_t px = (size_t)d[0] * d[1];
      if (px != count)
        return fail(ASC_ERR_INVALID, "asc_trial_background: " + std::to_string(count) + " floats for " +
                                         std::to_string(px) + " pixels");
      for (size_t i = 0; i < px; i++) {
        float v;
        if (bps == 8)
          v = (float)data[i] / 255u;
        else if (bps == 16)
          v = (float)((const uint16_t*)data.data())[i] / 65535u;
        else
          v = (float)((const uint32_t*)data.data())[i] / 4294967295u;
        if (v > dst[i]) dst[i] = v;
      }
    }
    return ASC_OK;
  });
}

ASC_API int asc_trial_close(asc_trial t)
{
  return guarded([&]() -> int {
    delete t;
    return ASC_OK;
  });
}

// ---- tracking results (.tra) ------------------------------------------------

ASC_API int asc_write_tra(const char* path, const double* poses, int n_frames, int n_volumes, int flags)
{
  return guarded([&]() -> int {
    if (!path || (!poses && n_frames * n_volumes > 0) || n_frames < 0 || n_volumes < 1)
      return fail(ASC_ERR_INVALID, "bad argument");
    std::FILE* f = std::fopen(path, "w");
    if (!f) return io_fail(std::string("Failed to save to file: ") + path);
    const char* s = (flags & ASC_TRA_COMMAS) ? "," : " ";
    const bool matrix = flags & ASC_TRA_MATRIX;
    for (int i = 0; i < n_frames; i++) {
      for (int j = 0; j < n_volumes; j++) {
        const double* p = poses + ((size_t)i * n_volumes + j) * 6;
        bool invalid = false;
        for (int k = 0; k < 6; k++) invalid |= std::isnan(p[k]);
        const int n = matrix ? 16 : 6;
        double v[16];
        if (!invalid && matrix) {
          const Frame fr = frame_from_xyzypr(p);
          // CoordFrame::to_matrix (CoordFrame.cpp:203-221): column-major 4x4
          for (int c = 0; c < 3; c++) {
            for (int r = 0; r < 3; r++) v[4 * c + r] = fr.r[3 * c + r];
            v[4 * c + 3] = 0.0;
            v[12 + c] = fr.t[c];
          }
          v[15] = 1.0;
          if (flags & ASC_TRA_CM)
            for (int c = 12; c < 15; c++) v[c] /= 10.0;
          if (flags & ASC_TRA_ROWS) {
            double w[16];
            for (int r = 0; r < 4; r++)
              for (int c = 0; c < 4; c++) w[4 * r + c] = v[4 * c + r];
            std::memcpy(v, w, sizeof v);
          }
        } else if (!invalid) {
          std::memcpy(v, p, 6 * sizeof(double));
          if (flags & ASC_TRA_CM)
            for (int c = 0; c < 3; c++) v[c] /= 10.0;
          if (flags & ASC_TRA_RADIANS)
            for (int c = 3; c < 6; c++) v[c] *= M_PI / 180.0;
        }
        for (int k = 0; k < n; k++) {
          if (k) std::fputs(s, f);
          if (invalid)
            std::fputs("NaN", f);
          else
            std::fprintf(f, "%.16f", v[k]);  // std::ios::fixed, precision 16
        }
        if (j != n_volumes - 1) std::fputs(s, f);
      }
      std::fputs("\n", f);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? ASC_OK : io_fail(std::string("Failed to save to file: ") + path);
  });
}

ASC_API int asc_read_tra(const char* path, int n_frames, int n_volumes, int flags, double* poses, int* frames_read)
{
  return guarded([&]() -> int {
    if (!path || !poses || n_frames < 0 || n_volumes < 1) return fail(ASC_ERR_INVALID, "bad argument");
    std::ifstream f(path);
    if (!f.is_open()) return io_fail(std::string("Tracking Data File Not Found: ") + path);
    const char s = (flags & ASC_TRA_COMMAS) ? ',' : ' ';
    const bool matrix = flags & ASC_TRA_MATRIX;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::fill(poses, poses + (size_t)n_frames * n_volumes * 6, nan);
    std::string line, value;
    int i = 0;
    for (; i < n_frames && get_line(f, line); ++i) {
      std::istringstream ls(line);
      for (int k = 0; k < n_volumes; k++) {
        double m[16] = {0};
        for (int j = 0; j < (matrix ? 16 : 6) && std::getline(ls, value, s); ++j) {
          std::istringstream vs(value);
          vs >> m[j];
        }
        if (value.compare(0, 3, "NaN") == 0) continue;
        if (matrix && (flags & ASC_TRA_ROWS)) {
          double n[16];
          std::memcpy(n, m, sizeof n);
          for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) m[4 * c + r] = n[4 * r + c];
        }
        if (flags & ASC_TRA_CM) {
          const int o = matrix ? 12 : 0;
          for (int c = 0; c < 3; c++) m[o + c] *= 10.0;
        }
        if ((flags & ASC_TRA_RADIANS) && !matrix)
          for (int c = 3; c < 6; c++) m[c] *= 180.0 / M_PI;
        double* p = poses + ((size_t)i * n_volumes + k) * 6;
        if (matrix) {
          // CoordFrame::from_matrix (CoordFrame.cpp:191-201), then to_xyzypr
          Frame fr;
          for (int c = 0; c < 3; c++) {
            for (int r = 0; r < 3; r++) fr.r[3 * c + r] = m[4 * c + r];
            fr.t[c] = m[12 + c];
          }
          frame_to_xyzypr(fr, p);
        } else {
          std::memcpy(p, m, 6 * sizeof(double));
        }
      }
    }
    if (frames_read) *frames_read = i;
    return ASC_OK;
  });
}

// ---- display output ----------------------------------------------------------

namespace {

// CRC-32 (ISO 3309 / ITU-T V.42, reflected polynomial 0xEDB88320), as the
// PNG specification's annex D computes it, table driven.
uint32_t png_crc(const uint8_t* p, size_t n, uint32_t crc = 0xFFFFFFFFu)
{
  static uint32_t table[256];
  static bool ready = false;
  if (!ready) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    ready = true;
  }
  for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return crc;
}

void put_be32(std::vector<uint8_t>& o, uint32_t v)
{
  for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}

void png_chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data)
{
  put_be32(o, (uint32_t)data.size());
  const size_t at = o.size();
  o.insert(o.end(), type, type + 4);
  o.insert(o.end(), data.begin(), data.end());
  put_be32(o, png_crc(o.data() + at, o.size() - at) ^ 0xFFFFFFFFu);
}

// An 8-bit greyscale (channels 1) or RGB (channels 3) PNG whose image data
// is a zlib stream (RFC 1950) of stored deflate blocks (RFC 1951 section
// 3.2.4), each scanline with filter type 0.
int write_png(const char* path, const uint8_t* pix, int w, int h, int channels)
{
  if (!path || !pix || w <= 0 || h <= 0 || (channels != 1 && channels != 3))
    return fail(ASC_ERR_INVALID, "bad argument");
  const size_t row = (size_t)w * channels;
  std::vector<uint8_t> raw;
  raw.reserve((row + 1) * h);
  for (int y = 0; y < h; y++) {
    raw.push_back(0);
    raw.insert(raw.end(), pix + row * y, pix + row * (y + 1));
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  size_t at = 0;
  do {
    const size_t len = std::min<size_t>(65535, raw.size() - at);
    z.push_back(at + len == raw.size() ? 1 : 0);
    z.push_back((uint8_t)(len & 0xFF));
    z.push_back((uint8_t)(len >> 8));
    z.push_back((uint8_t)(~len & 0xFF));
    z.push_back((uint8_t)((~len >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
    at += len;
  } while (at < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  put_be32(z, (b << 16) | a);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, (uint32_t)w);
  put_be32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, (uint8_t)(channels == 3 ? 2 : 0), 0, 0, 0});
  png_chunk(out, "IHDR", ihdr);
  png_chunk(out, "IDAT", z);
  png_chunk(out, "IEND", {});
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(std::string("cannot write ") + path);
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  return (std::fclose(f) == 0 && ok) ? ASC_OK : io_fail(std::string("write failed: ") + path);
}

}  // namespace

ASC_API int asc_write_png(const char* path, const uint8_t* pixels, int width, int height, int channels)
{
  return guarded([&]() -> int {
    return write_png(path, pixels, width, height, channels);
  });
}

// The composite as glDrawPixels(GL_RGB, GL_FLOAT) shows it: row 0 at the
// bottom, each component clamped to [0, 1] and rounded to 8 bits.
ASC_API int asc_write_view_png(const char* path, const float* rgb, int width, int height)
{
  return guarded([&]() -> int {
    if (!rgb || width <= 0 || height <= 0) return fail(ASC_ERR_INVALID, "bad argument");
    const size_t row = 3 * (size_t)width;
    std::vector<uint8_t> pix(row * height);
    for (int y = 0; y < height; y++) {
      const float* src = rgb + row * (height - 1 - y);
      for (size_t i = 0; i < row; i++) {
        const float v = src[i];
        pix[row * y + i] = v > 0.0f ? (uint8_t)std::floor(std::min(v, 1.0f) * 255.0f + 0.5f) : 0;
      }
    }
    return write_png(path, pix.data(), width, height, 3);
  });
}

// save_full_drr (Tracker.cpp:155-204): the image stretched to 0..255 over
// its own minimum and maximum (the maximum starting from the smallest
// positive float), (int) truncation, stored inverted (255 - value, air
// white) as an ASCII PGM, each value followed by a space, rows in memory
// order. An image with no range (maximum == minimum) writes 255 everywhere.
ASC_API int asc_write_full_drr_pgm(const char* path, const float* image, int width, int height)
{
  return guarded([&]() -> int {
    if (!path || !image || width <= 0 || height <= 0) return fail(ASC_ERR_INVALID, "bad argument");
    const size_t n = (size_t)width * height;
    float lo = std::numeric_limits<float>::max();
    float hi = std::numeric_limits<float>::min();
    for (size_t i = 0; i < n; i++) {
      if (image[i] > hi) hi = image[i];
      if (image[i] < lo) lo = image[i];
    }
    const float range = hi - lo;
    std::string text = "P2\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
    text.reserve(text.size() + 4 * n);
    for (size_t i = 0; i < n; i++) {
      unsigned char u = 0;
      if (range > 0.0f) u = (unsigned char)(int)(255.0f * (image[i] - lo) / range);
      text += std::to_string(255 - (int)u);
      text += ' ';
    }
    std::FILE* f = std::fopen(path, "w");
    if (!f) return io_fail(std::string("cannot write ") + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    return (std::fclose(f) == 0 && ok) ? ASC_OK : io_fail(std::string("write failed: ") + path);
  });
}
