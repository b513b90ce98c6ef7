// Replays malformed inputs through the host-side readers and the remote
// session's message handler, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_io_fuzz.py:
//   reader_fuzz tiff|cam|vie|tra|cfg <file>...   (each read: ok or an error status)
//   reader_fuzz messages <trial.cfg> <count>     (random protocol messages)
// Prints "ok N error M"; a memory error aborts with the sanitizer's report.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/autoscoper_hip.h"

// The library's error plumbing is internal to it; the sources linked here
// report through these.
namespace asc {
void set_error(const std::string&) {}
int fail(int code, const std::string&) { return code; }
}  // namespace asc

static int read_one(const std::string& kind, const char* path)
{
  if (kind == "tiff") {
    int w, h, n, b;
    int rc = asc_tiff_info(path, &w, &h, &n, &b);
    if (!rc && (size_t)w * h * n * (b / 8) <= (16u << 20)) {
      std::vector<unsigned char> buf((size_t)w * h * n * (b / 8));
      rc = asc_tiff_read(path, buf.data(), buf.size());
    }
    return rc;
  }
  if (kind == "cam") {
    asc_camera c;
    return asc_read_camera(path, &c);
  }
  if (kind == "vie") {
    asc_filter_settings s;
    return asc_read_filter_settings(path, &s);
  }
  if (kind == "tra") {
    std::vector<double> p(3 * 2 * 6);
    int frames;
    return asc_read_tra(path, 3, 2, ASC_TRA_COMMAS, p.data(), &frames);
  }
  asc_trial t;
  const int rc = asc_trial_open(path, &t);
  if (!rc) asc_trial_close(t);
  return rc;
}

int main(int argc, char** argv)
{
  if (argc < 3) return 2;
  const std::string kind = argv[1];
  int ok = 0, err = 0;
  if (kind == "messages") {
    asc_remote r;
    if (asc_remote_create(nullptr, &r)) return 2;
    std::mt19937 rng(11);
    std::vector<unsigned char> reply(1 << 16);
    const std::string cfg = argv[2];
    const int count = argc > 3 ? std::atoi(argv[3]) : 10000;
    for (int it = 0; it < count; it++) {
      std::vector<unsigned char> m;
      const int type = it % 50 == 0 ? 1 : (int)(rng() % 20);
      m.push_back((unsigned char)type);
      if (type == 1 && rng() % 2) {
        m.insert(m.end(), cfg.begin(), cfg.end());
      } else {
        const int len = (int)(rng() % 80);
        for (int k = 0; k < len; k++) m.push_back((unsigned char)(rng() % 256));
        if (type == 2 || type == 3 || type == 4 || type == 12) {
          // file names stay in the working directory
          const char* name = "fz.tra";
          m.resize(std::min<size_t>(m.size(), 29));
          m.insert(m.end(), name, name + std::strlen(name));
        }
      }
      size_t n;
      const int rc = asc_remote_handle(r, m.data(), m.size(), reply.data(), reply.size(), &n);
      (rc == 0 || rc == 1 ? ok : err)++;
    }
    asc_remote_destroy(r);
  } else {
    for (int i = 2; i < argc; i++) (read_one(kind, argv[i]) ? err : ok)++;
  }
  std::printf("ok %d error %d\n", ok, err);
  return 0;
}
