// Host-side internals shared by the translation units of libcaustic_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "caustic_amd.h"
#include "ct_device.h"

struct ct_scene {
  ct::DevScene dev;       // views into `buf`
  void* buf = nullptr;    // one device allocation holding every array
  size_t bytes = 0;
  int device = 0;
};

namespace ct {

int set_error(int code, const std::string& msg);
// Returns CT_OK, or records the HIP error (with `what`) and returns CT_EDEVICE.
int check_hip(hipError_t e, const char* what);

#define CT_HIP(call)                                \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) return ct::check_hip(e_, #call); \
  } while (0)

#define CT_LAUNCH_CHECK(what)                       \
  do {                                              \
    hipError_t e_ = hipGetLastError();              \
    if (e_ != hipSuccess) return ct::check_hip(e_, what); \
  } while (0)

// Number of workgroups for a grid-stride kernel over `n` items: enough waves
// to fill all 256 CUs several times over, never more than the work needs.
inline unsigned grid_for(size_t n, unsigned block, unsigned maxBlocks = 256 * 16) {
  size_t b = (n + block - 1) / block;
  if (b > maxBlocks) b = maxBlocks;
  return b == 0 ? 1u : (unsigned)b;
}

}  // namespace ct
