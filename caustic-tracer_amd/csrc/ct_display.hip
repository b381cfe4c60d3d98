// Display transform of an accumulated image: tonemapBuffer (tonemap.ts:41-62)
// on the device, so served and progressive frames leave HBM as 8-bit RGBA.
// Arithmetic is the reference's (binary64 per channel; pow from the device
// math library).
#include <cmath>

#include "ct_engine.h"

namespace ct {
namespace {

// acesFitted (tonemap.ts:7-16)
__device__ __forceinline__ double aces(double x) {
  const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  double v = (x * (a * x + b)) / (x * (c * x + d) + e);
  return v < 0 ? 0 : v > 1 ? 1 : v;
}

// reinhard (tonemap.ts:19-21)
__device__ __forceinline__ double reinhard(double x) { return x <= 0 ? 0 : x / (1 + x); }

// linearToSrgb (tonemap.ts:28-32)
__device__ __forceinline__ double to_srgb(double x) {
  if (x <= 0) return 0;
  if (x >= 1) return 1;
  return x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1 / 2.4) - 0.055;
}

// Math.round (ties toward +infinity) then the Uint8ClampedArray store.
__device__ __forceinline__ uint8_t to_byte(double v) {
  double f = floor(v);
  double r = (v - f >= 0.5) ? f + 1 : f;
  if (!(r > 0)) return 0;
  if (r > 255) return 255;
  return (uint8_t)r;
}

__global__ void __launch_bounds__(256)
k_tonemap(const float* __restrict__ accum, size_t npix, double inv, int mode,
          uchar4* __restrict__ rgba) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
    double c[3];
    for (int k = 0; k < 3; k++) {
      double x = (double)accum[3 * i + k] * inv;
      c[k] = to_srgb(mode == CT_TONEMAP_ACES ? aces(x) : reinhard(x)) * 255;
    }
    rgba[i] = make_uchar4(to_byte(c[0]), to_byte(c[1]), to_byte(c[2]), 255);
  }
}

}  // namespace

int launch_tonemap(const float* d_accum, size_t npix, double samples, double exposureEv,
                   int mode, uint8_t* d_rgba, hipStream_t st) {
  if (mode != CT_TONEMAP_ACES && mode != CT_TONEMAP_REINHARD)
    return set_error(CT_EINVAL, "tonemap: mode must be CT_TONEMAP_ACES or CT_TONEMAP_REINHARD");
  if (npix == 0) return CT_OK;
  // exposureScale(ev) / samples (tonemap.ts:34-36, 50)
  double inv = samples > 0 ? std::pow(2.0, exposureEv) / samples : 0;
  hipLaunchKernelGGL(k_tonemap, dim3(grid_for(npix, 256, 256 * 16)), dim3(256), 0, st, d_accum,
                     npix, inv, mode, (uchar4*)d_rgba);
  CT_LAUNCH_CHECK("k_tonemap");
  return CT_OK;
}

}  // namespace ct

extern "C" int ct_tonemap(const float* d_accum, uint32_t width, uint32_t height, double samples,
                          double exposureEv, int32_t mode, uint8_t* d_rgba, void* stream) {
  if (!d_accum || !d_rgba) return ct::set_error(CT_EINVAL, "ct_tonemap: null buffer");
  return ct::launch_tonemap(d_accum, (size_t)width * height, samples, exposureEv, mode, d_rgba,
                            (hipStream_t)stream);
}
