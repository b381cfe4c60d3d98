// The render engine's host interface shared by the translation units that
// drive it (the C-ABI entry points, the autotuner and the serving loop).
#pragma once

#include <vector>

#include "ct_internal.h"

namespace ct {

// Camera frame of camera.ts:8-22 (binary64).
struct Frame {
  double o[3], r[3], u[3], f[3];
  double tanHalfFov, aspect, lensRadius, focusDistance;
};

struct TraceOpts {
  int maxBounces, rr, nee;
};

// A job to render from pass `first` to `first + count`; `finish` multiplies
// the sum by `scale` when the job's last pass of this plan is added
// (renderImage's final 1/spp).
struct PlanJob {
  const ct_scene* scene;
  Frame F;
  TraceOpts opts;
  uint32_t width, height;
  uint32_t frameBase;  // frame index of pass 0
  uint32_t first, count;
  float* accum;
  bool finish;
  double scale;
  // FP32 precision form: persistent compensation terms (width*height*3) kept
  // by the caller across calls, or null for per-call terms folded into
  // `accum` when the call's last pass is added
  float* comp = nullptr;
  // radianceInto jobs: rays, RNG states in and out, radiance out (JobDev)
  const double* rays = nullptr;
  uint32_t* rngIO = nullptr;
  double* radiance = nullptr;
};

Frame make_frame(const double* frame16);

// Packs the jobs' passes into rounds of at most the tracer's path capacity
// and traces them on `st`; returns when the last round is done.
int run_plan(ct_tracer* T, std::vector<PlanJob> jobs, hipStream_t st);

// Sets n floats to zero with a kernel (safe under graph capture).
int zero_floats(float* p, size_t n, hipStream_t st);

// tonemapBuffer on the device: width*height*3 float32 -> width*height RGBA8.
int launch_tonemap(const float* d_accum, size_t npix, double samples, double exposureEv,
                   int mode, uint8_t* d_rgba, hipStream_t st);

// Path capacity of a tracer (paths in flight per round).
size_t tracer_capacity(const ct_tracer* T);

// Creates the tracer's side streams (second path group, NEE launches) now
// rather than at its first round.  A process gets few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default) and a stream created once they are all
// taken shares one, serialising its launches with the other stream's: a
// caller that creates its own round stream should call this right after, so
// that the round's four streams hold four queues and later streams share.
int tracer_streams(ct_tracer* T);

}  // namespace ct
