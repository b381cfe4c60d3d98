// Batched closest-hit and any-hit queries over a device-resident scene.
//
// One lane per ray.  Rays are read as three 16-byte loads (6 binary64); the
// scene is read straight from the packed float32 arrays: 2 x 16 B per node
// and 3 x 16 B per primitive, the algorithmic minimum of the layout.
#include "ct_internal.h"
#include "ct_qwalk.h"
#include "ct_wavefront.h"

namespace ct {

constexpr int QUERY_BLOCK = 256;
constexpr int ORDERED_STACK = 16;

__device__ __forceinline__ void load_ray(const double* __restrict__ rays, size_t i, double& ox,
                                         double& oy, double& oz, double& dx, double& dy,
                                         double& dz) {
  const double2* r = (const double2*)(rays + 6 * i);
  double2 a = r[0], b = r[1], c = r[2];
  ox = a.x; oy = a.y; oz = b.x; dx = b.y; dy = c.x; dz = c.y;
}

// Sum a per-lane count over the wave and add it to a 64-bit global counter.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
  unsigned long long x = v;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if ((threadIdx.x & 63) == 0 && x) atomicAdd(dst, x);
}

template <int WALK, bool COUNT>
__global__ void __launch_bounds__(QUERY_BLOCK)
k_closest(DevScene S, const double* __restrict__ rays, uint32_t n, double tMin,
          const double* __restrict__ tMax, int32_t* __restrict__ prim, double* __restrict__ tOut,
          double* __restrict__ uOut, double* __restrict__ vOut,
          unsigned long long* __restrict__ counts) {
  __shared__ int sNode[WALK == CT_WALK_ORDERED ? ORDERED_STACK * QUERY_BLOCK : 1];
  __shared__ float sEntry[WALK == CT_WALK_ORDERED ? ORDERED_STACK * QUERY_BLOCK : 1];
  TraversalCounts cnt = {0, 0, 0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double ox, oy, oz, dx, dy, dz;
    load_ray(rays, i, ox, oy, oz, dx, dy, dz);
    Hit h;
    const double t0 = tMax ? tMax[i] : __builtin_inf();
    h.t = t0;
    h.u = 0;
    h.v = 0;
    h.prim = -1;
    h.tie = false;
    if (WALK == CT_WALK_ORDERED) {
      closest_wave<ORDERED_STACK, COUNT>(S, ox, oy, oz, dx, dy, dz, tMin, h,
                                         sNode + threadIdx.x, sEntry + threadIdx.x,
                                         QUERY_BLOCK, cnt);
      settle_ties<COUNT>(S, ox, oy, oz, dx, dy, dz, tMin, t0, h, cnt);
    } else {  // intersectClosestStackless itself: ties to the lower packed index
      closest_stackless<COUNT, TIES_FIRST>(S, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
    }
    if (prim) prim[i] = h.prim;
    if (tOut) tOut[i] = h.prim < 0 ? __builtin_inf() : h.t;
    if (uOut) uOut[i] = h.u;
    if (vOut) vOut[i] = h.v;
  }
  if (COUNT) {
    wave_add(counts + 0, cnt.nodes);
    wave_add(counts + 1, cnt.tris);
    wave_add(counts + 2, cnt.spheres);
  }
}

__global__ void __launch_bounds__(QUERY_BLOCK)
k_occluded(DevScene S, const double* __restrict__ rays, uint32_t n, double tMin,
           const double* __restrict__ tMax, uint8_t* __restrict__ out) {
  TraversalCounts cnt = {0, 0, 0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double ox, oy, oz, dx, dy, dz;
    load_ray(rays, i, ox, oy, oz, dx, dy, dz);
    double tm = tMax ? tMax[i] : __builtin_inf();
    out[i] = occluded<false>(S, ox, oy, oz, dx, dy, dz, tMin, tm, cnt) ? 1 : 0;
  }
}

constexpr int WIDE_STACK = 24;

// u, v of the closest triangle hit, recomputed once after the walk: the same
// test on the same operands gives the same bits, and the walk does not carry
// two doubles per lane through every step.  A sphere hit gets 0, 0 (the
// reference leaves whatever the last triangle hit of its own visiting order
// set, intersect.ts:173-176).
__device__ __forceinline__ void hit_uv(const DevScene& S, int p, double ox, double oy, double oz,
                                       double dx, double dy, double dz, double tMin, double& u,
                                       double& v) {
  u = 0;
  v = 0;
  if (p < 0) return;
  PrimGeom g = load_prim(S, p);
  double t;
  if (g.a.w != 0.0f)
    ray_triangle(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z, g.c.x, g.c.y,
                 g.c.z, tMin, __builtin_inf(), t, u, v);
}

// Closest hit over the derived 4-wide nodes, organised like the path
// tracer's extend kernel: each lane walks its own rays (a fixed sequence
// i, i + stride, ...), parks leaves (up to PARK_SLOTS) until the wave's lanes
// all hold PARK_LEAVES of them or are done, then the parked leaves are tested together;
// a lane whose ray finishes takes its next ray at the top of the loop.
template <bool COUNT>
__global__ void __launch_bounds__(QUERY_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_closest_wide(DevScene S, const double* __restrict__ rays, uint32_t n, double tMin,
               const double* __restrict__ tMax, int32_t* __restrict__ prim,
               double* __restrict__ tOut, double* __restrict__ uOut, double* __restrict__ vOut,
               unsigned long long* __restrict__ counts) {
  __shared__ int sNodeAll[WIDE_STACK * QUERY_BLOCK];
  __shared__ qw::Entry sEntryAll[WIDE_STACK * QUERY_BLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  qw::Entry* sEntry = sEntryAll + threadIdx.x;
  const int lanes = QUERY_BLOCK;
  // 32-bit ray indices (n is): `next` advances by the grid's lane count and
  // stops at n without wrapping
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t next = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t i = 0;
  TraversalCounts cnt = {0, 0, 0, 0};
  bool have = false;
  double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0;
  qw::QRay Q = qw::qray(0, 0, 0, 1, 1, 1);
  Hit h;
  h.t = 0; h.u = 0; h.v = 0; h.prim = -1; h.tie = false;
  int cur = -1, sp = 0;
  Parked<PARK_SLOTS> pk;
  pk.n = 0;
  for (;;) {
    if (!have && next < n) {
      i = next;
      next = n - next > stride ? next + stride : n;
      load_ray(rays, i, ox, oy, oz, dx, dy, dz);
      h.t = tMax ? tMax[i] : __builtin_inf();
      h.u = 0; h.v = 0; h.prim = -1; h.tie = false;
      Q = qw::qray64(S, ox, oy, oz, dx, dy, dz);
      sp = 0;
      pk.n = 0;
      cur = 0;
      have = true;
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false;
      while (cur != -1) {
        if (cur <= -2) {
          if (pk.n == PARK_SLOTS) break;
          park(pk, -2 - cur);
          cur = -1;
        } else if (!qw::qwide_step<true, WIDE_STACK, COUNT>(S, Q, qw::cull_bound(h.t), cur, sp,
                                                           sNode, sEntry, lanes, cnt)) {
          overflow = true;
          cur = -1;
          break;
        }
        if (cur == -1) qw::qpop(cur, sp, sNode, sEntry, lanes, qw::cull_bound(h.t));
        if (__all(pk.n >= PARK_LEAVES || cur == -1)) break;
      }
#pragma unroll
      for (int k = 0; k < PARK_SLOTS; k++)  // constant slots: the words stay in registers
        if (k < pk.n) test_leaf<COUNT>(S, pk.w[k], ox, oy, oz, dx, dy, dz, tMin, h, cnt);
      pk.n = 0;
      if (overflow) {
        if (COUNT) cnt.fallbacks++;
        closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
        cur = -1;
        sp = 0;
      }
      if (cur == -1) {
        settle_ties<COUNT>(S, ox, oy, oz, dx, dy, dz, tMin, tMax ? tMax[i] : __builtin_inf(), h,
                           cnt);
        if (prim) prim[i] = h.prim;
        if (tOut) tOut[i] = h.prim < 0 ? __builtin_inf() : h.t;
        if (uOut || vOut) {
          double u, v;
          hit_uv(S, h.prim, ox, oy, oz, dx, dy, dz, tMin, u, v);
          if (uOut) uOut[i] = u;
          if (vOut) vOut[i] = v;
        }
        have = false;
      }
    }
  }
  if (COUNT) {
    wave_add(counts + 0, cnt.nodes);
    wave_add(counts + 1, cnt.tris);
    wave_add(counts + 2, cnt.spheres);
  }
}

// Workgroups the device holds at once for `kernel` (CUs x resident workgroups
// per CU; 2048 if the runtime cannot say).  The wide query kernel's lanes walk
// fixed ray strides, so a grid of exactly this size ends without a second,
// partial wave of workgroups running on part of the chip: 16M random rays on
// the 1M-triangle soup 805 -> 885 Mrays/s against a fixed 2048.
template <typename Kernel>
static unsigned resident_workgroups(Kernel kernel, int block) {
  int dev = 0, cus = 0, perCu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, block, 0) != hipSuccess ||
      cus <= 0 || perCu <= 0)
    return 2048;
  return (unsigned)(cus * perCu);
}

template <int WALK, bool COUNT>
static int launch_closest(const ct_scene* s, const double* rays, uint32_t n, double tMin,
                          const double* tMax, int32_t* prim, double* t, double* u, double* v,
                          unsigned long long* counts, hipStream_t stream) {
  if (WALK == CT_WALK_WIDE) {
    unsigned grid = grid_for(n, QUERY_BLOCK,
                             resident_workgroups(k_closest_wide<COUNT>, QUERY_BLOCK));
    hipLaunchKernelGGL((k_closest_wide<COUNT>), dim3(grid), dim3(QUERY_BLOCK), 0, stream, s->dev,
                       rays, n, tMin, tMax, prim, t, u, v, counts);
    CT_LAUNCH_CHECK("k_closest_wide");
    return CT_OK;
  }
  unsigned grid = grid_for(n, QUERY_BLOCK, 256 * 32);
  hipLaunchKernelGGL((k_closest<WALK, COUNT>), dim3(grid), dim3(QUERY_BLOCK), 0, stream, s->dev,
                     rays, n, tMin, tMax, prim, t, u, v, counts);
  CT_LAUNCH_CHECK("k_closest");
  return CT_OK;
}

}  // namespace ct

using namespace ct;

extern "C" {

int ct_closest_hit(const ct_scene* s, const double* rays, uint32_t n, double tMin,
                   const double* tMax, int32_t walk, int32_t* prim, double* t, double* u,
                   double* v, void* stream) {
  if (!s) return set_error(CT_EINVAL, "ct_closest_hit: null scene");
  if (n == 0) return CT_OK;
  if (!rays) return set_error(CT_EINVAL, "ct_closest_hit: null rays");
  hipStream_t st = (hipStream_t)stream;
  if (walk == CT_WALK_ORDERED)
    return launch_closest<CT_WALK_ORDERED, false>(s, rays, n, tMin, tMax, prim, t, u, v, nullptr, st);
  if (walk == CT_WALK_STACKLESS)
    return launch_closest<CT_WALK_STACKLESS, false>(s, rays, n, tMin, tMax, prim, t, u, v, nullptr, st);
  if (walk == CT_WALK_WIDE)
    return launch_closest<CT_WALK_WIDE, false>(s, rays, n, tMin, tMax, prim, t, u, v, nullptr, st);
  return set_error(CT_EINVAL, "ct_closest_hit: unknown walk");
}

int ct_closest_hit_counts(const ct_scene* s, const double* rays, uint32_t n, double tMin,
                          const double* tMax, int32_t walk, uint64_t counts[3], void* stream) {
  if (!s || !counts) return set_error(CT_EINVAL, "ct_closest_hit_counts: null argument");
  counts[0] = counts[1] = counts[2] = 0;
  if (n == 0) return CT_OK;
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* d = nullptr;
  CT_HIP(hipMallocAsync((void**)&d, 3 * sizeof(unsigned long long), st));
  unsigned long long zero[3] = {0, 0, 0};
  CT_HIP(hipMemcpyAsync(d, zero, sizeof zero, hipMemcpyHostToDevice, st));
  int rc;
  if (walk == CT_WALK_ORDERED)
    rc = launch_closest<CT_WALK_ORDERED, true>(s, rays, n, tMin, tMax, nullptr, nullptr, nullptr, nullptr, d, st);
  else if (walk == CT_WALK_STACKLESS)
    rc = launch_closest<CT_WALK_STACKLESS, true>(s, rays, n, tMin, tMax, nullptr, nullptr, nullptr, nullptr, d, st);
  else if (walk == CT_WALK_WIDE)
    rc = launch_closest<CT_WALK_WIDE, true>(s, rays, n, tMin, tMax, nullptr, nullptr, nullptr, nullptr, d, st);
  else
    rc = set_error(CT_EINVAL, "ct_closest_hit_counts: unknown walk");
  unsigned long long h[3] = {0, 0, 0};
  if (rc == CT_OK) CT_HIP(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, st));
  CT_HIP(hipFreeAsync(d, st));
  CT_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 3; k++) counts[k] = h[k];
  return rc;
}

int ct_occluded(const ct_scene* s, const double* rays, uint32_t n, double tMin,
                const double* tMax, uint8_t* out, void* stream) {
  if (!s) return set_error(CT_EINVAL, "ct_occluded: null scene");
  if (n == 0) return CT_OK;
  if (!rays || !out) return set_error(CT_EINVAL, "ct_occluded: null array");
  unsigned grid = grid_for(n, QUERY_BLOCK, 256 * 32);
  hipLaunchKernelGGL(k_occluded, dim3(grid), dim3(QUERY_BLOCK), 0, (hipStream_t)stream, s->dev,
                     rays, n, tMin, tMax, out);
  CT_LAUNCH_CHECK("k_occluded");
  return CT_OK;
}

}  // extern "C"
