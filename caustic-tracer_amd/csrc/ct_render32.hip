// FP32 precision form of the wavefront tracer.
//
// Same decomposition, queues, job table and RNG stream as the binary64
// kernels; what changes is the arithmetic and the node format:
//   * traversal reads the 8-bit quantized 4-wide nodes (64 B per node, half
//     the float32 wide node) and evaluates slab planes in FP32 with a
//     per-node error bound, so a box is entered whenever the float32 box it
//     replaces could be entered -- the walk is conservative and the closest
//     hit is decided by the FP32 primitive tests alone;
//   * intersection and shading run in FP32 (the reference's algorithm and
//     draw order, binary32 operations);
//   * each pixel's samples are added with Neumaier-compensated FP32 sums (a
//     float sum plus a float compensation term per channel).
// Random numbers keep the reference's u32 stream and draw order; a draw is
// the top 24 bits scaled by 2^-24 (in [0, 1) exactly in binary32).
#include <string.h>

#include "ct_qwalk.h"
#include "ct_wavefront.h"

namespace ct {
namespace f32 {

constexpr float T_MIN_F = 1e-4f;
constexpr float TWO_PI_F = 6.28318530717958647692f;
constexpr float INV_PI_F = 0.31830988618379067154f;
constexpr float SHADOW_SHORTEN_F = 0.999f;
using qw::CULL_SLACK;
using qw::SLAB_REL;
constexpr int QUERY_BLOCK = 256;
constexpr int QUERY_STACK = 24;

__device__ __forceinline__ float next(Rng& r) {
  return (float)(r.u32() >> 8) * 5.9604644775390625e-8f;
}
__device__ __forceinline__ int pick(Rng& r, int n) {
  float f = floorf(next(r) * (float)n);
  float m = (float)(n - 1);
  return (int)(m < f ? m : f);
}
__device__ __forceinline__ float max0(float x) { return x > 0 ? x : (x != x ? x : 0.0f); }
// The walk ray is the ray whose primitives are tested: no extra slack.
using Ray = qw::QRay;
__device__ __forceinline__ Ray make_ray(float ox, float oy, float oz, float dx, float dy,
                                        float dz) {
  return qw::qray(ox, oy, oz, dx, dy, dz);
}

struct Hit {
  float t, u, v;
  int prim;
};

// raySphere (intersect.ts:20-39) in binary32.
__device__ __forceinline__ float ray_sphere(const Ray& R, float cx, float cy, float cz, float r,
                                            float tMin, float tMax) {
  float lx = R.ox - cx, ly = R.oy - cy, lz = R.oz - cz;
  float a = R.dx * R.dx + R.dy * R.dy + R.dz * R.dz;
  float halfB = lx * R.dx + ly * R.dy + lz * R.dz;
  float c = lx * lx + ly * ly + lz * lz - r * r;
  float disc = halfB * halfB - a * c;
  if (disc < 0) return -1;
  float s = sqrtf(disc);
  float t = (-halfB - s) / a;
  if (t <= tMin || t >= tMax) {
    t = (-halfB + s) / a;
    if (t <= tMin || t >= tMax) return -1;
  }
  return t;
}

// rayTriangle (intersect.ts:52-85) in binary32.  Before the division a
// triangle is dropped when its barycentric numerators put u or v certainly
// outside [0, 1] (margins of 2^-20 relative and 1e-30 absolute cover the
// rounding of the full computation), i.e. only when the full test would drop
// it; otherwise the full computation decides.
__device__ __forceinline__ bool ray_triangle(const Ray& R, float ax, float ay, float az,
                                             float bx, float by, float bz, float cx, float cy,
                                             float cz, float tMin, float tMax, float& tOut,
                                             float& uOut, float& vOut) {
  float e1x = bx - ax, e1y = by - ay, e1z = bz - az;
  float e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
  float px = R.dy * e2z - R.dz * e2y;
  float py = R.dz * e2x - R.dx * e2z;
  float pz = R.dx * e2y - R.dy * e2x;
  float det = e1x * px + e1y * py + e1z * pz;
  if (det > -1e-12f && det < 1e-12f) return false;
  float tx = R.ox - ax, ty = R.oy - ay, tz = R.oz - az;
  float un = tx * px + ty * py + tz * pz;
  float qx = ty * e1z - tz * e1y;
  float qy = tz * e1x - tx * e1z;
  float qz = tx * e1y - ty * e1x;
  float vn = R.dx * qx + R.dy * qy + R.dz * qz;
  {
    float ad = fabsf(det);
    float su = det > 0 ? un : -un, sv = det > 0 ? vn : -vn;
    float lim = ad * 1.00000095367431640625f;
    if (su < -1e-30f || su > lim || sv < -1e-30f || su + sv > lim * 1.00000095367431640625f)
      return false;
  }
  float invDet = 1.0f / det;
  float u = un * invDet;
  if (u < 0 || u > 1) return false;
  float v = vn * invDet;
  if (v < 0 || u + v > 1) return false;
  float t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
  if (t <= tMin || t >= tMax) return false;
  tOut = t;
  uOut = u;
  vOut = v;
  return true;
}

// Leaf loop (strict '<' keeps the first primitive tested on exact ties).
template <bool COUNT>
__device__ __forceinline__ void test_leaf(const DevScene& S, int word, const Ray& R, float tMin,
                                          Hit& h, TraversalCounts& cnt) {
  int count = word & 7, start = word >> 3;
  for (int p = start; p < start + count; p++) {
    PrimGeom g = load_prim(S, p);
    if (g.a.w == 0.0f) {
      if (COUNT) cnt.spheres++;
      float t = ray_sphere(R, g.a.x, g.a.y, g.a.z, g.b.x, tMin, h.t);
      if (t > 0) {
        h.t = t;
        h.prim = p;
      }
    } else {
      if (COUNT) cnt.tris++;
      float t, u, v;
      if (ray_triangle(R, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z, g.c.x, g.c.y, g.c.z, tMin,
                       h.t, t, u, v)) {
        h.t = t;
        h.prim = p;
        h.u = u;
        h.v = v;
      }
    }
  }
}

template <bool COUNT>
__device__ __forceinline__ bool leaf_blocks(const DevScene& S, int word, const Ray& R,
                                            float tMin, float tMax, TraversalCounts& cnt) {
  int count = word & 7, start = word >> 3;
  for (int p = start; p < start + count; p++) {
    PrimGeom g = load_prim(S, p);
    if (g.a.w == 0.0f) {
      if (COUNT) cnt.spheres++;
      if (ray_sphere(R, g.a.x, g.a.y, g.a.z, g.b.x, tMin, tMax) > 0) return true;
    } else {
      if (COUNT) cnt.tris++;
      float t, u, v;
      if (ray_triangle(R, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z, g.c.x, g.c.y, g.c.z, tMin,
                       tMax, t, u, v))
        return true;
    }
  }
  return false;
}

// Conservative FP32 slab test of a float32 box (binary nodes of the stackless
// fallback): t = (plane - o)*inv has relative error below 3 ulp-halvings, so
// widening the interval by 2^-21 of its largest end keeps every box the exact
// test would enter.
__device__ __forceinline__ bool box_hit(const Ray& R, float4 lo, float4 hi, float tCull) {
  float tx0 = (lo.x - R.ox) * R.ix, tx1 = (hi.x - R.ox) * R.ix;
  float ty0 = (lo.y - R.oy) * R.iy, ty1 = (hi.y - R.oy) * R.iy;
  float tz0 = (lo.z - R.oz) * R.iz, tz1 = (hi.z - R.oz) * R.iz;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
  float E = SLAB_REL * fmaxf(fabsf(tn), fabsf(tf));
  tn = fmaxf(tn - E, 0.0f);
  tf = fminf(tf + E, tCull);
  return tn <= tf;
}

// Stackless closest hit over the binary nodes with exact leaf words: the
// fallback of the quantized walk when its LDS stack would overflow (boxes
// entered under the current bound; re-testing primitives cannot change a
// strict-'<' minimum).
template <bool COUNT>
__device__ __forceinline__ void closest_stackless(const DevScene& S, const Ray& R, float tMin,
                                                  Hit& h, TraversalCounts& cnt) {
  int node = 0;
  const int nodeCount = S.nodeCount;
  while (node < nodeCount) {
    float4 lo = gload(S.nodes + 2 * (size_t)node);
    float4 hi = gload(S.nodes + 2 * (size_t)node + 1);
    if (COUNT) cnt.nodes++;
    if (!box_hit(R, lo, hi, h.t * CULL_SLACK)) {
      node = (int)lo.w;
      continue;
    }
    if (hi.w >= 0.0f) test_leaf<COUNT>(S, (int)gload(S.leafWords + node), R, tMin, h, cnt);
    node++;
  }
}

template <bool COUNT>
__device__ __forceinline__ bool occluded_stackless(const DevScene& S, const Ray& R, float tMin,
                                                   float tMax, TraversalCounts& cnt) {
  int node = 0;
  const int nodeCount = S.nodeCount;
  const float tCull = tMax * CULL_SLACK;
  while (node < nodeCount) {
    float4 lo = gload(S.nodes + 2 * (size_t)node);
    float4 hi = gload(S.nodes + 2 * (size_t)node + 1);
    if (COUNT) cnt.nodes++;
    if (!box_hit(R, lo, hi, tCull)) {
      node = (int)lo.w;
      continue;
    }
    if (hi.w >= 0.0f && leaf_blocks<COUNT>(S, (int)gload(S.leafWords + node), R, tMin, tMax, cnt))
      return true;
    node++;
  }
  return false;
}

using qw::Entry;
using qw::entry_value;
using qw::qwide_step;

// Per-lane closest hit with the quantized walk (queries).
template <int STACK, bool COUNT>
__device__ void closest_q(const DevScene& S, const Ray& R, float tMin, Hit& h,
                          int* __restrict__ sNode, Entry* __restrict__ sEntry, int lanes,
                          TraversalCounts& cnt) {
  int cur = 0, sp = 0;
  for (;;) {
    if (cur >= 0) {
      if (!qwide_step<true, STACK, COUNT>(S, R, h.t * CULL_SLACK, cur, sp, sNode, sEntry, lanes,
                                          cnt)) {
        if (COUNT) cnt.fallbacks++;
        closest_stackless<COUNT>(S, R, tMin, h, cnt);
        return;
      }
    } else if (cur <= -2) {
      test_leaf<COUNT>(S, -2 - cur, R, tMin, h, cnt);
      cur = -1;
    } else {
      while (sp > 0) {
        sp--;
        if (entry_value(sEntry[sp * lanes]) <= h.t * CULL_SLACK) {
          cur = sNode[sp * lanes];
          break;
        }
      }
      if (cur == -1) return;
    }
  }
}

template <int STACK, bool COUNT>
__device__ bool occluded_q(const DevScene& S, const Ray& R, float tMin, float tMax,
                           int* __restrict__ sNode, int lanes, TraversalCounts& cnt) {
  int cur = 0, sp = 0;
  const float tCull = tMax * CULL_SLACK;
  for (;;) {
    if (cur >= 0) {
      if (!qwide_step<false, STACK, COUNT>(S, R, tCull, cur, sp, sNode, nullptr, lanes, cnt)) {
        if (COUNT) cnt.fallbacks++;
        return occluded_stackless<COUNT>(S, R, tMin, tMax, cnt);
      }
    } else if (cur <= -2) {
      if (leaf_blocks<COUNT>(S, -2 - cur, R, tMin, tMax, cnt)) return true;
      cur = -1;
    } else {
      if (sp == 0) return false;
      sp--;
      cur = sNode[sp * lanes];
    }
  }
}

// ---------------------------------------------------------------------------
// Query kernels (batched closest hit / any hit in FP32)

template <bool BRUTE, bool COUNT>
__global__ void __launch_bounds__(QUERY_BLOCK)
k_closest32(DevScene S, const float* __restrict__ rays, uint32_t n, float tMin,
            const float* __restrict__ tMax, int32_t* __restrict__ prim, float* __restrict__ tOut,
            float* __restrict__ uOut, float* __restrict__ vOut,
            unsigned long long* __restrict__ counts) {
  __shared__ int sNode[BRUTE ? 1 : QUERY_STACK * QUERY_BLOCK];
  __shared__ Entry sEntry[BRUTE ? 1 : QUERY_STACK * QUERY_BLOCK];
  TraversalCounts cnt = {0, 0, 0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float* r = rays + 6 * i;
    Ray R = make_ray(r[0], r[1], r[2], r[3], r[4], r[5]);
    Hit h;
    h.t = tMax ? tMax[i] : __builtin_inff();
    h.u = 0;
    h.v = 0;
    h.prim = -1;
    if (BRUTE) {
      for (int p = 0; p < S.primCount; p++) test_leaf<false>(S, p * 8 + 1, R, tMin, h, cnt);
    } else {
      closest_q<QUERY_STACK, COUNT>(S, R, tMin, h, sNode + threadIdx.x, sEntry + threadIdx.x,
                                    QUERY_BLOCK, cnt);
    }
    if (prim) prim[i] = h.prim;
    if (tOut) tOut[i] = h.prim < 0 ? __builtin_inff() : h.t;
    if (uOut) uOut[i] = h.u;
    if (vOut) vOut[i] = h.v;
  }
  if (COUNT && !BRUTE) {
    wave_add64(counts + 0, cnt.nodes);
    wave_add64(counts + 1, cnt.tris);
    wave_add64(counts + 2, cnt.spheres);
    wave_add64(counts + 3, cnt.fallbacks);
  }
}

template <bool BRUTE>
__global__ void __launch_bounds__(QUERY_BLOCK)
k_occluded32(DevScene S, const float* __restrict__ rays, uint32_t n, float tMin,
             const float* __restrict__ tMax, uint8_t* __restrict__ out) {
  __shared__ int sNode[BRUTE ? 1 : QUERY_STACK * QUERY_BLOCK];
  TraversalCounts cnt = {0, 0, 0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float* r = rays + 6 * i;
    Ray R = make_ray(r[0], r[1], r[2], r[3], r[4], r[5]);
    float tm = tMax ? tMax[i] : __builtin_inff();
    bool b = false;
    if (BRUTE) {
      for (int p = 0; p < S.primCount && !b; p++) b = leaf_blocks<false>(S, p * 8 + 1, R, tMin, tm, cnt);
    } else {
      b = occluded_q<QUERY_STACK, false>(S, R, tMin, tm, sNode + threadIdx.x, QUERY_BLOCK, cnt);
    }
    out[i] = b ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Engine kernels

__global__ void k_clear32(uint32_t* ctr, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
}

// generateRay (camera.ts:62-72) with the pixel/lens draws of tracer.ts:373-384.
template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_generate32(PathSoA32 P, const JobDev* __restrict__ jobs, int njobs, size_t base,
             size_t count) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const size_t pid = base + i;
    int j = job_of<MULTI, false>(jobs, njobs, pid);
    const JobDev& J = jobs[j];
    const Frame& F = J.F;
    const uint32_t width = J.width, height = J.height, npix = J.npix;
    size_t local = pid - J.pathBase;
    uint32_t pass = (uint32_t)(local / npix);
    uint32_t pixel = (uint32_t)(local - (size_t)pass * npix);
    uint32_t y = pixel / width, x = pixel - y * width;
    Rng rng;
    rng.s = seed_for(pixel, J.frameBase + pass);
    float sx = ((float)x + next(rng)) / (float)width;
    float sy = 1.0f - ((float)y + next(rng)) / (float)height;
    float lx = 0, ly = 0;
    const float lens = (float)F.lensRadius;
    if (lens > 0) {
      float r = sqrtf(next(rng));
      float theta = TWO_PI_F * next(rng);
      float sn, cs;
      sincosf(theta, &sn, &cs);
      lx = r * cs;
      ly = r * sn;
    }
    const float thf = (float)F.tanHalfFov, fd = (float)F.focusDistance;
    float px = (2 * sx - 1) * thf * (float)F.aspect;
    float py = (2 * sy - 1) * thf;
    float o[3], d[3];
    for (int k = 0; k < 3; k++) {
      float ok = (float)F.o[k], rk = (float)F.r[k], uk = (float)F.u[k], fk = (float)F.f[k];
      float focus = ok + (rk * px + uk * py + fk) * fd;
      o[k] = ok + (rk * (lx * lens) + uk * (ly * lens));
      d[k] = focus - o[k];
    }
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len > 0) {
      float s = 1.0f / len;
      d[0] *= s; d[1] *= s; d[2] *= s;
    } else {
      d[0] = d[1] = d[2] = 0;
    }
    P.ox[pid] = o[0]; P.oy[pid] = o[1]; P.oz[pid] = o[2];
    P.dx[pid] = d[0]; P.dy[pid] = d[1]; P.dz[pid] = d[2];
    P.br[pid] = 1; P.bg[pid] = 1; P.bb[pid] = 1;
    P.Lr[pid] = 0; P.Lg[pid] = 0; P.Lb[pid] = 0;
    P.prevPdf[pid] = 0;
    P.specular[pid] = 1;
    P.rng[pid] = rng.s;
    if (MULTI) P.jobOf[pid] = (uint32_t)j;
    P.q0[i] = (uint32_t)pid;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.ctr[0] = (uint32_t)count;
}

// lightPdfScalar (tracer.ts:61-90)
__device__ float light_pdf(const DevScene& S, float px, float py, float pz, float hx, float hy,
                           float hz, float hnx, float hny, float hnz, int prim) {
  int lightCount = S.lightCount;
  if (lightCount == 0) return 0;
  float tx = hx - px, ty = hy - py, tz = hz - pz;
  float dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return 0;
  float inv = 1.0f / sqrtf(dist2);
  float wx = tx * inv, wy = ty * inv, wz = tz * inv;
  PrimGeom g = load_prim(S, prim);
  if (g.a.w == 0.0f) {
    float cx = g.a.x - px, cy = g.a.y - py, cz = g.a.z - pz;
    float r = g.b.x;
    float d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      float cosMax = sqrtf(max0(1 - (r * r) / d2));
      float solid = TWO_PI_F * (1 - cosMax);
      return solid > 0 ? 1.0f / (solid * lightCount) : 0;
    }
  }
  float cosL = fabsf(hnx * wx + hny * wy + hnz * wz);
  if (cosL < 1e-6f) return 0;
  return dist2 / (g.c.w * cosL * lightCount);
}

// basisInto (materials.ts:20-42)
__device__ __forceinline__ void basis(float nx, float ny, float nz, float* b) {
  float hx = 1, hy = 0, hz = 0;
  if (fabsf(nx) > 0.9f) {
    hx = 0;
    hy = 1;
  }
  float tx = hy * nz - hz * ny;
  float ty = hz * nx - hx * nz;
  float tz = hx * ny - hy * nx;
  float inv = 1.0f / sqrtf(tx * tx + ty * ty + tz * tz);
  tx *= inv;
  ty *= inv;
  tz *= inv;
  b[0] = tx; b[1] = ty; b[2] = tz;
  b[3] = ny * tz - nz * ty;
  b[4] = nz * tx - nx * tz;
  b[5] = nx * ty - ny * tx;
}

// unitVectorInto (materials.ts:67-74)
__device__ __forceinline__ void unit_vector(float u1, float u2, float* out) {
  float z = 1 - 2 * u1;
  float r = sqrtf(z * z < 1 ? 1 - z * z : 0);
  float sn, cs;
  sincosf(TWO_PI_F * u2, &sn, &cs);
  out[0] = r * cs;
  out[1] = r * sn;
  out[2] = z;
}

struct LightSample {
  float w[3], dist, pdf, e[3];
};

// sampleLightInto (tracer.ts:93-180); consumes its draws even when the sample
// turns out degenerate.
__device__ bool sample_light(const DevScene& S, float px, float py, float pz, Rng& rng,
                             LightSample& L) {
  int lightCount = S.lightCount;
  if (lightCount == 0) return false;
  int prim = gload(S.lights + pick(rng, lightCount));
  const LightGeom g = load_light(S, prim);
  float4 em = gload(S.mats + (size_t)g.obj * 4 + 3);
  L.e[0] = em.x; L.e[1] = em.y; L.e[2] = em.z;
  if (g.a.w == 0.0f) {
    float cx = g.a.x - px, cy = g.a.y - py, cz = g.a.z - pz;
    float r = g.b.x;
    float d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      float cosMax = sqrtf(max0(1 - (r * r) / d2));
      float cosT = 1 - next(rng) * (1 - cosMax);
      float sinT = sqrtf(max0(1 - cosT * cosT));
      float phi = TWO_PI_F * next(rng);
      float invD = 1.0f / sqrtf(d2);
      float ax = cx * invD, ay = cy * invD, az = cz * invD;
      float b[6];
      basis(ax, ay, az, b);
      float sn, cs;
      sincosf(phi, &sn, &cs);
      float sx = sinT * cs, sy = sinT * sn;
      float wx = b[0] * sx + b[3] * sy + ax * cosT;
      float wy = b[1] * sx + b[4] * sy + ay * cosT;
      float wz = b[2] * sx + b[5] * sy + az * cosT;
      float invW = 1.0f / sqrtf(wx * wx + wy * wy + wz * wz);
      wx *= invW;
      wy *= invW;
      wz *= invW;
      float halfB = -(cx * wx + cy * wy + cz * wz);
      float disc = halfB * halfB - (d2 - r * r);
      if (disc < 0) return false;
      float dist = -halfB - sqrtf(disc);
      if (dist <= 0) return false;
      L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
      L.dist = dist;
      L.pdf = 1.0f / (TWO_PI_F * (1 - cosMax) * lightCount);
      return true;
    }
    float u1 = next(rng);
    float u2 = next(rng);
    float nv[3];
    unit_vector(u1, u2, nv);
    float tx = cx + nv[0] * r, ty = cy + nv[1] * r, tz = cz + nv[2] * r;
    float dist2 = tx * tx + ty * ty + tz * tz;
    if (dist2 <= 0) return false;
    float dist = sqrtf(dist2);
    float wx = tx / dist, wy = ty / dist, wz = tz / dist;
    float cosL = fabsf(nv[0] * wx + nv[1] * wy + nv[2] * wz);
    if (cosL < 1e-6f) return false;
    L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
    L.dist = dist;
    L.pdf = dist2 / (g.c.w * cosL * lightCount);
    return true;
  }
  float u1 = next(rng);
  float u2 = next(rng);
  if (u1 + u2 > 1) {
    u1 = 1 - u1;
    u2 = 1 - u2;
  }
  float ax = g.a.x, ay = g.a.y, az = g.a.z;
  float qx = ax + (g.b.x - ax) * u1 + (g.c.x - ax) * u2;
  float qy = ay + (g.b.y - ay) * u1 + (g.c.y - ay) * u2;
  float qz = az + (g.b.z - az) * u1 + (g.c.z - az) * u2;
  float tx = qx - px, ty = qy - py, tz = qz - pz;
  float dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return false;
  float dist = sqrtf(dist2);
  float wx = tx / dist, wy = ty / dist, wz = tz / dist;
  const float4 nrm = g.n;
  float cosL = fabsf(nrm.x * wx + nrm.y * wy + nrm.z * wz);
  if (cosL < 1e-6f) return false;
  L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
  L.dist = dist;
  L.pdf = dist2 / (g.c.w * cosL * lightCount);
  return true;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {
  float a2 = a * a;
  float b2 = b * b;
  float denom = a2 + b2;
  return denom > 0 ? a2 / denom : 0;
}

// schlick / dielectricReflectance (materials.ts:77-95)
__device__ __forceinline__ float schlick(float cosTheta, float etaI, float etaT) {
  float q = (etaI - etaT) / (etaI + etaT);
  float r0 = q * q;
  float c = cosTheta < 1 ? cosTheta : 1;
  c = c > 0 ? c : 0;
  float m = 1 - c;
  return r0 + (1 - r0) * m * m * m * m * m;
}

__device__ __forceinline__ float dielectric_reflectance(float cosI, float etaI, float etaT) {
  float eta = etaI / etaT;
  float sin2T = eta * eta * (1 - cosI * cosI);
  if (sin2T >= 1) return 1;
  if (etaI > etaT) return schlick(sqrtf(1 - sin2T), etaI, etaT);
  return schlick(cosI, etaI, etaT);
}

// One bounce of radianceInto (tracer.ts:191-328) for every live path.
template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_shade32(PathSoA32 P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ qin,
          const uint32_t* __restrict__ countIn, uint32_t* __restrict__ qout,
          uint32_t* __restrict__ countOut, uint32_t* __restrict__ shadowCount, int bounce) {
  const uint32_t n = *countIn;
  uint32_t stride = gridDim.x * blockDim.x;
  uint32_t iters = (n + stride - 1) / stride;
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t it = 0; it < iters; it++, slot += stride) {
    bool valid = slot < n;
    uint32_t pid = valid ? qin[slot] : 0;
    bool alive = false, shadow = false;
    float wdir[3] = {0, 0, 0}, sTmax = 0, sc[3] = {0, 0, 0};
    if (valid) {
      const JobDev& J = jobs[MULTI ? P.jobOf[pid] : 0];
      const DevScene& S = J.S;
      const TraceOpts opts = J.opts;
      const bool useNee = opts.nee && S.lightCount > 0;
      float ox = P.ox[pid], oy = P.oy[pid], oz = P.oz[pid];
      float dx = P.dx[pid], dy = P.dy[pid], dz = P.dz[pid];
      float br = P.br[pid], bg = P.bg[pid], bb = P.bb[pid];
      float Lr = P.Lr[pid], Lg = P.Lg[pid], Lb = P.Lb[pid];
      float prevPdf = P.prevPdf[pid];
      bool specularBounce = P.specular[pid] != 0;
      Rng rng;
      rng.s = P.rng[pid];
      int prim = P.hitPrim[pid];
      bool cont = true;
      float ndx = 0, ndy = 0, ndz = 0, px = ox, py = oy, pz = oz;
      if (prim < 0) {
        if (S.skyKind == 1) {
          float t = 0.5f * (dy + 1);
          float st = (float)S.skyStrength;
          Lr += br * ((float)S.horizon[0] + ((float)S.zenith[0] - (float)S.horizon[0]) * t) * st;
          Lg += bg * ((float)S.horizon[1] + ((float)S.zenith[1] - (float)S.horizon[1]) * t) * st;
          Lb += bb * ((float)S.horizon[2] + ((float)S.zenith[2] - (float)S.horizon[2]) * t) * st;
        }
        cont = false;
      } else {
        float t = P.hitT[pid];
        px = ox + dx * t;
        py = oy + dy * t;
        pz = oz + dz * t;
        const HitGeom g = load_hit(S, prim);
        float nx, ny, nz;
        if (g.a.w == 0.0f) {
          float invR = 1.0f / g.b.x;
          nx = (px - g.a.x) * invR;
          ny = (py - g.a.y) * invR;
          nz = (pz - g.a.z) * invR;
        } else {
          nx = g.n.x;
          ny = g.n.y;
          nz = g.n.z;
        }
        bool frontFace = dx * nx + dy * ny + dz * nz < 0;
        float sx = frontFace ? nx : -nx;
        float sy = frontFace ? ny : -ny;
        float sz = frontFace ? nz : -nz;
        size_t m = (size_t)g.obj * 4;
        float4 m0 = gload(S.mats + m), m1 = gload(S.mats + m + 1), m3 = gload(S.mats + m + 3);
        float er = m3.x, eg = m3.y, eb = m3.z;
        if (er > 0 || eg > 0 || eb > 0) {
          float w = 1;
          if (useNee && !specularBounce) {
            float pl = light_pdf(S, ox, oy, oz, px, py, pz, nx, ny, nz, prim);
            w = power_heuristic(prevPdf, pl);
          }
          Lr += br * er * w;
          Lg += bg * eg * w;
          Lb += bb * eb * w;
        }
        float type = m0.w;
        if (type == 3.0f) {
          cont = false;  // MAT_EMISSIVE
        } else if (type == 0.0f) {  // MAT_LAMBERTIAN
          float ar = m0.x, ag = m0.y, ab = m0.z;
          if (m1.z > 0.5f) {
            float cs = m1.y;
            float sum = floorf(px * cs + 0.5f) + floorf(py * cs + 0.5f) + floorf(pz * cs + 0.5f);
            // checker parity of the integer sum, exactly (see the binary64 kernel)
            if (sum - 2.0f * floorf(0.5f * sum) != 0) {  // !checkerPrimary
              float4 m2 = gload(S.mats + m + 2);
              ar = m2.x;
              ag = m2.y;
              ab = m2.z;
            }
          }
          LightSample ls;
          if (useNee && sample_light(S, px, py, pz, rng, ls)) {
            float pdfL = ls.pdf;
            float cosS = sx * ls.w[0] + sy * ls.w[1] + sz * ls.w[2];
            if (cosS > 0 && pdfL > 0) {
              float pdfB = cosS * INV_PI_F;
              float k = (cosS * INV_PI_F * power_heuristic(pdfL, pdfB)) / pdfL;
              shadow = true;
              wdir[0] = ls.w[0];
              wdir[1] = ls.w[1];
              wdir[2] = ls.w[2];
              sTmax = ls.dist * SHADOW_SHORTEN_F;
              sc[0] = br * ar * ls.e[0] * k;
              sc[1] = bg * ag * ls.e[1] * k;
              sc[2] = bb * ab * ls.e[2] * k;
            }
          }
          // cosineHemisphereInto (materials.ts:47-64)
          float u1 = next(rng);
          float u2 = next(rng);
          float r = sqrtf(u1);
          float sn, cs2;
          sincosf(TWO_PI_F * u2, &sn, &cs2);
          float lx = r * cs2;
          float ly = r * sn;
          float lz = sqrtf(u1 < 1 ? 1 - u1 : 0);
          float b[6];
          basis(sx, sy, sz, b);
          float ddx = b[0] * lx + b[3] * ly + sx * lz;
          float ddy = b[1] * lx + b[4] * ly + sy * lz;
          float ddz = b[2] * lx + b[5] * ly + sz * lz;
          float inv = 1.0f / sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
          ndx = ddx * inv;
          ndy = ddy * inv;
          ndz = ddz * inv;
          prevPdf = max0(sx * ndx + sy * ndy + sz * ndz) * INV_PI_F;
          br *= ar;
          bg *= ag;
          bb *= ab;
          specularBounce = false;
        } else if (type == 1.0f) {  // MAT_METAL: metalInto (materials.ts:102-124)
          float u1 = next(rng);
          float u2 = next(rng);
          float u3 = next(rng);
          float rough = m1.x;
          float k = 2 * (dx * sx + dy * sy + dz * sz);
          float rx = dx - k * sx;
          float ry = dy - k * sy;
          float rz = dz - k * sz;
          if (rough > 0) {
            float uv[3];
            unit_vector(u1, u2, uv);
            float rad = cbrtf(u3) * rough;
            rx += uv[0] * rad;
            ry += uv[1] * rad;
            rz += uv[2] * rad;
          }
          float inv = 1.0f / sqrtf(rx * rx + ry * ry + rz * rz);
          ndx = rx * inv;
          ndy = ry * inv;
          ndz = rz * inv;
          if (ndx * sx + ndy * sy + ndz * sz <= 0) {
            cont = false;
          } else {
            br *= m0.x;
            bg *= m0.y;
            bb *= m0.z;
            specularBounce = true;
          }
        } else {  // MAT_DIELECTRIC: dielectricInto (materials.ts:131-166)
          float u = next(rng);
          float ior = m1.x;
          float etaI = frontFace ? 1.0f : ior;
          float etaT = frontFace ? ior : 1.0f;
          float eta = etaI / etaT;
          float cosI = -(dx * sx + dy * sy + dz * sz);
          if (cosI > 1) cosI = 1;
          float refl = dielectric_reflectance(cosI, etaI, etaT);
          float sin2T = eta * eta * (1 - cosI * cosI);
          if (u < refl || sin2T > 1) {
            float k = 2 * (dx * sx + dy * sy + dz * sz);
            ndx = dx - k * sx;
            ndy = dy - k * sy;
            ndz = dz - k * sz;
          } else {
            float cosT = sqrtf(1 - sin2T);
            float s2 = eta * cosI - cosT;
            float tx = dx * eta + sx * s2;
            float ty = dy * eta + sy * s2;
            float tz = dz * eta + sz * s2;
            float inv = 1.0f / sqrtf(tx * tx + ty * ty + tz * tz);
            ndx = tx * inv;
            ndy = ty * inv;
            ndz = tz * inv;
          }
          br *= m0.x;
          bg *= m0.y;
          bb *= m0.z;
          specularBounce = true;
        }
        if (cont && opts.rr && bounce >= 3) {  // tracer.ts:305-312
          float maxB = br > bg ? (br > bb ? br : bb) : bg > bb ? bg : bb;
          float pSurvive = maxB > 1 ? 1 : maxB < 0.05f ? 0.05f : maxB;
          if (next(rng) > pSurvive) {
            cont = false;
          } else {
            br /= pSurvive;
            bg /= pSurvive;
            bb /= pSurvive;
          }
        }
      }
      P.Lr[pid] = Lr;
      P.Lg[pid] = Lg;
      P.Lb[pid] = Lb;
      P.ox[pid] = px;
      P.oy[pid] = py;
      P.oz[pid] = pz;
      if (cont && bounce < opts.maxBounces) {
        alive = true;
        P.dx[pid] = ndx;
        P.dy[pid] = ndy;
        P.dz[pid] = ndz;
        P.br[pid] = br;
        P.bg[pid] = bg;
        P.bb[pid] = bb;
        P.prevPdf[pid] = prevPdf;
        P.specular[pid] = specularBounce ? 1 : 0;
        P.rng[pid] = rng.s;
      }
    }
    wave_push(alive, pid, qout, countOut);
    uint32_t sslot = wave_reserve(shadow, shadowCount);
    if (shadow) {
      P.sPath[sslot] = pid;
      P.sdx[sslot] = wdir[0];
      P.sdy[sslot] = wdir[1];
      P.sdz[sslot] = wdir[2];
      P.sTmax[sslot] = sTmax;
      P.sCr[sslot] = sc[0];
      P.sCg[sslot] = sc[1];
      P.sCb[sslot] = sc[2];
    }
  }
}

// Persistent closest hit over the quantized nodes, organised like the
// binary64 wave walk: lanes park leaves (up to PARK_SLOTS32) until every lane of
// the wave holds PARK_LEAVES of them, then the parked leaves are tested
// together; lanes refill from the queue as their rays finish.
template <int STACK, bool COUNT, bool MULTI, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_extend32(PathSoA32 P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ q,
           const uint32_t* __restrict__ count, uint32_t* __restrict__ fetch) {
  // single-scene launches read the top levels of the 4-wide tree from LDS
  constexpr int TOP = MULTI ? 0 : TOP_NODES;
  __shared__ int sNodeAll[STACK * BLOCK];
  __shared__ Entry sEntryAll[STACK * BLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  Entry* sEntry = sEntryAll + threadIdx.x;
  const int lanes = BLOCK;
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  bool have = false, exhausted = false;
  uint32_t pid = 0;
  DevScene S = jobs[0].S;
  float4* top = nullptr;
  if constexpr (TOP > 0) {
    __shared__ float4 sTop[4 * (TOP > 0 ? TOP : 1)];
    top = sTop;
    qw::stage_top<TOP>(top, S);
  }
  Ray R = make_ray(0, 0, 0, 1, 1, 1);
  Hit h;
  h.t = 0; h.u = 0; h.v = 0; h.prim = -1;
  int cur = -1, sp = 0;
  Parked<PARK_SLOTS32> pk;
  pk.n = 0;
  for (;;) {
    if (!exhausted) {
      uint32_t slot = refill(!have, n, fetch, exhausted);
      if (slot != 0xffffffffu) {
        pid = q[slot];
        if (MULTI) S = jobs[P.jobOf[pid]].S;
        R = make_ray(P.ox[pid], P.oy[pid], P.oz[pid], P.dx[pid], P.dy[pid], P.dz[pid]);
        h.t = __builtin_inff(); h.u = 0; h.v = 0; h.prim = -1;
        have = true;
        sp = 0;
        pk.n = 0;
        cur = 0;
      }
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false;
      while (cur != -1) {
        if (cur <= -2) {
          if (pk.n == PARK_SLOTS32) break;
          park(pk, -2 - cur);
          cur = -1;
        } else if (!qwide_step<true, STACK, COUNT, TOP>(S, R, h.t * CULL_SLACK, cur, sp, sNode,
                                                        sEntry, lanes, cnt, top)) {
          overflow = true;
          cur = -1;
          break;
        }
        if (cur == -1) {
          while (sp > 0) {
            sp--;
            if (entry_value(sEntry[sp * lanes]) <= h.t * CULL_SLACK) {
              cur = sNode[sp * lanes];
              break;
            }
          }
        }
        if (__all(pk.n >= PARK_LEAVES || cur == -1)) break;
      }
#pragma unroll
      for (int k = 0; k < PARK_SLOTS32; k++)  // constant slots: the words stay in registers
        if (k < pk.n) test_leaf<COUNT>(S, pk.w[k], R, T_MIN_F, h, cnt);
      pk.n = 0;
      if (overflow) {
        if (COUNT) cnt.fallbacks++;
        closest_stackless<COUNT>(S, R, T_MIN_F, h, cnt);
        cur = -1;
        sp = 0;
      }
      if (cur == -1) {
        P.hitPrim[pid] = h.prim;
        P.hitT[pid] = h.t;
        have = false;
      }
    }
  }
  if (COUNT) {
    wave_add64(P.ctr64 + 0, cnt.nodes);
    wave_add64(P.ctr64 + 1, cnt.tris);
    wave_add64(P.ctr64 + 2, cnt.spheres);
    wave_add64(P.ctr64 + 6, cnt.fallbacks);
  }
}

// Persistent any hit of the NEE rays; adds the pending term when unoccluded.
template <int STACK, bool COUNT, bool MULTI, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_shadow32(PathSoA32 P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ count,
           uint32_t* __restrict__ fetch) {
  constexpr int TOP = MULTI ? 0 : TOP_NODES;
  __shared__ int sNodeAll[STACK * BLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  const int lanes = BLOCK;
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  bool have = false, exhausted = false, blocked = false;
  uint32_t slotCur = 0, pid = 0;
  DevScene S = jobs[0].S;
  float4* top = nullptr;
  if constexpr (TOP > 0) {
    __shared__ float4 sTop[4 * (TOP > 0 ? TOP : 1)];
    top = sTop;
    qw::stage_top<TOP>(top, S);
  }
  Ray R = make_ray(0, 0, 0, 1, 1, 1);
  float tMax = 0, tCull = 0;
  int cur = -1, sp = 0;
  Parked<SHADOW_PARK_SLOTS32> pk;
  pk.n = 0;
  for (;;) {
    if (!exhausted) {
      uint32_t slot = refill(!have, n, fetch, exhausted);
      if (slot != 0xffffffffu) {
        slotCur = slot;
        pid = P.sPath[slot];
        if (MULTI) S = jobs[P.jobOf[pid]].S;
        R = make_ray(P.ox[pid], P.oy[pid], P.oz[pid], P.sdx[slot], P.sdy[slot], P.sdz[slot]);
        tMax = P.sTmax[slot];
        tCull = tMax * CULL_SLACK;
        have = true;
        blocked = false;
        sp = 0;
        pk.n = 0;
        cur = 0;
      }
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false;
      while (cur != -1) {
        if (cur <= -2) {
          if (pk.n == SHADOW_PARK_SLOTS32) break;
          park(pk, -2 - cur);
          cur = -1;
        } else if (!qwide_step<false, STACK, COUNT, TOP>(S, R, tCull, cur, sp, sNode, nullptr,
                                                         lanes, cnt, top)) {
          overflow = true;
          cur = -1;
          break;
        }
        if (cur == -1 && sp > 0) {
          sp--;
          cur = sNode[sp * lanes];
        }
        if (__all(pk.n >= SHADOW_PARK_LEAVES || cur == -1)) break;
      }
#pragma unroll
      for (int k = 0; k < SHADOW_PARK_SLOTS32; k++)
        if (k < pk.n && !blocked) blocked = leaf_blocks<COUNT>(S, pk.w[k], R, T_MIN_F, tMax, cnt);
      pk.n = 0;
      if (overflow && !blocked) {
        if (COUNT) cnt.fallbacks++;
        blocked = occluded_stackless<COUNT>(S, R, T_MIN_F, tMax, cnt);
      }
      if (blocked || overflow) {
        cur = -1;
        sp = 0;
      }
      if (cur == -1) {
        if (!blocked) {
          P.Lr[pid] += P.sCr[slotCur];
          P.Lg[pid] += P.sCg[slotCur];
          P.Lb[pid] += P.sCb[slotCur];
        }
        have = false;
      }
    }
  }
  if (COUNT) {
    wave_add64(P.ctr64 + 3, cnt.nodes);
    wave_add64(P.ctr64 + 4, cnt.tris);
    wave_add64(P.ctr64 + 5, cnt.spheres);
    wave_add64(P.ctr64 + 7, cnt.fallbacks);
  }
}

// Neumaier's compensated addition of x into (s, c).
__device__ __forceinline__ void comp_add(float& s, float& c, float x) {
  float t = s + x;
  if (fabsf(s) >= fabsf(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

// Adds each pass's sample in pass order into the compensated sum, dropping
// non-finite samples (tracer.ts:386-392); with doScale the image becomes
// (sum + compensation) * scale, with fold the compensation is added into the
// sum (end of a call that owns no persistent compensation buffer).
template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_finalize32(PathSoA32 P, const JobDev* __restrict__ jobs, int njobs, size_t totalPix) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gi < totalPix; gi += stride) {
    const JobDev& J = jobs[job_of<MULTI, true>(jobs, njobs, gi)];
    uint32_t pix = (uint32_t)(gi - J.pixelBase);
    float* acc = J.accum + 3 * (size_t)pix;
    float* cmp = J.comp + 3 * (size_t)pix;
    float s0 = acc[0], s1 = acc[1], s2 = acc[2];
    float c0 = cmp[0], c1 = cmp[1], c2 = cmp[2];
    for (uint32_t s = 0; s < J.passes; s++) {
      size_t pid = J.pathBase + (size_t)s * J.npix + pix;
      float cr = P.Lr[pid], cg = P.Lg[pid], cb = P.Lb[pid];
      if (isfinite(cr) && isfinite(cg) && isfinite(cb)) {
        comp_add(s0, c0, cr);
        comp_add(s1, c1, cg);
        comp_add(s2, c2, cb);
      }
    }
    if (J.doScale) {
      float sc = (float)J.scale;
      s0 = (s0 + c0) * sc;
      s1 = (s1 + c1) * sc;
      s2 = (s2 + c2) * sc;
      c0 = c1 = c2 = 0;
    } else if (J.fold) {
      s0 = s0 + c0;
      s1 = s1 + c1;
      s2 = s2 + c2;
      c0 = c1 = c2 = 0;
    }
    acc[0] = s0; acc[1] = s1; acc[2] = s2;
    cmp[0] = c0; cmp[1] = c1; cmp[2] = c2;
  }
}

// FP64 sum of an FP32 image per channel (the furnace energy balance): one
// binary64 partial per workgroup, reduced in a fixed order.
__global__ void __launch_bounds__(256)
k_sum_f64(const float* __restrict__ img, size_t npix, double* __restrict__ partial) {
  __shared__ double red[3][256];
  double a[3] = {0, 0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
    a[0] += img[3 * i];
    a[1] += img[3 * i + 1];
    a[2] += img[3 * i + 2];
  }
  for (int k = 0; k < 3; k++) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) partial[3 * (size_t)blockIdx.x + k] = red[k][0];
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x < 3) {
    double s = 0;
    for (int b = 0; b < n; b++) s += partial[3 * (size_t)b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

template <int STACK, bool COUNT, bool MULTI, int BLOCK>
static void launch_one(unsigned blocks, size_t paths, hipStream_t st, const PathSoA32& P,
                       const JobDev* jobs, const uint32_t* q, const uint32_t* c, uint32_t* f,
                       const uint32_t* sc, uint32_t* sf, bool shadowPhase) {
  unsigned grid = grid_for(paths, BLOCK, blocks ? blocks : 1u << 30);
  if (!shadowPhase)
    hipLaunchKernelGGL((k_extend32<STACK, COUNT, MULTI, BLOCK>), dim3(grid), dim3(BLOCK), 0, st,
                       P, jobs, q, c, f);
  else
    hipLaunchKernelGGL((k_shadow32<STACK, COUNT, MULTI, BLOCK>), dim3(grid), dim3(BLOCK), 0, st,
                       P, jobs, sc, sf);
}

template <int STACK, bool MULTI>
static void launch_trace(bool counting, int block, unsigned blocks, size_t paths, hipStream_t st,
                         const PathSoA32& P, const JobDev* jobs, const uint32_t* q,
                         const uint32_t* c, uint32_t* f, const uint32_t* sc, uint32_t* sf,
                         bool shadowPhase) {
  if (counting)
    launch_one<STACK, true, MULTI, 256>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else if (block == 64)
    launch_one<STACK, false, MULTI, 64>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else if (block == 128)
    launch_one<STACK, false, MULTI, 128>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else
    launch_one<STACK, false, MULTI, 256>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
}

template <bool MULTI>
static void dispatch_trace(int stack, bool counting, int block, unsigned blocks, size_t paths,
                           hipStream_t st, const PathSoA32& P, const JobDev* jobs,
                           const uint32_t* q, const uint32_t* c, uint32_t* f, const uint32_t* sc,
                           uint32_t* sf, bool shadowPhase) {
  if (stack == 8)
    launch_trace<8, MULTI>(counting, block, blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else if (stack == 16)
    launch_trace<16, MULTI>(counting, block, blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else
    launch_trace<24, MULTI>(counting, block, blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
}

template <bool MULTI>
static int group32_body(const RoundLaunch& R, const PathSoA32& P, size_t base, size_t count,
                        hipStream_t st, const SideStream& side);

// A group's launches; its NEE side stream is joined back to `st` on every path out.
template <bool MULTI>
static int group32(const RoundLaunch& R, const PathSoA32& P, size_t base, size_t count,
                   hipStream_t st, const SideStream& side) {
  int rc = group32_body<MULTI>(R, P, base, count, st, side);
  if (side.side) {
    int rj = check_hip(side_join(side, st), "side_join");
    if (!rc) rc = rj;
  }
  return rc;
}

template <bool MULTI>
static int group32_body(const RoundLaunch& R, const PathSoA32& P, size_t base, size_t count,
                        hipStream_t st, const SideStream& side) {
  const JobDev* jobs = R.jobs;
  unsigned gridPaths = grid_for(count, RBLOCK, 256 * 16);
  hipLaunchKernelGGL((k_generate32<MULTI>), dim3(gridPaths), dim3(RBLOCK), 0, st, P, jobs,
                     R.njobs, base, count);
  CT_LAUNCH_CHECK("k_generate32");
  for (int b = 0; b < R.nb; b++) {
    uint32_t* qin = (b & 1) ? P.q1 : P.q0;
    uint32_t* qout = (b & 1) ? P.q0 : P.q1;
    uint32_t* cin = P.ctr + b;
    uint32_t* cout = P.ctr + b + 1;
    uint32_t* scount = P.ctr + MAX_BOUNCE_SLOTS + b;
    uint32_t* fExt = P.ctr + 2 * MAX_BOUNCE_SLOTS + b;
    uint32_t* fSh = P.ctr + 3 * MAX_BOUNCE_SLOTS + b;
    if (R.mark) R.mark(R.markCtx, st, 0, true);
    dispatch_trace<MULTI>(R.stack, R.counting, R.block, R.persistBlocks, count, st, P, jobs,
                          qin, cin, fExt, scount, fSh, false);
    CT_LAUNCH_CHECK("k_extend32");
    if (R.mark) R.mark(R.markCtx, st, 0, false);
    if (side.side && b > 0) CT_HIP(side_join(side, st));  // NEE rays of bounce b-1 done
    if (R.mark) R.mark(R.markCtx, st, 1, true);
    hipLaunchKernelGGL((k_shade32<MULTI>), dim3(gridPaths), dim3(RBLOCK), 0, st, P, jobs, qin,
                       cin, qout, cout, scount, b);
    CT_LAUNCH_CHECK("k_shade32");
    if (R.mark) R.mark(R.markCtx, st, 1, false);
    hipStream_t sst = st;
    if (side.side) {
      CT_HIP(side_fork(side, st));
      sst = side.side;
    }
    if (R.mark) R.mark(R.markCtx, sst, 2, true);
    dispatch_trace<MULTI>(R.stack, R.counting, R.block, R.persistBlocks, count, sst, P, jobs,
                          qin, cin, fExt, scount, fSh, true);
    CT_LAUNCH_CHECK("k_shadow32");
    if (R.mark) R.mark(R.markCtx, sst, 2, false);
  }
  return CT_OK;
}

template <bool MULTI>
static int part32(const RoundLaunch& R, int part, hipStream_t st) {
  const PathSoA32& P = R.P32;
  const RoundStreams& rs = R.rs;
  if (part == ROUND_CLEAR) {
    hipLaunchKernelGGL(k_clear32, dim3(8), dim3(256), 0, st, P.ctr, 2 * GROUP_COUNTERS);
    CT_LAUNCH_CHECK("k_clear32");
  } else if (part == ROUND_FINAL) {
    unsigned gridPix = grid_for(R.pix, RBLOCK, 256 * 16);
    hipLaunchKernelGGL((k_finalize32<MULTI>), dim3(gridPix), dim3(RBLOCK), 0, st, P, R.jobs,
                       R.njobs, R.pix);
    CT_LAUNCH_CHECK("k_finalize32");
  } else {
    const int g = part - ROUND_GROUP0;
    size_t base, count;
    group_range(rs, R.paths, g, base, count);
    return group32<MULTI>(R, group_view(P, base, g), base, count, st, rs.shadow[g]);
  }
  return CT_OK;
}

}  // namespace f32

int launch_part32(const RoundLaunch& R, bool multi, int part, hipStream_t st) {
  return multi ? f32::part32<true>(R, part, st) : f32::part32<false>(R, part, st);
}

}  // namespace ct

using namespace ct;

extern "C" {

int ct_closest_hit32(const ct_scene* s, const float* rays, uint32_t n, float tMin,
                     const float* tMax, int32_t walk, int32_t* prim, float* t, float* u, float* v,
                     void* stream) {
  if (!s) return set_error(CT_EINVAL, "ct_closest_hit32: null scene");
  if (n == 0) return CT_OK;
  if (!rays) return set_error(CT_EINVAL, "ct_closest_hit32: null rays");
  hipStream_t st = (hipStream_t)stream;
  unsigned grid = grid_for(n, f32::QUERY_BLOCK, 256 * 32);
  if (walk == CT_WALK_WIDE)
    hipLaunchKernelGGL((f32::k_closest32<false, false>), dim3(grid), dim3(f32::QUERY_BLOCK), 0,
                       st, s->dev, rays, n, tMin, tMax, prim, t, u, v, nullptr);
  else if (walk == CT_WALK_BRUTE)
    hipLaunchKernelGGL((f32::k_closest32<true, false>), dim3(grid), dim3(f32::QUERY_BLOCK), 0,
                       st, s->dev, rays, n, tMin, tMax, prim, t, u, v, nullptr);
  else
    return set_error(CT_EINVAL, "ct_closest_hit32: walk must be CT_WALK_WIDE or CT_WALK_BRUTE");
  CT_LAUNCH_CHECK("k_closest32");
  return CT_OK;
}

int ct_closest_hit32_counts(const ct_scene* s, const float* rays, uint32_t n, float tMin,
                            const float* tMax, uint64_t* d_counts4, void* stream) {
  if (!s || !d_counts4) return set_error(CT_EINVAL, "ct_closest_hit32_counts: null argument");
  if (n == 0) return CT_OK;
  hipStream_t st = (hipStream_t)stream;
  unsigned grid = grid_for(n, f32::QUERY_BLOCK, 256 * 32);
  hipLaunchKernelGGL((f32::k_closest32<false, true>), dim3(grid), dim3(f32::QUERY_BLOCK), 0, st,
                     s->dev, rays, n, tMin, tMax, nullptr, nullptr, nullptr, nullptr,
                     (unsigned long long*)d_counts4);
  CT_LAUNCH_CHECK("k_closest32");
  return CT_OK;
}

int ct_occluded32(const ct_scene* s, const float* rays, uint32_t n, float tMin,
                  const float* tMax, int32_t walk, uint8_t* out, void* stream) {
  if (!s) return set_error(CT_EINVAL, "ct_occluded32: null scene");
  if (n == 0) return CT_OK;
  if (!rays || !out) return set_error(CT_EINVAL, "ct_occluded32: null array");
  hipStream_t st = (hipStream_t)stream;
  unsigned grid = grid_for(n, f32::QUERY_BLOCK, 256 * 32);
  if (walk == CT_WALK_WIDE)
    hipLaunchKernelGGL((f32::k_occluded32<false>), dim3(grid), dim3(f32::QUERY_BLOCK), 0, st,
                       s->dev, rays, n, tMin, tMax, out);
  else if (walk == CT_WALK_BRUTE)
    hipLaunchKernelGGL((f32::k_occluded32<true>), dim3(grid), dim3(f32::QUERY_BLOCK), 0, st,
                       s->dev, rays, n, tMin, tMax, out);
  else
    return set_error(CT_EINVAL, "ct_occluded32: walk must be CT_WALK_WIDE or CT_WALK_BRUTE");
  CT_LAUNCH_CHECK("k_occluded32");
  return CT_OK;
}

int ct_image_sum_f64(const float* d_img, uint64_t npix, double* h_sum3, void* stream) {
  if (!d_img || !h_sum3) return set_error(CT_EINVAL, "ct_image_sum_f64: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int blocks = 1024;
  double* d = nullptr;
  CT_HIP(hipMallocAsync((void**)&d, (3 * blocks + 3) * sizeof(double), st));
  hipLaunchKernelGGL(f32::k_sum_f64, dim3(blocks), dim3(256), 0, st, d_img, (size_t)npix, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(f32::k_sum_partials, dim3(1), dim3(64), 0, st, d, blocks, d + 3 * blocks);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h_sum3, d + 3 * blocks, 3 * sizeof(double),
                                          hipMemcpyDeviceToHost, st);
  hipError_t e2 = hipFreeAsync(d, st);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return check_hip(e, "ct_image_sum_f64");
}

}  // extern "C"
