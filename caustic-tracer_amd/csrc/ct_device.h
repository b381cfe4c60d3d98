// Device-side primitives of the gfx950 path tracer.
//
// Arithmetic contract: the scene stays in its packed float32 layout in HBM
// (pack.ts:1-17, bvh.ts:1-16); every quantity derived from it is computed in
// binary64, in the operation order of the reference (a JS engine computes in
// doubles).  CDNA4 issues v_fma/v_mul/v_add_f64 at the same lane rate as
// unpacked FP32, so the double-precision twin costs registers, not issue
// slots, and closest-hit IDs, t, u, v come out bit-identical to the
// reference's.  The library is compiled with -ffp-contract=off so no a*b+c is
// fused; f64 division and sqrt are the correctly rounded expansions.  Which
// boxes a walk enters is not part of that contract, only that it enters every
// box holding a primitive the reference's walk would test: the 4-wide walk
// (ct_qwalk.h) decides boxes in FP32 with conservative slack.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "caustic_amd.h"

namespace ct {

constexpr int PRIM_STRIDE = 16;  // floats per primitive record (pack.ts:30)
constexpr int MAT_STRIDE = 16;   // floats per material record (pack.ts:31)
constexpr int NODE_STRIDE = 8;   // floats per BVH node (bvh.ts:18)
constexpr double T_MIN = 1e-4;   // intersect.ts:14
constexpr double PI_D = 3.141592653589793;
constexpr double TWO_PI = 2 * PI_D;
constexpr double INV_PI = 1 / PI_D;
constexpr double SHADOW_SHORTEN = 1 - 1e-3;  // tracer.ts:36
// Quantized 4-wide nodes numbered breadth first at upload: the root and the
// levels below it, up to this many nodes, form one block at the start of the
// array, which single-scene traversal launches copy into LDS.  16 nodes (1 KB)
// keep the 24-entry stack kernels within 31 KB of LDS per 256-lane workgroup;
// 32 nodes (32 KB) cost the binary64 extend launches 5% on the 1M-triangle
// soup, which reads as one workgroup fewer per CU.
constexpr int TOP_NODES = 16;

// Read-only view of a device-resident packed scene.
struct DevScene {
  const float4* __restrict__ prims;  // 4 float4 per primitive
  const float4* __restrict__ nodes;  // 2 float4 per node
  const float4* __restrict__ mats;   // 4 float4 per material
  const int32_t* __restrict__ lights;
  const uint32_t* __restrict__ primObject;
  // Exact leaf words (start*8 + count; 0xffffffff on interior nodes), one per
  // reference node: the float32 words of the reference layout are exact only
  // below 2^21 primitives.
  const uint32_t* __restrict__ leafWords;
  // 4-wide traversal nodes derived from `nodes` at upload, quantized, 4
  // float4 (64 B) each: [origin.xyz, scale.x | scale.y, scale.z, qlo.x, qlo.y |
  // qlo.z, qhi.x, qhi.y, qhi.z | child code[4]]; q* hold one byte per child; a
  // child box is origin + q*scale (power-of-two scales), rounded outward at
  // upload from the float32 box of the reference node it stands for.  Child
  // code >= 0 another wide node, <= -2 a leaf (-2 - leaf word), -1 an empty
  // slot.
  const float4* __restrict__ qwide;
  int wideCount;
  int binaryExact;  // every float32 leaf word equals its exact word
  int primCount;
  int nodeCount;
  int lightCount;
  int skyKind;  // 0 none, 1 gradient
  double zenith[3], horizon[3], skyStrength;
};

// ---------------------------------------------------------------------------
// RNG: PCG RXS-M-XS seeding + xorshift32 stream (rng.ts:10-64)

__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t shifted = (state >> ((state >> 28) + 4u)) ^ state;
  uint32_t word = shifted * 277803737u;
  return (word >> 22) ^ word;
}

__device__ __forceinline__ uint32_t seed_for(uint32_t pixel, uint32_t frame) {
  uint32_t a = pcg_hash(pixel);
  uint32_t b = pcg_hash(frame ^ 0x9e3779b9u);
  uint32_t s = pcg_hash(a ^ (b * 0x85ebca6bu));
  return s == 0 ? 0x1234567u : s;
}

struct Rng {
  uint32_t s;
  __device__ __forceinline__ uint32_t u32() {
    uint32_t x = s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    s = x;
    return x;
  }
  // u32 / 2^32 exactly (a power-of-two scale is exact in binary64)
  __device__ __forceinline__ double next() { return (double)u32() * 2.3283064365386963e-10; }
  __device__ __forceinline__ int pick(int n) {
    double f = floor(next() * n);
    double m = (double)(n - 1);
    return (int)(m < f ? m : f);
  }
};

// ---------------------------------------------------------------------------
// Packed-record access

struct PrimGeom {
  float4 a;  // a.xyz, type
  float4 b;  // b.xyz (sphere: radius in x), object
  float4 c;  // c.xyz, area
};

// Loads through the global address space.  A scene view read out of device
// memory (the job table) carries generic pointers, and a load through one is a
// FLAT access, which also counts against the LDS wait counter: LDS stack reads
// would wait for node fetches in flight.  The scene arrays are global memory,
// so their loads go through address-space-1 pointers of native vector types
// (a load of HIP's float4 class would convert back to a generic reference).
typedef float ct_nf4 __attribute__((ext_vector_type(4)));
typedef int ct_ni4 __attribute__((ext_vector_type(4)));
#define CT_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ float4 gload(const float4* p) {
  ct_nf4 v = *(const CT_GLOBAL ct_nf4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 gload(const int4* p) {
  ct_ni4 v = *(const CT_GLOBAL ct_ni4*)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
// LDS block of the top traversal nodes, read and written through
// address-space-3 pointers (ds_read/ds_write, never FLAT).
#define CT_LDS __attribute__((address_space(3)))
__device__ __forceinline__ float4 lload(const float4* p) {
  ct_nf4 v = *(const CT_LDS ct_nf4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void lstore(float4* p, float4 x) {
  ct_nf4 v = {x.x, x.y, x.z, x.w};
  *(CT_LDS ct_nf4*)p = v;
}
__device__ __forceinline__ uint32_t gload(const uint32_t* p) {
  return *(const CT_GLOBAL uint32_t*)p;
}
__device__ __forceinline__ int32_t gload(const int32_t* p) {
  return *(const CT_GLOBAL int32_t*)p;
}

// Loads through the constant address space: the scene arrays are not written
// while a kernel runs, and a load whose address is the same on every lane of
// a wave then goes through the scalar cache (other addresses: vector loads).
#define CT_CONST __attribute__((address_space(4)))
__device__ __forceinline__ float4 cload(const float4* p) {
  ct_nf4 v = *(const CT_CONST ct_nf4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// UNIFORM: callers whose primitive index is the same on every lane of a wave
// (the scans over all primitives) read the record through the scalar cache.
template <bool UNIFORM = false>
__device__ __forceinline__ PrimGeom load_prim(const DevScene& S, int p) {
  const float4* r = S.prims + (size_t)p * 4;
  PrimGeom g;
  if (UNIFORM) {
    g.a = cload(r);
    g.b = cload(r + 1);
    g.c = cload(r + 2);
    return g;
  }
  g.a = gload(r);
  g.b = gload(r + 1);
  g.c = gload(r + 2);
  // The third vertex is needed by triangles only, so the compiler would sink
  // its load into the triangle branch, issued only after the type test has
  // waited on the first two: a second memory round trip per triangle.  An
  // empty statement that takes its registers keeps the three loads together.
  __asm__ volatile("" ::"v"(g.c.x), "v"(g.c.y), "v"(g.c.z));
  return g;
}

// What shading reads of the primitive a path hit: the record's first two
// float4 (a sphere's centre and radius), its stored normal (triangles) and
// its object index, all loads issued together (the normal would otherwise
// wait behind the type test, and the object index behind the normal).
struct HitGeom {
  float4 a, b, n;
  uint32_t obj;
};

// A sampled light's whole record (a, b, c: sphere or triangle; n: stored
// normal) and its object index, all loads issued together.
struct LightGeom {
  float4 a, b, c, n;
  uint32_t obj;
};
__device__ __forceinline__ LightGeom load_light(const DevScene& S, int prim) {
  const float4* r = S.prims + (size_t)prim * 4;
  LightGeom g;
  g.a = gload(r);
  g.b = gload(r + 1);
  g.c = gload(r + 2);
  g.n = gload(r + 3);
  g.obj = gload(S.primObject + prim);
  __asm__ volatile("" ::"v"(g.a.w), "v"(g.b.x), "v"(g.c.x), "v"(g.n.x), "v"(g.obj));
  return g;
}
__device__ __forceinline__ HitGeom load_hit(const DevScene& S, int prim) {
  const float4* r = S.prims + (size_t)prim * 4;
  HitGeom g;
  g.a = gload(r);
  g.b = gload(r + 1);
  g.n = gload(r + 3);
  g.obj = gload(S.primObject + prim);
  __asm__ volatile("" ::"v"(g.a.w), "v"(g.b.x), "v"(g.n.x), "v"(g.obj));
  return g;
}

// Math.max(0, x): +0 for x <= 0 (including -0), NaN stays NaN.
__device__ __forceinline__ double max0(double x) { return x > 0 ? x : (x != x ? x : 0.0); }

// ---------------------------------------------------------------------------
// Intersection (intersect.ts:20-123), binary64 over float32 inputs

// TIE: a miss whose root equals tMax (a hit the strict bound rejects only
// because it ties the current closest one) returns -2 instead of -1.
template <bool TIE = false>
__device__ __forceinline__ double ray_sphere(double ox, double oy, double oz, double dx,
                                             double dy, double dz, double cx, double cy,
                                             double cz, double r, double tMin, double tMax) {
  double lx = ox - cx, ly = oy - cy, lz = oz - cz;
  double a = dx * dx + dy * dy + dz * dz;
  double halfB = lx * dx + ly * dy + lz * dz;
  double c = lx * lx + ly * ly + lz * lz - r * r;
  double disc = halfB * halfB - a * c;
  if (disc < 0) return -1;
  double s = sqrt(disc);
  double t = (-halfB - s) / a;
  if (t <= tMin || t >= tMax) {
    bool tied = TIE && t == tMax;
    t = (-halfB + s) / a;
    if (t <= tMin || t >= tMax) return (tied || (TIE && t == tMax)) ? -2 : -1;
  }
  return t;
}

// Two-sided Moller-Trumbore; returns true and t,u,v on a hit in (tMin, tMax).
// The arithmetic of every value returned is the reference's.  Before the
// division the barycentric numerators are checked against |det|: a triangle
// is dropped early only when u = num_u/det is certainly < 0 or > 1 (margins
// of 1e-14 relative and 1e-250 absolute keep rounding and underflow out of
// the decision), or likewise for v and u + v, i.e. only when the reference's
// own tests would reject it; otherwise the reference's computation runs.
// TIE: returns 2 instead of 0 for a hit rejected only because t == tMax.
template <bool TIE = false>
__device__ __forceinline__ int ray_triangle(double ox, double oy, double oz, double dx,
                                            double dy, double dz, double ax, double ay,
                                            double az, double bx, double by, double bz,
                                            double cx, double cy, double cz, double tMin,
                                            double tMax, double& tOut, double& uOut,
                                            double& vOut) {
  double e1x = bx - ax, e1y = by - ay, e1z = bz - az;
  double e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
  double px = dy * e2z - dz * e2y;
  double py = dz * e2x - dx * e2z;
  double pz = dx * e2y - dy * e2x;
  double det = e1x * px + e1y * py + e1z * pz;
  if (det > -1e-12 && det < 1e-12) return false;
  double tx = ox - ax, ty = oy - ay, tz = oz - az;
  double un = tx * px + ty * py + tz * pz;
  double qx = ty * e1z - tz * e1y;
  double qy = tz * e1x - tx * e1z;
  double qz = tx * e1y - ty * e1x;
  double vn = dx * qx + dy * qy + dz * qz;
  {
    double ad = fabs(det);
    double su = det > 0 ? un : -un, sv = det > 0 ? vn : -vn;
    double lim = ad * (1 + 1e-14);
    if (su < -1e-250 || su > lim || sv < -1e-250 || su + sv > lim * (1 + 1e-14)) return false;
  }
  double invDet = 1 / det;
  double u = un * invDet;
  if (u < 0 || u > 1) return false;
  double v = vn * invDet;
  if (v < 0 || u + v > 1) return false;
  double t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
  if (t <= tMin || t >= tMax) return (TIE && t == tMax) ? 2 : 0;
  tOut = t;
  uOut = u;
  vOut = v;
  return 1;
}

// Slab test; returns +inf on a miss (entry before tMax required), else max(tNear,0).
__device__ __forceinline__ double ray_aabb(double ox, double oy, double oz, double idx,
                                           double idy, double idz, float4 lo, float4 hi,
                                           double tMax) {
  double t0 = ((double)lo.x - ox) * idx;
  double t1 = ((double)hi.x - ox) * idx;
  double tNear = t0 < t1 ? t0 : t1;
  double tFar = t0 < t1 ? t1 : t0;
  t0 = ((double)lo.y - oy) * idy;
  t1 = ((double)hi.y - oy) * idy;
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0;
    if (t1 < tFar) tFar = t1;
  } else {
    if (t1 > tNear) tNear = t1;
    if (t0 < tFar) tFar = t0;
  }
  t0 = ((double)lo.z - oz) * idz;
  t1 = ((double)hi.z - oz) * idz;
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0;
    if (t1 < tFar) tFar = t1;
  } else {
    if (t1 > tNear) tNear = t1;
    if (t0 < tFar) tFar = t0;
  }
  if (tNear > tFar || tFar < 0 || tNear > tMax) return __builtin_inf();
  return tNear < 0 ? 0 : tNear;
}

__device__ __forceinline__ double safe_inv(double d) { return 1 / (d == 0 ? 1e-20 : d); }

struct Hit {
  double t, u, v;
  int prim;
  bool tie;  // a primitive was hit at exactly t (TIES_DETECT)
};

// How a closest-hit walk settles primitives hit at exactly the same t.
enum TieMode : int {
  TIES_FIRST = 0,    // the first tested stays (intersectClosestStackless: its DFS
                     // order makes that the lower packed index)
  TIES_DETECT = 1,   // mark the tie in Hit::tie; settle_ties re-walks the ray
  TIES_ORDERED = 2,  // the one intersectClosest reaches first (ordered_first)
};

struct TraversalCounts {
  uint32_t nodes, tris, spheres, fallbacks;
};

// Which of two primitives hit at the same t the reference's ordered walk
// (intersect.ts:202-282) reaches first.  Its leaves are visited depth first,
// at each interior node the child with the smaller slab entry first (tl <= tr:
// left), and on an exact tie the primitive found first stays (strict '<'), so
// the winner lies under the child taken first at the two primitives' lowest
// common ancestor -- or, in one leaf, is the lower packed index.  Subtrees
// hold contiguous packed ranges (DFS pre-order), so the ancestor is found by
// descending with the first primitive of each right subtree (the leaf at the
// end of its left spine).  Both children of that ancestor contain a hit at
// the tied t, so both entries are finite and the reference's comparison is
// made on them, with its own slab arithmetic.  Called only on exact ties.
__device__ __forceinline__ bool ordered_first(const DevScene& S, int p, int q,
                                                        double ox, double oy, double oz,
                                                        double dx, double dy, double dz) {
  const double ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
  int c = 0;
  for (;;) {
    if (gload(S.nodes + 2 * (size_t)c + 1).w >= 0.0f) return p < q;  // one leaf
    int L = c + 1;
    float4 llo = gload(S.nodes + 2 * (size_t)L), lhi = gload(S.nodes + 2 * (size_t)L + 1);
    int R = (int)llo.w;
    int x = R;
    while (gload(S.nodes + 2 * (size_t)x + 1).w < 0.0f) x++;  // left spine of R
    int mid = (int)(gload(S.leafWords + x) >> 3);
    bool pl = p < mid, ql = q < mid;
    if (pl == ql) {
      c = pl ? L : R;
      continue;
    }
    float4 rlo = gload(S.nodes + 2 * (size_t)R), rhi = gload(S.nodes + 2 * (size_t)R + 1);
    double tl = ray_aabb(ox, oy, oz, ix, iy, iz, llo, lhi, __builtin_inf());
    double tr = ray_aabb(ox, oy, oz, ix, iy, iz, rlo, rhi, __builtin_inf());
    return pl == (tl <= tr);
  }
}

// The next binary64 above a positive t (+inf stays): hits up to and including
// the current closest t come back from the primitive tests, so exact ties can
// be resolved.
__device__ __forceinline__ double tie_bound(double t) {
  return t < __builtin_inf() ? __longlong_as_double(__double_as_longlong(t) + 1) : t;
}

// Leaf primitive loop shared by the closest-hit walks (intersect.ts:155-193,
// 324-351): strict '<' keeps the first prim tested on exact ties; a sphere hit
// leaves u, v as the last triangle hit set them.  Walks that stand for
// intersectClosest but do not share its visiting order (the 4-wide and parked
// walks, the FMA-slab binary walk) run TIES_DETECT: a primitive whose hit the
// strict bound rejects only because it equals the current t sets Hit::tie
// (a hit exactly at a query's own tMax sets it too: the second walk then
// finds nothing more), and settle_ties walks that ray again with
// TIES_ORDERED, where such a tie goes to the primitive intersectClosest
// reaches first (ordered_first) whatever the order of the tests.
template <bool COUNT, int TIES = TIES_DETECT, bool UNIFORM = false>
__device__ __forceinline__ void test_leaf(const DevScene& S, int leafWord, double ox, double oy,
                                          double oz, double dx, double dy, double dz,
                                          double tMin, Hit& h, TraversalCounts& cnt) {
  int li = leafWord;
  int count = li & 7;
  int start = li >> 3;
  for (int p = start; p < start + count; p++) {
    PrimGeom g = load_prim<UNIFORM>(S, p);
    // TIES_ORDERED sees hits up to and including h.t; the others keep the
    // strict bound, TIES_DETECT noting a primitive rejected only by a tie
    const double tMax = TIES == TIES_ORDERED ? tie_bound(h.t) : h.t;
    constexpr bool DET = TIES == TIES_DETECT;
    double t, u = 0, v = 0;
    bool hit;
    if (g.a.w == 0.0f) {  // PRIM_SPHERE
      if (COUNT) cnt.spheres++;
      t = ray_sphere<DET>(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, tMin, tMax);
      hit = t > 0;
      if (DET && t == -2) h.tie = true;
    } else {
      if (COUNT) cnt.tris++;
      int r = ray_triangle<DET>(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z,
                                g.c.x, g.c.y, g.c.z, tMin, tMax, t, u, v);
      hit = r == 1;
      if (DET && r == 2) h.tie = true;
    }
    if (!hit) continue;
    bool take = TIES != TIES_ORDERED || t < h.t;
    if (!take && h.prim >= 0)  // t == h.t: an exact tie with the current hit
      take = ordered_first(S, p, h.prim, ox, oy, oz, dx, dy, dz);
    if (take) {
      h.t = t;
      h.prim = p;
      if (g.a.w != 0.0f) {
        h.u = u;
        h.v = v;
      }
    }
  }
}

// Stackless skip-link closest hit (intersect.ts:290-365): node hit -> node+1,
// miss -> skip.  h.t enters as tMax and leaves as the closest t (prim -1 = miss).
// TIES: TIES_FIRST for intersectClosestStackless itself, whose DFS order
// gives ties to the lower packed index; TIES_DETECT as the fallback of a walk
// that stands for intersectClosest; TIES_ORDERED inside settle_ties.
template <bool COUNT, int TIES = TIES_DETECT>
__device__ __forceinline__ void closest_stackless(const DevScene& S, double ox, double oy,
                                                  double oz, double dx, double dy, double dz,
                                                  double tMin, Hit& h, TraversalCounts& cnt) {
  double idx = safe_inv(dx), idy = safe_inv(dy), idz = safe_inv(dz);
  int node = 0;
  const int nodeCount = S.nodeCount;
  while (node < nodeCount) {
    float4 lo = gload(S.nodes + 2 * (size_t)node);
    float4 hi = gload(S.nodes + 2 * (size_t)node + 1);
    if (COUNT) cnt.nodes++;
    double entry = ray_aabb(ox, oy, oz, idx, idy, idz, lo, hi, h.t);
    if (entry == __builtin_inf()) {
      node = (int)lo.w;
      continue;
    }
    if (hi.w >= 0.0f)
      test_leaf<COUNT, TIES>(S, (int)gload(S.leafWords + node), ox, oy, oz, dx, dy, dz, tMin, h,
                             cnt);
    node++;
  }
}

// A ray of a TIES_DETECT walk that met an exact tie is walked again from
// (tMin, tMax) with the ties settled in intersectClosest's order; rays
// without one keep their hit.  Exact ties need shared edges or vertices hit
// at bit-identical t, so the second walk is rare.
template <bool COUNT>
__device__ __forceinline__ void settle_ties(const DevScene& S, double ox, double oy, double oz,
                                            double dx, double dy, double dz, double tMin,
                                            double tMax, Hit& h, TraversalCounts& cnt) {
  if (!h.tie) return;
  h.t = tMax;
  h.u = 0;
  h.v = 0;
  h.prim = -1;
  h.tie = false;
  closest_stackless<COUNT, TIES_ORDERED>(S, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
}

// Stackless any-hit on (tMin, tMax) (intersect.ts:368-420).
template <bool COUNT>
__device__ __forceinline__ bool occluded(const DevScene& S, double ox, double oy, double oz,
                                         double dx, double dy, double dz, double tMin,
                                         double tMax, TraversalCounts& cnt) {
  double idx = safe_inv(dx), idy = safe_inv(dy), idz = safe_inv(dz);
  int node = 0;
  const int nodeCount = S.nodeCount;
  while (node < nodeCount) {
    float4 lo = gload(S.nodes + 2 * (size_t)node);
    float4 hi = gload(S.nodes + 2 * (size_t)node + 1);
    if (COUNT) cnt.nodes++;
    double entry = ray_aabb(ox, oy, oz, idx, idy, idz, lo, hi, tMax);
    if (entry == __builtin_inf()) {
      node = (int)lo.w;
      continue;
    }
    if (hi.w >= 0.0f) {
      int li = (int)gload(S.leafWords + node);
      int count = li & 7;
      int start = li >> 3;
      for (int p = start; p < start + count; p++) {
        PrimGeom g = load_prim(S, p);
        if (g.a.w == 0.0f) {
          if (COUNT) cnt.spheres++;
          if (ray_sphere(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, tMin, tMax) > 0)
            return true;
        } else {
          if (COUNT) cnt.tris++;
          double t, u, v;
          if (ray_triangle(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z,
                           g.c.x, g.c.y, g.c.z, tMin, tMax, t, u, v))
            return true;
        }
      }
    }
    node++;
  }
  return false;
}

// Largest float not above x (stack entries are stored rounded down so that a
// pop skips a subtree only when the binary64 entry would have been skipped).
__device__ __forceinline__ float f32_round_down(double x) {
  float f = (float)x;
  if ((double)f > x) {  // step one float toward -inf
    uint32_t b = __float_as_uint(f);
    if (f > 0.0f) b -= 1;
    else if (f < 0.0f) b += 1;
    else b = 0x80000001u;  // -smallest subnormal
    f = __uint_as_float(b);
  }
  return f;
}

// ---------------------------------------------------------------------------
// Wave-friendly ordered walk.
//
// Same slab values as ray_aabb ((b - o) * inv in binary64) reduced with
// min/max instead of the reference's compare-and-assign chains; the two agree
// on every hit/miss decision and entry value (they can differ only in the
// sign of a zero entry, which no comparison distinguishes).
__device__ __forceinline__ bool box_enter(double ox, double oy, double oz, double ix, double iy,
                                          double iz, float4 lo, float4 hi, double tMax,
                                          double& entry) {
  double tx0 = ((double)lo.x - ox) * ix, tx1 = ((double)hi.x - ox) * ix;
  double ty0 = ((double)lo.y - oy) * iy, ty1 = ((double)hi.y - oy) * iy;
  double tz0 = ((double)lo.z - oz) * iz, tz1 = ((double)hi.z - oz) * iz;
  double tn = fmax(fmax(fmin(tx0, tx1), fmin(ty0, ty1)), fmax(fmin(tz0, tz1), 0.0));
  double tf = fmin(fmin(fmax(tx0, tx1), fmax(ty0, ty1)), fmin(fmax(tz0, tz1), tMax));
  entry = tn;
  return tn <= tf;
}

// Slab test with one fused multiply-add per plane: t = b*inv - (o*inv).  Its
// values differ from (b - o)*inv by about 1e-16 relative (a few 1e-15 scene
// units), while every box is padded by at least 1e-6 beyond the primitives it
// holds (pack.ts:153,177, bvh.ts:140-147).  A ray that hits a primitive inside
// a box therefore crosses that box with a margin no rounding difference can
// remove: boxes holding a closer hit are accepted exactly when the reference
// accepts them, and the closest hit is unchanged.  oi = o * inv per ray.
__device__ __forceinline__ bool box_enter_fma(double ix, double iy, double iz, double oix,
                                              double oiy, double oiz, float4 lo, float4 hi,
                                              double tMax, double& entry) {
  double tx0 = fma((double)lo.x, ix, -oix), tx1 = fma((double)hi.x, ix, -oix);
  double ty0 = fma((double)lo.y, iy, -oiy), ty1 = fma((double)hi.y, iy, -oiy);
  double tz0 = fma((double)lo.z, iz, -oiz), tz1 = fma((double)hi.z, iz, -oiz);
  double tn = fmax(fmax(fmin(tx0, tx1), fmin(ty0, ty1)), fmax(fmin(tz0, tz1), 0.0));
  double tf = fmin(fmin(fmax(tx0, tx1), fmax(ty0, ty1)), fmin(fmax(tz0, tz1), tMax));
  entry = tn;
  return tn <= tf;
}

// Stack entry distance as bfloat16 rounded toward zero (entries are >= 0):
// the top 16 bits of the float32 rounded down -- a lower bound of the binary64
// entry, so a pop skips a subtree only when the reference would have.
__device__ __forceinline__ uint16_t entry_bf16_down(double x) {
  return (uint16_t)(__float_as_uint(f32_round_down(x)) >> 16);
}
__device__ __forceinline__ double entry_from_bf16(uint16_t b) {
  return (double)__uint_as_float((uint32_t)b << 16);
}

// Work item encoding shared by the stack and the current-node register:
// >= 0 interior node index, <= -2 a leaf (-2 - leafWord), -1 nothing.
__device__ __forceinline__ int enc_child(const DevScene& S, int idx, float leafWord) {
  return leafWord >= 0.0f ? -2 - (int)gload(S.leafWords + idx) : idx;
}

// Closest hit with the ordered walk of intersect.ts:202-282 (nearer child
// first, far child deferred with its entry, deferred entries >= closest
// skipped) organised for 64-wide waves: interior steps run in an inner loop;
// a lane that reaches a leaf parks it and keeps descending until every lane
// of the wave holds a leaf (or has nothing left), then all parked leaves are
// tested together.  Parking a leaf only delays when `closest` shrinks, so the
// walk visits a superset of what the reference visits and returns the same
// minimum.  Stack overflow falls back to the stackless walk from the root.
template <int STACK, bool COUNT>
__device__ __forceinline__ void closest_wave(const DevScene& S, double ox, double oy, double oz,
                                            double dx, double dy, double dz, double tMin, Hit& h,
                                            int* __restrict__ sNode, float* __restrict__ sEntry,
                                            int lanes, TraversalCounts& cnt) {
  const double ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
  if (COUNT) cnt.nodes++;
  int cur = -1;
  {
    float4 lo = gload(S.nodes + 0), hi = gload(S.nodes + 1);
    double e;
    if (box_enter(ox, oy, oz, ix, iy, iz, lo, hi, h.t, e)) cur = enc_child(S, 0, hi.w);
  }
  int sp = 0;
  int pend = -1;  // parked leaf word
  bool overflow = false;
  while (cur != -1 || pend >= 0) {
    // descend
    while (cur != -1) {
      if (cur <= -2) {  // reached a leaf
        if (pend >= 0) break;
        pend = -2 - cur;
        cur = -1;
      } else {
        int left = cur + 1;
        float4 llo = gload(S.nodes + 2 * (size_t)left);
        float4 lhi = gload(S.nodes + 2 * (size_t)left + 1);
        int right = (int)llo.w;
        float4 rlo = gload(S.nodes + 2 * (size_t)right);
        float4 rhi = gload(S.nodes + 2 * (size_t)right + 1);
        if (COUNT) cnt.nodes += 2;
        double tl, tr;
        bool hl = box_enter(ox, oy, oz, ix, iy, iz, llo, lhi, h.t, tl);
        bool hr = box_enter(ox, oy, oz, ix, iy, iz, rlo, rhi, h.t, tr);
        int el = enc_child(S, left, lhi.w), er = enc_child(S, right, rhi.w);
        if (hl && hr) {
          if (sp == STACK) {
            overflow = true;
            cur = -1;
            break;
          }
          bool leftFirst = tl <= tr;
          sNode[sp * lanes] = leftFirst ? er : el;
          sEntry[sp * lanes] = f32_round_down(leftFirst ? tr : tl);
          sp++;
          cur = leftFirst ? el : er;
        } else if (hl) {
          cur = el;
        } else if (hr) {
          cur = er;
        } else {
          cur = -1;
        }
      }
      if (cur == -1) {  // pop, skipping entries the bound already excludes
        while (sp > 0) {
          sp--;
          if ((double)sEntry[sp * lanes] < h.t) {
            cur = sNode[sp * lanes];
            break;
          }
        }
      }
      if (__all(pend >= 0 || cur == -1)) break;
    }
    if (overflow) break;
    // test the parked leaf
    if (pend >= 0) {
      test_leaf<COUNT>(S, pend, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
      pend = -1;
    }
  }
  if (overflow) {
    if (COUNT) cnt.fallbacks++;
    if (pend >= 0) test_leaf<COUNT>(S, pend, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
    closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, tMin, h, cnt);
  }
}

// One interior step of the binary ordered walk: test both children of
// reference node `cur`, continue with the nearer, defer the other on the LDS
// stack with a bf16 lower bound of its entry (PUSH_ENTRY) or without one (any
// hit).  Returns false when the stack would overflow.
template <bool PUSH_ENTRY, int STACK, bool COUNT>
__device__ __forceinline__ bool interior_step(const DevScene& S, double ix, double iy, double iz,
                                              double oix, double oiy, double oiz, double tMax,
                                              int& cur, int& sp, int* __restrict__ sNode,
                                              uint16_t* __restrict__ sEntry, int lanes,
                                              TraversalCounts& cnt) {
  int left = cur + 1;
  float4 llo = gload(S.nodes + 2 * (size_t)left), lhi = gload(S.nodes + 2 * (size_t)left + 1);
  int right = (int)llo.w;
  float4 rlo = gload(S.nodes + 2 * (size_t)right), rhi = gload(S.nodes + 2 * (size_t)right + 1);
  if (COUNT) cnt.nodes += 2;
  double tl, tr;
  bool hl = box_enter_fma(ix, iy, iz, oix, oiy, oiz, llo, lhi, tMax, tl);
  bool hr = box_enter_fma(ix, iy, iz, oix, oiy, oiz, rlo, rhi, tMax, tr);
  int el = enc_child(S, left, lhi.w), er = enc_child(S, right, rhi.w);
  if (hl && hr) {
    if (sp == STACK) return false;
    bool leftFirst = tl <= tr;
    sNode[sp * lanes] = leftFirst ? er : el;
    if (PUSH_ENTRY) sEntry[sp * lanes] = entry_bf16_down(leftFirst ? tr : tl);
    sp++;
    cur = leftFirst ? el : er;
  } else if (hl) {
    cur = el;
  } else if (hr) {
    cur = er;
  } else {
    cur = -1;
  }
  return true;
}

// Code of the binary walk's root: the root box is tested first, as
// intersect.ts:220-227 does.
__device__ __forceinline__ int walk_root(const DevScene& S, double ix, double iy, double iz,
                                         double oix, double oiy, double oiz, double tMax,
                                         TraversalCounts& cnt, bool count) {
  if (count) cnt.nodes++;
  float4 lo = gload(S.nodes + 0), hi = gload(S.nodes + 1);
  double e;
  return box_enter_fma(ix, iy, iz, oix, oiy, oiz, lo, hi, tMax, e) ? enc_child(S, 0, hi.w) : -1;
}

}  // namespace ct
