// Ordered walk over the quantized 4-wide nodes, shared by the binary64 and
// the FP32 kernels.
//
// A quantized node holds, per axis, an origin and a power-of-two scale, and
// per child one byte for its lower and one for its upper plane: the child box
// is origin + q*scale, rounded outward from the float32 box of the reference
// node it stands for (DevScene::qwide).  Slab planes are evaluated in FP32 as
// t = q*B + A' with B = scale*inv (exact) and A' = (origin - o)*inv -+ E, where
// the per-axis slack E bounds every rounding of the evaluation, so near
// planes come out below and far planes above their exact values.
//
// The walk ray carries a further per-axis slack D, in units of t:
//   * D = 0 when the walk ray is the ray whose primitives are tested (FP32
//     kernels);
//   * for a binary64 ray, D covers the distance between the binary64 ray and
//     the float32 walk ray at every parameter where a primitive hit can lie.
//     A hit point lies in its primitive's padded bounds, hence in every box on
//     the path from the root to its leaf and in the root box, so the float32
//     walk enters all those boxes at the hit parameter and the binary64
//     primitive tests see every primitive the binary64 walk would.
#pragma once

#include "ct_device.h"

namespace ct {
namespace qw {

// Per-node rounding slack: t = q*B + A' with A = (origin - o)*inv and
// A' = A -+ E, E = 2^-21 * (255|B| + |A|) + D.  The rounding of A, of 1/d, of
// A' and of the FMA together stays below about 6.1 ulp-halvings (2^-24) of
// 255|B| + |A| (+ 2 of D, which carries its own 2^-10 margin), so every near
// plane lies below and every far plane above its exact value.  The slack is
// per axis: an axis where d is tiny gets a huge slack on its own planes only,
// which then lie far outside the other axes' intervals.  The same constant
// bounds the purely relative error (3 roundings) of a float32 box test.
constexpr float SLAB_REL = 4.76837158203125e-7f;
// Boxes and deferred entries are culled only beyond closest*(1 + 2^-14): the
// FP32 primitive tests' rounding, and the rounding of a binary64 closest
// distance to float32, cannot make the walk skip a primitive whose computed
// hit would be nearer.
constexpr float CULL_SLACK = 1.00006103515625f;

struct QRay {
  float ox, oy, oz, dx, dy, dz, ix, iy, iz;
  float Dx, Dy, Dz;  // per-axis slack of the planes (t units)
};

// 1/d with |d| below 1e-20 treated as +-1e-20 (intersect.ts:142 uses 1e-20 for d == 0).
__device__ __forceinline__ float safe_inv_f(float d) {
  return 1.0f / (fabsf(d) < 1e-20f ? (d < 0.0f ? -1e-20f : 1e-20f) : d);
}

__device__ __forceinline__ QRay qray(float ox, float oy, float oz, float dx, float dy, float dz) {
  QRay r;
  r.ox = ox; r.oy = oy; r.oz = oz;
  r.dx = dx; r.dy = dy; r.dz = dz;
  r.ix = safe_inv_f(dx); r.iy = safe_inv_f(dy); r.iz = safe_inv_f(dz);
  r.Dx = 0.0f; r.Dy = 0.0f; r.Dz = 0.0f;
  return r;
}

// Slack of one axis for a binary64 ray: the walk ray moves along
// deff = 1/inv (its effective direction, exactly what its planes use) from
// of = float(o); at parameters t <= T the two rays are at most
// |o - of| + T*|d - deff| apart on this axis, which is Delta*|inv| in t.
__device__ __forceinline__ float axis_slack(double o, float of, double d, float inv, double T) {
  const double deff = 1.0 / (double)inv;
  const double delta = fabs(o - (double)of) + T * fabs(d - deff);
  return (float)(delta * fabs((double)inv) * 1.0009765625);  // 2^-10 margin
}

// Walk ray of a binary64 ray.  Every point x of the root box satisfies
// |x - o| <= M = |o - c|_1 + |h|_1 (c its centre, h its half extents), so a
// hit parameter is at most M/|d|; T = 2M/|d| leaves a factor two for the
// padding and rounding of the hit point.  A zero direction gives T = inf and
// an infinite slack (every box entered; no primitive test can hit).
__device__ __forceinline__ QRay qray64(const DevScene& S, double ox, double oy, double oz,
                                       double dx, double dy, double dz) {
  QRay r = qray((float)ox, (float)oy, (float)oz, (float)dx, (float)dy, (float)dz);
  const float4 lo = gload(S.nodes), hi = gload(S.nodes + 1);
  const double cx = 0.5 * ((double)lo.x + (double)hi.x), hx = 0.5 * ((double)hi.x - (double)lo.x);
  const double cy = 0.5 * ((double)lo.y + (double)hi.y), hy = 0.5 * ((double)hi.y - (double)lo.y);
  const double cz = 0.5 * ((double)lo.z + (double)hi.z), hz = 0.5 * ((double)hi.z - (double)lo.z);
  const double M = fabs(ox - cx) + fabs(oy - cy) + fabs(oz - cz) + fabs(hx) + fabs(hy) + fabs(hz);
  const double T = 2.0 * M / sqrt(dx * dx + dy * dy + dz * dz);
  r.Dx = axis_slack(ox, r.ox, dx, r.ix, T);
  r.Dy = axis_slack(oy, r.oy, dy, r.iy, T);
  r.Dz = axis_slack(oz, r.oz, dz, r.iz, T);
  return r;
}

__device__ __forceinline__ float qbyte(uint32_t w, int k) {
  return (float)((w >> (8 * k)) & 0xffu);
}

// Deferred entries keep one byte: a lower bound of the entry distance with
// the float's exponent and top three mantissa bits (steps of at most 12.5%)
// over [2^-15, 1.75 * 2^16]; 0 stands for 0, and distances beyond the range
// keep the top code, itself a lower bound.  Truncating the mantissa rounds
// down, so a pop never skips an entry the culling bound does not exclude.
using Entry = uint8_t;
constexpr int ENTRY_BASE = (127 - 15) << 3;  // (exponent, 3 mantissa bits) of 2^-15
__device__ __forceinline__ Entry entry_down(float x) {
  int b = (int)(__float_as_uint(x) >> 20) - ENTRY_BASE + 1;
  return (Entry)min(max(b, 0), 255);
}
__device__ __forceinline__ float entry_value(Entry e) {
  return e == 0 ? 0.0f : __uint_as_float((uint32_t)(e - 1 + ENTRY_BASE) << 20);
}

// Culling bound of a walk whose closest distance (or segment end) is t.
__device__ __forceinline__ float cull_bound(float t) { return t * CULL_SLACK; }
__device__ __forceinline__ float cull_bound(double t) { return (float)t * CULL_SLACK; }

// Copies the first min(TOP, wideCount) quantized nodes of S (the top levels:
// they are numbered breadth first) into the workgroup's LDS block `top`; every
// lane of the workgroup must call it, before its first walk step.
template <int TOP>
__device__ __forceinline__ void stage_top(float4* __restrict__ top, const DevScene& S) {
  if constexpr (TOP > 0) {
    const int n = 4 * min(TOP, S.wideCount);
    for (int i = threadIdx.x; i < n; i += blockDim.x) lstore(top + i, gload(S.qwide + i));
    __syncthreads();
  }
}

// One step of the ordered walk over quantized 4-wide node `cur`: the four
// child boxes are tested (near/far planes chosen by the ray's direction
// signs), the nearest hit child continues, the others go on the LDS stack
// nearest-last with a one-byte lower bound of their entry (PUSH_ENTRY) or without
// one (any hit).  Returns false when the stack would overflow.  With TOP > 0
// a step in which every active lane of the wave is at a node below TOP reads
// the LDS copy `top` (stage_top).
template <bool PUSH_ENTRY, int STACK, bool COUNT, int TOP = 0>
__device__ __forceinline__ bool qwide_step(const DevScene& S, const QRay& R, float tCull, int& cur,
                                           int& sp, int* __restrict__ sNode,
                                           Entry* __restrict__ sEntry, int lanes,
                                           TraversalCounts& cnt,
                                           const float4* __restrict__ top = nullptr) {
  const float INF = __builtin_inff();
  float4 w0, w1, w2;
  int4 cc;
  if (TOP > 0 && __all(cur < TOP)) {  // wave-uniform: no lane waits on both paths
    const float4* w = top + 4 * cur;
    w0 = lload(w); w1 = lload(w + 1); w2 = lload(w + 2);
    const float4 c4 = lload(w + 3);
    cc = make_int4(__float_as_int(c4.x), __float_as_int(c4.y), __float_as_int(c4.z),
                   __float_as_int(c4.w));
  } else {
    const float4* w = S.qwide + 4 * (size_t)cur;
    w0 = gload(w); w1 = gload(w + 1); w2 = gload(w + 2);
    cc = gload((const int4*)(w + 3));
  }
  if (COUNT) cnt.nodes += 4;
  float Bx = w0.w * R.ix, By = w1.x * R.iy, Bz = w1.y * R.iz;
  float Ax = (w0.x - R.ox) * R.ix, Ay = (w0.y - R.oy) * R.iy, Az = (w0.z - R.oz) * R.iz;
  float Ex = fmaf(SLAB_REL, fmaf(255.0f, fabsf(Bx), fabsf(Ax)), R.Dx);
  float Ey = fmaf(SLAB_REL, fmaf(255.0f, fabsf(By), fabsf(Ay)), R.Dy);
  float Ez = fmaf(SLAB_REL, fmaf(255.0f, fabsf(Bz), fabsf(Az)), R.Dz);
  float Anx = Ax - Ex, Any = Ay - Ey, Anz = Az - Ez;  // near planes: lower bounds
  float Afx = Ax + Ex, Afy = Ay + Ey, Afz = Az + Ez;  // far planes: upper bounds
  uint32_t lx = __float_as_uint(w1.z), ly = __float_as_uint(w1.w), lz = __float_as_uint(w2.x);
  uint32_t hx = __float_as_uint(w2.y), hy = __float_as_uint(w2.z), hz = __float_as_uint(w2.w);
  uint32_t nx = R.ix < 0 ? hx : lx, fx = R.ix < 0 ? lx : hx;
  uint32_t ny = R.iy < 0 ? hy : ly, fy = R.iy < 0 ? ly : hy;
  uint32_t nz = R.iz < 0 ? hz : lz, fz = R.iz < 0 ? lz : hz;
  float e[4];
  int c[4] = {cc.x, cc.y, cc.z, cc.w};
  int n = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    float tn = fmaxf(fmaxf(fmaf(qbyte(nx, k), Bx, Anx), fmaf(qbyte(ny, k), By, Any)),
                     fmaf(qbyte(nz, k), Bz, Anz));
    float tf = fminf(fminf(fmaf(qbyte(fx, k), Bx, Afx), fmaf(qbyte(fy, k), By, Afy)),
                     fmaf(qbyte(fz, k), Bz, Afz));
    tn = fmaxf(tn, 0.0f);
    tf = fminf(tf, tCull);
    bool hit = tn <= tf && c[k] != -1;
    e[k] = hit ? tn : INF;
    n += hit ? 1 : 0;
  }
#define CT_CSWAP(a, b)                                        \
  {                                                           \
    bool sw_ = e[b] < e[a];                                   \
    float lo_ = sw_ ? e[b] : e[a], hi_ = sw_ ? e[a] : e[b];   \
    int cl_ = sw_ ? c[b] : c[a], ch_ = sw_ ? c[a] : c[b];     \
    e[a] = lo_; e[b] = hi_; c[a] = cl_; c[b] = ch_;           \
  }
  CT_CSWAP(0, 1) CT_CSWAP(2, 3) CT_CSWAP(0, 2) CT_CSWAP(1, 3) CT_CSWAP(1, 2)
#undef CT_CSWAP
  if (n == 0) {
    cur = -1;
    return true;
  }
  if (sp + n - 1 > STACK) return false;
  for (int k = 3; k >= 1; k--) {
    if (k < n) {
      sNode[sp * lanes] = c[k];
      if (PUSH_ENTRY) sEntry[sp * lanes] = entry_down(e[k]);
      sp++;
    }
  }
  cur = c[0];
  return true;
}

// Pops deferred work, skipping entries beyond the culling bound.
__device__ __forceinline__ void qpop(int& cur, int& sp, const int* __restrict__ sNode,
                                     const Entry* __restrict__ sEntry, int lanes,
                                     float tCull) {
  while (sp > 0) {
    sp--;
    if (entry_value(sEntry[sp * lanes]) <= tCull) {
      cur = sNode[sp * lanes];
      return;
    }
  }
}

}  // namespace qw
}  // namespace ct
