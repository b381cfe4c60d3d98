// Wavefront path tracer: the per-path loop of radianceInto (tracer.ts:191-328)
// unrolled into bounce-synchronous kernels over SoA path state in HBM.
//
//   generate   camera ray + PCG seed per path             tracer.ts:371-384
//   extend     closest hit per live path                  tracer.ts:210
//   shade      emission+MIS, BSDF sample, NEE setup, RR    tracer.ts:211-322
//              -> live queue for the next bounce, shadow queue for this one
//   shadow     any-hit of the NEE ray, adds its pending contribution
//   finalize   per pixel, samples added in pass order     tracer.ts:386-392
//
// A path is (pass, pixel); all passes of a batch run together and the
// finalize kernel adds their radiance in pass order, so the float32 sums are
// the ones accumulatePass produces pass after pass.  The NEE term of bounce b
// is added by the shadow kernel before bounce b+1 is shaded, keeping the
// reference's order of additions into L.  Live paths are compacted with a
// wave64 ballot + one atomic per wave.
#include <string.h>

#include <vector>

#include "ct_internal.h"

namespace ct {

constexpr int RBLOCK = 256;
constexpr int MAX_BOUNCE_SLOTS = 256;  // counters per kind (bounces + 1 <= this)
constexpr int EXT_STACK = 16;

struct Frame {
  double o[3], r[3], u[3], f[3];
  double tanHalfFov, aspect, lensRadius, focusDistance;
};

struct PathSoA {
  double *ox, *oy, *oz, *dx, *dy, *dz;
  double *br, *bg, *bb;
  double *Lr, *Lg, *Lb;
  double *prevPdf, *hitT;
  uint32_t* rng;
  int32_t* hitPrim;
  uint8_t* specular;
  uint32_t *q0, *q1;
  // shadow queue (one slot per live path at most)
  uint32_t* sPath;
  double *sdx, *sdy, *sdz, *sTmax, *sCr, *sCg, *sCb;
  // counters: live[b] at [b], shadow[b] at [MAX_BOUNCE_SLOTS + b], ray-fetch
  // cursors of the persistent extend/shadow launches at [2*MAX + b] and
  // [3*MAX + b]; 64-bit work counters at ctr64
  uint32_t* ctr;
  unsigned long long* ctr64;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Append `pid` to a queue for every lane with `take`; one atomic per wave.
__device__ __forceinline__ void wave_push(bool take, uint32_t pid, uint32_t* q, uint32_t* count) {
  unsigned long long active = __ballot(1);
  unsigned long long m = __ballot(take);
  if (m == 0) return;
  int lane = lane_id();
  int leader = __ffsll((long long)active) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (take) {
    unsigned long long below = m & ((1ull << lane) - 1ull);
    q[base + (uint32_t)__popcll(below)] = pid;
  }
}

__device__ __forceinline__ uint32_t wave_reserve(bool take, uint32_t* count) {
  unsigned long long active = __ballot(1);
  unsigned long long m = __ballot(take);
  int lane = lane_id();
  int leader = __ffsll((long long)active) - 1;
  uint32_t base = 0;
  if (m && lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  unsigned long long below = m & ((1ull << lane) - 1ull);
  return base + (uint32_t)__popcll(below);
}

__device__ __forceinline__ void wave_add64(unsigned long long* dst, uint32_t v) {
  unsigned long long x = v;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if ((threadIdx.x & 63) == 0 && x) atomicAdd(dst, x);
}

__global__ void k_clear(uint32_t* ctr, int n, unsigned long long* ctr64, int n64) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
  if (i < n64) ctr64[i] = 0;
}

// camera.ts:62-72 generateRay with the pixel/lens draws of tracer.ts:373-384.
__global__ void __launch_bounds__(RBLOCK)
k_generate(PathSoA P, Frame F, uint32_t width, uint32_t height, uint32_t passes,
           uint32_t frameBase) {
  const uint32_t npix = width * height;
  const size_t total = (size_t)npix * passes;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t pid = (size_t)blockIdx.x * blockDim.x + threadIdx.x; pid < total; pid += stride) {
    uint32_t pass = (uint32_t)(pid / npix);
    uint32_t pixel = (uint32_t)(pid - (size_t)pass * npix);
    uint32_t y = pixel / width, x = pixel - y * width;
    Rng rng;
    rng.s = seed_for(pixel, frameBase + pass);
    double sx = (x + rng.next()) / width;
    double sy = 1 - (y + rng.next()) / height;
    double lx = 0, ly = 0;
    if (F.lensRadius > 0) {
      double r = sqrt(rng.next());
      double theta = TWO_PI * rng.next();
      lx = r * cos(theta);
      ly = r * sin(theta);
    }
    double px = (2 * sx - 1) * F.tanHalfFov * F.aspect;
    double py = (2 * sy - 1) * F.tanHalfFov;
    double o[3], d[3];
    for (int k = 0; k < 3; k++) {
      double focus = F.o[k] + (F.r[k] * px + F.u[k] * py + F.f[k]) * F.focusDistance;
      o[k] = F.o[k] + (F.r[k] * (lx * F.lensRadius) + F.u[k] * (ly * F.lensRadius));
      d[k] = focus - o[k];
    }
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len > 0) {
      double s = 1 / len;
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
    P.q0[pid] = (uint32_t)pid;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.ctr[0] = (uint32_t)total;
}

template <int WALK, int STACK, bool COUNT>
__global__ void __launch_bounds__(RBLOCK)
k_extend(PathSoA P, DevScene S, const uint32_t* __restrict__ q, const uint32_t* __restrict__ count) {
  __shared__ int sNode[WALK == CT_WALK_ORDERED ? STACK * RBLOCK : 1];
  __shared__ float sEntry[WALK == CT_WALK_ORDERED ? STACK * RBLOCK : 1];
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < n; slot += stride) {
    uint32_t pid = q[slot];
    Hit h;
    h.t = __builtin_inf();
    h.u = h.v = 0;
    h.prim = -1;
    double ox = P.ox[pid], oy = P.oy[pid], oz = P.oz[pid];
    double dx = P.dx[pid], dy = P.dy[pid], dz = P.dz[pid];
    if (WALK == CT_WALK_ORDERED)
      closest_wave<STACK, COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, sNode + threadIdx.x,
                                 sEntry + threadIdx.x, RBLOCK, cnt);
    else
      closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
    P.hitPrim[pid] = h.prim;
    P.hitT[pid] = h.t;
  }
  if (COUNT) {
    wave_add64(P.ctr64 + 0, cnt.nodes);
    wave_add64(P.ctr64 + 1, cnt.tris);
    wave_add64(P.ctr64 + 2, cnt.spheres);
    wave_add64(P.ctr64 + 6, cnt.fallbacks);
  }
}

// lightPdfScalar (tracer.ts:61-90)
__device__ double light_pdf(const DevScene& S, double px, double py, double pz, double hx,
                            double hy, double hz, double hnx, double hny, double hnz, int prim) {
  int lightCount = S.lightCount;
  if (lightCount == 0) return 0;
  double tx = hx - px, ty = hy - py, tz = hz - pz;
  double dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return 0;
  double inv = 1 / sqrt(dist2);
  double wx = tx * inv, wy = ty * inv, wz = tz * inv;
  PrimGeom g = load_prim(S, prim);
  if (g.a.w == 0.0f) {
    double cx = (double)g.a.x - px, cy = (double)g.a.y - py, cz = (double)g.a.z - pz;
    double r = g.b.x;
    double d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      double cosMax = sqrt(max0(1 - (r * r) / d2));
      double solid = TWO_PI * (1 - cosMax);
      return solid > 0 ? 1 / (solid * lightCount) : 0;
    }
  }
  double cosL = fabs(hnx * wx + hny * wy + hnz * wz);
  if (cosL < 1e-6) return 0;
  return dist2 / ((double)g.c.w * cosL * lightCount);
}

// basisInto (materials.ts:20-42)
__device__ __forceinline__ void basis(double nx, double ny, double nz, double* b) {
  double hx = 1, hy = 0, hz = 0;
  if (fabs(nx) > 0.9) {
    hx = 0;
    hy = 1;
  }
  double tx = hy * nz - hz * ny;
  double ty = hz * nx - hx * nz;
  double tz = hx * ny - hy * nx;
  double inv = 1 / sqrt(tx * tx + ty * ty + tz * tz);
  tx *= inv;
  ty *= inv;
  tz *= inv;
  b[0] = tx; b[1] = ty; b[2] = tz;
  b[3] = ny * tz - nz * ty;
  b[4] = nz * tx - nx * tz;
  b[5] = nx * ty - ny * tx;
}

// unitVectorInto (materials.ts:67-74)
__device__ __forceinline__ void unit_vector(double u1, double u2, double* out) {
  double z = 1 - 2 * u1;
  double r = sqrt(z * z < 1 ? 1 - z * z : 0);
  double phi = TWO_PI * u2;
  out[0] = r * cos(phi);
  out[1] = r * sin(phi);
  out[2] = z;
}

struct LightSample {
  double w[3], dist, pdf, e[3];
};

// sampleLightInto (tracer.ts:93-180): consumes its draws even when the sample
// turns out degenerate.
__device__ bool sample_light(const DevScene& S, double px, double py, double pz, Rng& rng,
                             LightSample& L) {
  int lightCount = S.lightCount;
  if (lightCount == 0) return false;
  int prim = S.lights[rng.pick(lightCount)];
  PrimGeom g = load_prim(S, prim);
  float4 em = S.mats[(size_t)S.primObject[prim] * 4 + 3];
  L.e[0] = em.x; L.e[1] = em.y; L.e[2] = em.z;
  if (g.a.w == 0.0f) {
    double cx = (double)g.a.x - px, cy = (double)g.a.y - py, cz = (double)g.a.z - pz;
    double r = g.b.x;
    double d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      double cosMax = sqrt(max0(1 - (r * r) / d2));
      double cosT = 1 - rng.next() * (1 - cosMax);
      double sinT = sqrt(max0(1 - cosT * cosT));
      double phi = TWO_PI * rng.next();
      double invD = 1 / sqrt(d2);
      double ax = cx * invD, ay = cy * invD, az = cz * invD;
      double b[6];
      basis(ax, ay, az, b);
      double sx = sinT * cos(phi), sy = sinT * sin(phi);
      double wx = b[0] * sx + b[3] * sy + ax * cosT;
      double wy = b[1] * sx + b[4] * sy + ay * cosT;
      double wz = b[2] * sx + b[5] * sy + az * cosT;
      double invW = 1 / sqrt(wx * wx + wy * wy + wz * wz);
      wx *= invW;
      wy *= invW;
      wz *= invW;
      double halfB = -(cx * wx + cy * wy + cz * wz);
      double disc = halfB * halfB - (d2 - r * r);
      if (disc < 0) return false;
      double dist = -halfB - sqrt(disc);
      if (dist <= 0) return false;
      L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
      L.dist = dist;
      L.pdf = 1 / (TWO_PI * (1 - cosMax) * lightCount);
      return true;
    }
    double u1 = rng.next();
    double u2 = rng.next();
    double nv[3];
    unit_vector(u1, u2, nv);
    double tx = cx + nv[0] * r, ty = cy + nv[1] * r, tz = cz + nv[2] * r;
    double dist2 = tx * tx + ty * ty + tz * tz;
    if (dist2 <= 0) return false;
    double dist = sqrt(dist2);
    double wx = tx / dist, wy = ty / dist, wz = tz / dist;
    double cosL = fabs(nv[0] * wx + nv[1] * wy + nv[2] * wz);
    if (cosL < 1e-6) return false;
    L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
    L.dist = dist;
    L.pdf = dist2 / ((double)g.c.w * cosL * lightCount);
    return true;
  }
  double u1 = rng.next();
  double u2 = rng.next();
  if (u1 + u2 > 1) {
    u1 = 1 - u1;
    u2 = 1 - u2;
  }
  double ax = g.a.x, ay = g.a.y, az = g.a.z;
  double qx = ax + ((double)g.b.x - ax) * u1 + ((double)g.c.x - ax) * u2;
  double qy = ay + ((double)g.b.y - ay) * u1 + ((double)g.c.y - ay) * u2;
  double qz = az + ((double)g.b.z - az) * u1 + ((double)g.c.z - az) * u2;
  double tx = qx - px, ty = qy - py, tz = qz - pz;
  double dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return false;
  double dist = sqrt(dist2);
  double wx = tx / dist, wy = ty / dist, wz = tz / dist;
  float4 nrm = S.prims[(size_t)prim * 4 + 3];
  double cosL = fabs((double)nrm.x * wx + (double)nrm.y * wy + (double)nrm.z * wz);
  if (cosL < 1e-6) return false;
  L.w[0] = wx; L.w[1] = wy; L.w[2] = wz;
  L.dist = dist;
  L.pdf = dist2 / ((double)g.c.w * cosL * lightCount);
  return true;
}

__device__ __forceinline__ double power_heuristic(double a, double b) {
  double a2 = a * a;
  double b2 = b * b;
  double denom = a2 + b2;
  return denom > 0 ? a2 / denom : 0;
}

// schlick / dielectricReflectance (materials.ts:77-95)
__device__ __forceinline__ double schlick(double cosTheta, double etaI, double etaT) {
  double q = (etaI - etaT) / (etaI + etaT);
  double r0 = q * q;
  double c = cosTheta < 1 ? cosTheta : 1;
  c = c > 0 ? c : 0;
  double m = 1 - c;
  return r0 + (1 - r0) * m * m * m * m * m;
}

__device__ __forceinline__ double dielectric_reflectance(double cosI, double etaI, double etaT) {
  double eta = etaI / etaT;
  double sin2T = eta * eta * (1 - cosI * cosI);
  if (sin2T >= 1) return 1;
  if (etaI > etaT) return schlick(sqrt(1 - sin2T), etaI, etaT);
  return schlick(cosI, etaI, etaT);
}

struct TraceOpts {
  int maxBounces, rr, nee;
};

// One bounce of radianceInto for every live path.
__global__ void __launch_bounds__(RBLOCK)
k_shade(PathSoA P, DevScene S, const uint32_t* __restrict__ qin,
        const uint32_t* __restrict__ countIn, uint32_t* __restrict__ qout,
        uint32_t* __restrict__ countOut, uint32_t* __restrict__ shadowCount, int bounce,
        TraceOpts opts) {
  const uint32_t n = *countIn;
  const bool useNee = opts.nee && S.lightCount > 0;
  uint32_t stride = gridDim.x * blockDim.x;
  // Every lane of a wave runs the same number of iterations so the wave-wide
  // ballots below see the whole wave.
  uint32_t iters = (n + stride - 1) / stride;
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t it = 0; it < iters; it++, slot += stride) {
    bool valid = slot < n;
    uint32_t pid = valid ? qin[slot] : 0;
    bool alive = false, shadow = false;
    double wdir[3] = {0, 0, 0}, sTmax = 0, sc[3] = {0, 0, 0};
    if (valid) {
      double ox = P.ox[pid], oy = P.oy[pid], oz = P.oz[pid];
      double dx = P.dx[pid], dy = P.dy[pid], dz = P.dz[pid];
      double br = P.br[pid], bg = P.bg[pid], bb = P.bb[pid];
      double Lr = P.Lr[pid], Lg = P.Lg[pid], Lb = P.Lb[pid];
      double prevPdf = P.prevPdf[pid];
      bool specularBounce = P.specular[pid] != 0;
      Rng rng;
      rng.s = P.rng[pid];
      int prim = P.hitPrim[pid];
      bool cont = true;
      double ndx = 0, ndy = 0, ndz = 0, px = ox, py = oy, pz = oz;
      if (prim < 0) {
        if (S.skyKind == 1) {
          double t = 0.5 * (dy + 1);
          Lr += br * (S.horizon[0] + (S.zenith[0] - S.horizon[0]) * t) * S.skyStrength;
          Lg += bg * (S.horizon[1] + (S.zenith[1] - S.horizon[1]) * t) * S.skyStrength;
          Lb += bb * (S.horizon[2] + (S.zenith[2] - S.horizon[2]) * t) * S.skyStrength;
        }
        cont = false;
      } else {
        double t = P.hitT[pid];
        px = ox + dx * t;
        py = oy + dy * t;
        pz = oz + dz * t;
        PrimGeom g = load_prim(S, prim);
        double nx, ny, nz;
        if (g.a.w == 0.0f) {
          double invR = 1 / (double)g.b.x;
          nx = (px - (double)g.a.x) * invR;
          ny = (py - (double)g.a.y) * invR;
          nz = (pz - (double)g.a.z) * invR;
        } else {
          float4 nrm = S.prims[(size_t)prim * 4 + 3];
          nx = nrm.x;
          ny = nrm.y;
          nz = nrm.z;
        }
        bool frontFace = dx * nx + dy * ny + dz * nz < 0;
        double sx = frontFace ? nx : -nx;
        double sy = frontFace ? ny : -ny;
        double sz = frontFace ? nz : -nz;
        size_t m = (size_t)S.primObject[prim] * 4;
        float4 m0 = S.mats[m], m1 = S.mats[m + 1], m3 = S.mats[m + 3];
        double er = m3.x, eg = m3.y, eb = m3.z;
        if (er > 0 || eg > 0 || eb > 0) {
          double w = 1;
          if (useNee && !specularBounce) {
            double pl = light_pdf(S, ox, oy, oz, px, py, pz, nx, ny, nz, prim);
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
          double ar = m0.x, ag = m0.y, ab = m0.z;
          if (m1.z > 0.5f) {
            double cs = m1.y;
            double sum = floor(px * cs + 0.5) + floor(py * cs + 0.5) + floor(pz * cs + 0.5);
            if (fmod(fmod(sum, 2.0) + 2.0, 2.0) != 0) {  // !checkerPrimary
              float4 m2 = S.mats[m + 2];
              ar = m2.x;
              ag = m2.y;
              ab = m2.z;
            }
          }
          LightSample ls;
          if (useNee && sample_light(S, px, py, pz, rng, ls)) {
            double pdfL = ls.pdf;
            double cosS = sx * ls.w[0] + sy * ls.w[1] + sz * ls.w[2];
            if (cosS > 0 && pdfL > 0) {
              double pdfB = cosS * INV_PI;
              double k = (cosS * INV_PI * power_heuristic(pdfL, pdfB)) / pdfL;
              shadow = true;
              wdir[0] = ls.w[0];
              wdir[1] = ls.w[1];
              wdir[2] = ls.w[2];
              sTmax = ls.dist * SHADOW_SHORTEN;
              sc[0] = br * ar * ls.e[0] * k;
              sc[1] = bg * ag * ls.e[1] * k;
              sc[2] = bb * ab * ls.e[2] * k;
            }
          }
          // cosineHemisphereInto (materials.ts:47-64)
          double u1 = rng.next();
          double u2 = rng.next();
          double r = sqrt(u1);
          double phi = TWO_PI * u2;
          double lx = r * cos(phi);
          double ly = r * sin(phi);
          double lz = sqrt(u1 < 1 ? 1 - u1 : 0);
          double b[6];
          basis(sx, sy, sz, b);
          double ddx = b[0] * lx + b[3] * ly + sx * lz;
          double ddy = b[1] * lx + b[4] * ly + sy * lz;
          double ddz = b[2] * lx + b[5] * ly + sz * lz;
          double inv = 1 / sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
          ndx = ddx * inv;
          ndy = ddy * inv;
          ndz = ddz * inv;
          prevPdf = max0(sx * ndx + sy * ndy + sz * ndz) * INV_PI;
          br *= ar;
          bg *= ag;
          bb *= ab;
          specularBounce = false;
        } else if (type == 1.0f) {  // MAT_METAL: metalInto (materials.ts:102-124)
          double u1 = rng.next();
          double u2 = rng.next();
          double u3 = rng.next();
          double rough = m1.x;
          double k = 2 * (dx * sx + dy * sy + dz * sz);
          double rx = dx - k * sx;
          double ry = dy - k * sy;
          double rz = dz - k * sz;
          if (rough > 0) {
            double uv[3];
            unit_vector(u1, u2, uv);
            double rad = cbrt(u3) * rough;
            rx += uv[0] * rad;
            ry += uv[1] * rad;
            rz += uv[2] * rad;
          }
          double inv = 1 / sqrt(rx * rx + ry * ry + rz * rz);
          ndx = rx * inv;
          ndy = ry * inv;
          ndz = rz * inv;
          if (ndx * sx + ndy * sy + ndz * sz <= 0) {
            cont = false;
          } else {
            br *= (double)m0.x;
            bg *= (double)m0.y;
            bb *= (double)m0.z;
            specularBounce = true;
          }
        } else {  // MAT_DIELECTRIC: dielectricInto (materials.ts:131-166)
          double u = rng.next();
          double ior = m1.x;
          double etaI = frontFace ? 1 : ior;
          double etaT = frontFace ? ior : 1;
          double eta = etaI / etaT;
          double cosI = -(dx * sx + dy * sy + dz * sz);
          if (cosI > 1) cosI = 1;
          double refl = dielectric_reflectance(cosI, etaI, etaT);
          double sin2T = eta * eta * (1 - cosI * cosI);
          if (u < refl || sin2T > 1) {
            double k = 2 * (dx * sx + dy * sy + dz * sz);
            ndx = dx - k * sx;
            ndy = dy - k * sy;
            ndz = dz - k * sz;
          } else {
            double cosT = sqrt(1 - sin2T);
            double s2 = eta * cosI - cosT;
            double tx = dx * eta + sx * s2;
            double ty = dy * eta + sy * s2;
            double tz = dz * eta + sz * s2;
            double inv = 1 / sqrt(tx * tx + ty * ty + tz * tz);
            ndx = tx * inv;
            ndy = ty * inv;
            ndz = tz * inv;
          }
          br *= (double)m0.x;
          bg *= (double)m0.y;
          bb *= (double)m0.z;
          specularBounce = true;
        }
        if (cont && opts.rr && bounce >= 3) {  // tracer.ts:305-312
          double maxB = br > bg ? (br > bb ? br : bb) : bg > bb ? bg : bb;
          double pSurvive = maxB > 1 ? 1 : maxB < 0.05 ? 0.05 : maxB;
          if (rng.next() > pSurvive) {
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
      // the shadow ray starts at the hit point, which is also the next origin
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

// Any-hit of the NEE ray; adds the pending term when unoccluded (tracer.ts:268-282).
template <int WALK, int STACK, bool COUNT>
__global__ void __launch_bounds__(RBLOCK)
k_shadow(PathSoA P, DevScene S, const uint32_t* __restrict__ count) {
  __shared__ int sNode[WALK == CT_WALK_ORDERED ? STACK * RBLOCK : 1];
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < n; slot += stride) {
    uint32_t pid = P.sPath[slot];
    double ox = P.ox[pid], oy = P.oy[pid], oz = P.oz[pid];
    bool blocked;
    if (WALK == CT_WALK_ORDERED)
      blocked = occluded_wave<STACK, COUNT>(S, ox, oy, oz, P.sdx[slot], P.sdy[slot],
                                            P.sdz[slot], T_MIN, P.sTmax[slot],
                                            sNode + threadIdx.x, RBLOCK, cnt);
    else
      blocked = occluded<COUNT>(S, ox, oy, oz, P.sdx[slot], P.sdy[slot], P.sdz[slot], T_MIN,
                                P.sTmax[slot], cnt);
    if (!blocked) {
      P.Lr[pid] += P.sCr[slot];
      P.Lg[pid] += P.sCg[slot];
      P.Lb[pid] += P.sCb[slot];
    }
  }
  if (COUNT) {
    wave_add64(P.ctr64 + 3, cnt.nodes);
    wave_add64(P.ctr64 + 4, cnt.tris);
    wave_add64(P.ctr64 + 5, cnt.spheres);
    wave_add64(P.ctr64 + 7, cnt.fallbacks);
  }
}


// Persistent closest-hit: each lane pulls the next ray from the bounce's queue
// as soon as its current ray is finished (wave-aggregated fetch), so lanes do
// not idle until the slowest ray of their wave ends.  The walk is the
// wave-organised ordered walk of closest_wave (leaf parking, bf16 entries),
// over the reference's binary nodes or the derived 4-wide nodes (WIDE).
template <bool WIDE, int STACK, bool COUNT>
__global__ void __launch_bounds__(RBLOCK)
k_extend_persist(PathSoA P, DevScene S, const uint32_t* __restrict__ q,
                 const uint32_t* __restrict__ count, uint32_t* __restrict__ fetch) {
  __shared__ int sNodeAll[STACK * RBLOCK];
  __shared__ uint16_t sEntryAll[STACK * RBLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  uint16_t* sEntry = sEntryAll + threadIdx.x;
  const int lanes = RBLOCK;
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  bool have = false, exhausted = false;
  uint32_t pid = 0;
  double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0;
  double oix = 0, oiy = 0, oiz = 0;
  Hit h;
  h.t = 0; h.u = 0; h.v = 0; h.prim = -1;
  int cur = -1, sp = 0, pend = -1;
  const int lane = __lane_id();
  for (;;) {
    if (!exhausted) {
      bool need = !have;
      unsigned long long m = __ballot(need);
      if (m) {
        int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(fetch, (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        if (base + (uint32_t)__popcll(m) >= n) exhausted = true;
        if (need) {
          uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
          if (slot < n) {
            pid = q[slot];
            ox = P.ox[pid]; oy = P.oy[pid]; oz = P.oz[pid];
            dx = P.dx[pid]; dy = P.dy[pid]; dz = P.dz[pid];
            ix = safe_inv(dx); iy = safe_inv(dy); iz = safe_inv(dz);
            oix = ox * ix; oiy = oy * iy; oiz = oz * iz;
            h.t = __builtin_inf(); h.u = 0; h.v = 0; h.prim = -1;
            sp = 0; pend = -1;
            cur = walk_root<WIDE>(S, ix, iy, iz, oix, oiy, oiz, h.t, cnt, COUNT);
            have = true;
          }
        }
      }
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false;
      while (cur != -1) {
        if (cur <= -2) {
          if (pend >= 0) break;
          pend = -2 - cur;
          cur = -1;
        } else if (!interior_step<WIDE, true, STACK, COUNT>(S, ix, iy, iz, oix, oiy, oiz, h.t,
                                                            cur, sp, sNode, sEntry, lanes, cnt)) {
          overflow = true;
          cur = -1;
          break;
        }
        if (cur == -1) {
          while (sp > 0) {
            sp--;
            if (entry_from_bf16(sEntry[sp * lanes]) < h.t) {
              cur = sNode[sp * lanes];
              break;
            }
          }
        }
        if (__all(pend >= 0 || cur == -1)) break;
      }
      if (pend >= 0) {
        test_leaf<COUNT>(S, (float)pend, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
        pend = -1;
      }
      if (overflow) {
        if (COUNT) cnt.fallbacks++;
        closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
        cur = -1;
        sp = 0;
      }
      if (cur == -1 && pend < 0) {
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

// Persistent any-hit of the NEE rays (same lane refill and walk organisation).
template <bool WIDE, int STACK, bool COUNT>
__global__ void __launch_bounds__(RBLOCK)
k_shadow_persist(PathSoA P, DevScene S, const uint32_t* __restrict__ count,
                 uint32_t* __restrict__ fetch) {
  __shared__ int sNodeAll[STACK * RBLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  const int lanes = RBLOCK;
  const uint32_t n = *count;
  TraversalCounts cnt = {0, 0, 0, 0};
  bool have = false, exhausted = false;
  uint32_t slotCur = 0, pid = 0;
  double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0, tMax = 0;
  double oix = 0, oiy = 0, oiz = 0;
  int cur = -1, sp = 0, pend = -1;
  const int lane = __lane_id();
  for (;;) {
    if (!exhausted) {
      bool need = !have;
      unsigned long long m = __ballot(need);
      if (m) {
        int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(fetch, (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        if (base + (uint32_t)__popcll(m) >= n) exhausted = true;
        if (need) {
          uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
          if (slot < n) {
            slotCur = slot;
            pid = P.sPath[slot];
            ox = P.ox[pid]; oy = P.oy[pid]; oz = P.oz[pid];
            dx = P.sdx[slot]; dy = P.sdy[slot]; dz = P.sdz[slot];
            tMax = P.sTmax[slot];
            ix = safe_inv(dx); iy = safe_inv(dy); iz = safe_inv(dz);
            oix = ox * ix; oiy = oy * iy; oiz = oz * iz;
            sp = 0; pend = -1;
            cur = walk_root<WIDE>(S, ix, iy, iz, oix, oiy, oiz, tMax, cnt, COUNT);
            have = true;
          }
        }
      }
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false, blocked = false;
      while (cur != -1) {
        if (cur <= -2) {
          if (pend >= 0) break;
          pend = -2 - cur;
          cur = -1;
        } else if (!interior_step<WIDE, false, STACK, COUNT>(S, ix, iy, iz, oix, oiy, oiz, tMax,
                                                             cur, sp, sNode, nullptr, lanes,
                                                             cnt)) {
          overflow = true;
          cur = -1;
          break;
        }
        if (cur == -1 && sp > 0) {
          sp--;
          cur = sNode[sp * lanes];
        }
        if (__all(pend >= 0 || cur == -1)) break;
      }
      if (pend >= 0) {
        int count = pend & 7, start = pend >> 3;
        for (int p = start; p < start + count && !blocked; p++) {
          PrimGeom g = load_prim(S, p);
          if (g.a.w == 0.0f) {
            if (COUNT) cnt.spheres++;
            blocked = ray_sphere(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, T_MIN, tMax) > 0;
          } else {
            if (COUNT) cnt.tris++;
            double t, u, v;
            blocked = ray_triangle(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y,
                                   g.b.z, g.c.x, g.c.y, g.c.z, T_MIN, tMax, t, u, v);
          }
        }
        pend = -1;
      }
      if (overflow && !blocked) {
        if (COUNT) cnt.fallbacks++;
        blocked = occluded<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, tMax, cnt);
      }
      if (blocked || overflow) {
        cur = -1;
        sp = 0;
      }
      if (cur == -1 && pend < 0) {
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

template <int WALK, int STACK>
static void launch_extend(bool counting, unsigned grid, hipStream_t st, const PathSoA& P,
                          const DevScene& S, const uint32_t* q, const uint32_t* c,
                          uint32_t* fetch = nullptr) {
  if (WALK != CT_WALK_STACKLESS && fetch) {
    constexpr bool W = WALK == CT_WALK_WIDE;
    if (counting)
      hipLaunchKernelGGL((k_extend_persist<W, STACK, true>), dim3(grid), dim3(RBLOCK), 0, st, P, S, q, c, fetch);
    else
      hipLaunchKernelGGL((k_extend_persist<W, STACK, false>), dim3(grid), dim3(RBLOCK), 0, st, P, S, q, c, fetch);
    return;
  }
  if constexpr (WALK != CT_WALK_WIDE) {
    if (counting)
      hipLaunchKernelGGL((k_extend<WALK, STACK, true>), dim3(grid), dim3(RBLOCK), 0, st, P, S, q, c);
    else
      hipLaunchKernelGGL((k_extend<WALK, STACK, false>), dim3(grid), dim3(RBLOCK), 0, st, P, S, q, c);
  }
}

template <int WALK, int STACK>
static void launch_shadow(bool counting, unsigned grid, hipStream_t st, const PathSoA& P,
                          const DevScene& S, const uint32_t* c, uint32_t* fetch = nullptr) {
  if (WALK != CT_WALK_STACKLESS && fetch) {
    constexpr bool W = WALK == CT_WALK_WIDE;
    if (counting)
      hipLaunchKernelGGL((k_shadow_persist<W, STACK, true>), dim3(grid), dim3(RBLOCK), 0, st, P, S, c, fetch);
    else
      hipLaunchKernelGGL((k_shadow_persist<W, STACK, false>), dim3(grid), dim3(RBLOCK), 0, st, P, S, c, fetch);
    return;
  }
  if constexpr (WALK != CT_WALK_WIDE) {
    if (counting)
      hipLaunchKernelGGL((k_shadow<WALK, STACK, true>), dim3(grid), dim3(RBLOCK), 0, st, P, S, c);
    else
      hipLaunchKernelGGL((k_shadow<WALK, STACK, false>), dim3(grid), dim3(RBLOCK), 0, st, P, S, c);
  }
}

// Stack depths the ordered walks are instantiated for (LDS per 256-lane block:
// 6 B/entry for closest hit, 4 B/entry for any hit).
static bool stack_supported(int d) { return d == 8 || d == 12 || d == 16 || d == 24; }

template <int WALK>
static void dispatch_extend_w(int stack, bool counting, unsigned grid, hipStream_t st,
                              const PathSoA& P, const DevScene& S, const uint32_t* q,
                              const uint32_t* c, uint32_t* f) {
  switch (stack) {
    case 8: launch_extend<WALK, 8>(counting, grid, st, P, S, q, c, f); break;
    case 12: launch_extend<WALK, 12>(counting, grid, st, P, S, q, c, f); break;
    case 16: launch_extend<WALK, 16>(counting, grid, st, P, S, q, c, f); break;
    default: launch_extend<WALK, 24>(counting, grid, st, P, S, q, c, f); break;
  }
}

template <int WALK>
static void dispatch_shadow_w(int stack, bool counting, unsigned grid, hipStream_t st,
                              const PathSoA& P, const DevScene& S, const uint32_t* c,
                              uint32_t* f) {
  switch (stack) {
    case 8: launch_shadow<WALK, 8>(counting, grid, st, P, S, c, f); break;
    case 12: launch_shadow<WALK, 12>(counting, grid, st, P, S, c, f); break;
    case 16: launch_shadow<WALK, 16>(counting, grid, st, P, S, c, f); break;
    default: launch_shadow<WALK, 24>(counting, grid, st, P, S, c, f); break;
  }
}

// The wide walk always runs persistent (f must be the launch's fetch cursor).
static void dispatch_extend(int walk, int stack, bool counting, unsigned grid, hipStream_t st,
                            const PathSoA& P, const DevScene& S, const uint32_t* q,
                            const uint32_t* c, uint32_t* f) {
  if (walk == CT_WALK_STACKLESS)
    launch_extend<CT_WALK_STACKLESS, 1>(counting, grid, st, P, S, q, c);
  else if (walk == CT_WALK_WIDE)
    dispatch_extend_w<CT_WALK_WIDE>(stack, counting, grid, st, P, S, q, c, f);
  else
    dispatch_extend_w<CT_WALK_ORDERED>(stack, counting, grid, st, P, S, q, c, f);
}

static void dispatch_shadow(int walk, int stack, bool counting, unsigned grid, hipStream_t st,
                            const PathSoA& P, const DevScene& S, const uint32_t* c, uint32_t* f) {
  if (walk == CT_WALK_STACKLESS)
    launch_shadow<CT_WALK_STACKLESS, 1>(counting, grid, st, P, S, c);
  else if (walk == CT_WALK_WIDE)
    dispatch_shadow_w<CT_WALK_WIDE>(stack, counting, grid, st, P, S, c, f);
  else
    dispatch_shadow_w<CT_WALK_ORDERED>(stack, counting, grid, st, P, S, c, f);
}

// Adds each pass's sample in pass order, dropping non-finite samples
// (tracer.ts:386-392); optional final scale (tracer.ts:352-353).
__global__ void __launch_bounds__(RBLOCK)
k_finalize(PathSoA P, uint32_t npix, uint32_t passes, float* __restrict__ accum, int doScale,
           double scale) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < npix; pix += stride) {
    float a0 = accum[3 * (size_t)pix], a1 = accum[3 * (size_t)pix + 1],
          a2 = accum[3 * (size_t)pix + 2];
    for (uint32_t s = 0; s < passes; s++) {
      size_t pid = (size_t)s * npix + pix;
      double cr = P.Lr[pid], cg = P.Lg[pid], cb = P.Lb[pid];
      if (isfinite(cr) && isfinite(cg) && isfinite(cb)) {
        a0 = (float)((double)a0 + cr);
        a1 = (float)((double)a1 + cg);
        a2 = (float)((double)a2 + cb);
      }
    }
    if (doScale) {
      a0 = (float)((double)a0 * scale);
      a1 = (float)((double)a1 * scale);
      a2 = (float)((double)a2 * scale);
    }
    accum[3 * (size_t)pix] = a0;
    accum[3 * (size_t)pix + 1] = a1;
    accum[3 * (size_t)pix + 2] = a2;
  }
}

__global__ void k_zero(float* p, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

}  // namespace ct

using namespace ct;

struct ct_tracer {
  void* buf = nullptr;
  size_t capPaths = 0;
  PathSoA P{};
  bool profiling = false;
  bool counting = false;
  int walk = CT_WALK_WIDE;           // closest-hit walk of extension rays
  int shadowWalk = CT_WALK_WIDE;     // any-hit walk of NEE rays
  int stack = 24;                    // LDS stack entries of the ordered walks
  bool persistent = true;            // lanes refill with new rays as they finish
  unsigned persistBlocks = 256 * 8;  // workgroups of a persistent launch
  // stats of the last run
  int lastBounces = 0;
  uint64_t lastPaths = 0;
  uint64_t lastExtension = 0;
  uint64_t lastShadow = 0;
  std::vector<hipEvent_t> events;  // pairs per launch when profiling
  std::vector<int> eventKind;      // 0 extend, 1 shade, 2 shadow
  hipStream_t lastStream = nullptr;
};

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

static int tracer_reserve(ct_tracer* T, size_t paths) {
  if (paths <= T->capPaths) return CT_OK;
  if (T->buf) {
    CT_HIP(hipFree(T->buf));
    T->buf = nullptr;
    T->capPaths = 0;
  }
  size_t n = paths;
  size_t d = al(n * 8), u32 = al(n * 4), u8 = al(n);
  size_t total = 21 * d + 5 * u32 + u8 + al(4 * MAX_BOUNCE_SLOTS * 4) + al(16 * 8);
  CT_HIP(hipMalloc(&T->buf, total));
  char* p = (char*)T->buf;
  auto takeD = [&]() { double* r = (double*)p; p += d; return r; };
  auto takeU = [&]() { uint32_t* r = (uint32_t*)p; p += u32; return r; };
  PathSoA& P = T->P;
  P.ox = takeD(); P.oy = takeD(); P.oz = takeD();
  P.dx = takeD(); P.dy = takeD(); P.dz = takeD();
  P.br = takeD(); P.bg = takeD(); P.bb = takeD();
  P.Lr = takeD(); P.Lg = takeD(); P.Lb = takeD();
  P.prevPdf = takeD(); P.hitT = takeD();
  P.sdx = takeD(); P.sdy = takeD(); P.sdz = takeD(); P.sTmax = takeD();
  P.sCr = takeD(); P.sCg = takeD(); P.sCb = takeD();
  P.rng = takeU();
  P.hitPrim = (int32_t*)takeU();
  P.q0 = takeU(); P.q1 = takeU(); P.sPath = takeU();
  P.specular = (uint8_t*)p; p += u8;
  P.ctr = (uint32_t*)p; p += al(4 * MAX_BOUNCE_SLOTS * 4);
  P.ctr64 = (unsigned long long*)p;
  T->capPaths = paths;
  return CT_OK;
}

static Frame make_frame(const double* f) {
  Frame F;
  for (int k = 0; k < 3; k++) {
    F.o[k] = f[k];
    F.r[k] = f[3 + k];
    F.u[k] = f[6 + k];
    F.f[k] = f[9 + k];
  }
  F.tanHalfFov = f[12];
  F.aspect = f[13];
  F.lensRadius = f[14];
  F.focusDistance = f[15];
  return F;
}

static void clear_events(ct_tracer* T) {
  for (hipEvent_t e : T->events) hipEventDestroy(e);
  T->events.clear();
  T->eventKind.clear();
}

static void mark(ct_tracer* T, hipStream_t st, int kind, bool begin) {
  if (!T->profiling) return;
  hipEvent_t e;
  hipEventCreate(&e);
  hipEventRecord(e, st);
  T->events.push_back(e);
  if (begin) T->eventKind.push_back(kind);
}

// Trace `passes` passes of width x height (frame indices frameBase..+passes-1)
// and add them into accum (optionally scaled at the end).
static int run_passes(ct_tracer* T, const ct_scene* s, const Frame& F, uint32_t width,
                      uint32_t height, uint32_t passes, uint32_t frameBase, TraceOpts opts,
                      float* accum, bool doScale, double scale, hipStream_t st, bool firstBatch) {
  size_t npix = (size_t)width * height;
  size_t paths = npix * passes;
  int rc = tracer_reserve(T, paths);
  if (rc) return rc;
  PathSoA P = T->P;
  int nb = opts.maxBounces + 1;
  // live/shadow queue counts restart per batch; the 64-bit work counters keep
  // summing over the batches of one render
  hipLaunchKernelGGL(k_clear, dim3(4), dim3(256), 0, st, P.ctr, 4 * MAX_BOUNCE_SLOTS, P.ctr64,
                     firstBatch ? 16 : 0);
  CT_LAUNCH_CHECK("k_clear");
  unsigned gridPaths = grid_for(paths, RBLOCK, 256 * 16);
  hipLaunchKernelGGL(k_generate, dim3(gridPaths), dim3(RBLOCK), 0, st, P, F, width, height,
                     passes, frameBase);
  CT_LAUNCH_CHECK("k_generate");
  for (int b = 0; b < nb; b++) {
    uint32_t* qin = (b & 1) ? P.q1 : P.q0;
    uint32_t* qout = (b & 1) ? P.q0 : P.q1;
    uint32_t* cin = P.ctr + b;
    uint32_t* cout = P.ctr + b + 1;
    uint32_t* scount = P.ctr + MAX_BOUNCE_SLOTS + b;
    mark(T, st, 0, true);
    bool persist = T->persistent || T->walk == CT_WALK_WIDE || T->shadowWalk == CT_WALK_WIDE;
    uint32_t* fExt = persist ? P.ctr + 2 * MAX_BOUNCE_SLOTS + b : nullptr;
    uint32_t* fSh = persist ? P.ctr + 3 * MAX_BOUNCE_SLOTS + b : nullptr;
    unsigned gridTrace = persist ? grid_for(paths, RBLOCK, T->persistBlocks) : gridPaths;
    dispatch_extend(T->walk, T->stack, T->counting, gridTrace, st, P, s->dev, qin, cin, fExt);
    CT_LAUNCH_CHECK("k_extend");
    mark(T, st, 0, false);
    mark(T, st, 1, true);
    hipLaunchKernelGGL(k_shade, dim3(gridPaths), dim3(RBLOCK), 0, st, P, s->dev, qin, cin, qout,
                       cout, scount, b, opts);
    CT_LAUNCH_CHECK("k_shade");
    mark(T, st, 1, false);
    mark(T, st, 2, true);
    dispatch_shadow(T->shadowWalk, T->stack, T->counting, gridTrace, st, P, s->dev, scount, fSh);
    CT_LAUNCH_CHECK("k_shadow");
    mark(T, st, 2, false);
  }
  unsigned gridPix = grid_for(npix, RBLOCK, 256 * 16);
  hipLaunchKernelGGL(k_finalize, dim3(gridPix), dim3(RBLOCK), 0, st, P, (uint32_t)npix, passes,
                     accum, doScale ? 1 : 0, scale);
  CT_LAUNCH_CHECK("k_finalize");
  return CT_OK;
}

// Largest number of passes per batch so path state stays within a fixed budget.
static uint32_t passes_per_batch(size_t npix, uint32_t passes) {
  const size_t maxPaths = size_t(1) << 25;  // 33.5M paths ~ 6.4 GB of path state
  size_t per = npix ? maxPaths / npix : passes;
  if (per < 1) per = 1;
  return (uint32_t)(per < passes ? per : passes);
}

struct BatchTotals {
  uint64_t ext = 0, shadow = 0;
};

static int read_batch_counts(ct_tracer* T, hipStream_t st, int nb, BatchTotals& tot) {
  std::vector<uint32_t> h(2 * MAX_BOUNCE_SLOTS);
  CT_HIP(hipMemcpyAsync(h.data(), T->P.ctr, h.size() * 4, hipMemcpyDeviceToHost, st));
  CT_HIP(hipStreamSynchronize(st));
  for (int b = 0; b < nb; b++) {
    tot.ext += h[b];
    tot.shadow += h[MAX_BOUNCE_SLOTS + b];
  }
  return CT_OK;
}

extern "C" {

int ct_tracer_create(ct_tracer** out) {
  if (!out) return set_error(CT_EINVAL, "ct_tracer_create: null out");
  *out = new ct_tracer();
  return CT_OK;
}

int ct_tracer_destroy(ct_tracer* T) {
  if (!T) return CT_OK;
  clear_events(T);
  hipError_t e = hipSuccess;
  if (T->buf) e = hipFree(T->buf);
  delete T;
  return check_hip(e, "ct_tracer_destroy");
}

int ct_tracer_set_profiling(ct_tracer* T, int32_t enable) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  T->profiling = enable != 0;
  return CT_OK;
}

int ct_tracer_set_counting(ct_tracer* T, int32_t enable) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  T->counting = enable != 0;
  return CT_OK;
}

int ct_tracer_set_traversal(ct_tracer* T, int32_t shadowWalk, int32_t stackDepth) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (shadowWalk != CT_WALK_ORDERED && shadowWalk != CT_WALK_STACKLESS &&
      shadowWalk != CT_WALK_WIDE)
    return set_error(CT_EINVAL, "ct_tracer_set_traversal: unknown walk");
  if (!stack_supported(stackDepth))
    return set_error(CT_EINVAL, "ct_tracer_set_traversal: stack depth must be 8, 12, 16 or 24");
  T->shadowWalk = shadowWalk;
  T->stack = stackDepth;
  return CT_OK;
}

int ct_tracer_set_schedule(ct_tracer* T, int32_t persistent, int32_t blocks) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (persistent && (blocks < 1 || blocks > 65535))
    return set_error(CT_EINVAL, "ct_tracer_set_schedule: blocks must be in [1, 65535]");
  T->persistent = persistent != 0;
  if (persistent) T->persistBlocks = (unsigned)blocks;
  return CT_OK;
}

int ct_tracer_set_walk(ct_tracer* T, int32_t walk) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (walk != CT_WALK_ORDERED && walk != CT_WALK_STACKLESS && walk != CT_WALK_WIDE)
    return set_error(CT_EINVAL, "ct_tracer_set_walk: unknown walk");
  T->walk = walk;
  return CT_OK;
}

static int render_common(ct_tracer* T, const ct_scene* s, const double* frame16, uint32_t width,
                         uint32_t height, uint32_t spp, uint32_t frameBase, TraceOpts opts,
                         float* out, bool average, hipStream_t st) {
  if (!T || !s || !frame16 || !out) return set_error(CT_EINVAL, "render: null argument");
  if (opts.maxBounces < 0 || opts.maxBounces + 2 > MAX_BOUNCE_SLOTS)
    return set_error(CT_ELIMIT, "render: maxBounces must be in [0, 254]");
  if ((size_t)width * height > (size_t)1 << 26)
    return set_error(CT_ELIMIT, "render: more than 2^26 pixels");
  clear_events(T);
  T->lastStream = st;
  size_t npix = (size_t)width * height;
  if (npix == 0 || spp == 0) return CT_OK;
  Frame F = make_frame(frame16);
  int nb = opts.maxBounces + 1;
  uint32_t per = passes_per_batch(npix, spp);
  BatchTotals tot;
  if (average) {
    unsigned g = grid_for(npix * 3, RBLOCK, 256 * 16);
    hipLaunchKernelGGL(k_zero, dim3(g), dim3(RBLOCK), 0, st, out, npix * 3);
    CT_LAUNCH_CHECK("k_zero");
  }
  for (uint32_t s0 = 0; s0 < spp; s0 += per) {
    uint32_t passes = (spp - s0) < per ? (spp - s0) : per;
    bool last = s0 + passes >= spp;
    int rc = run_passes(T, s, F, width, height, passes, frameBase + s0, opts, out,
                        average && last, average ? 1.0 / spp : 1.0, st, s0 == 0);
    if (rc) return rc;
    rc = read_batch_counts(T, st, nb, tot);
    if (rc) return rc;
  }
  T->lastExtension = tot.ext;
  T->lastShadow = tot.shadow;
  T->lastBounces = nb;
  T->lastPaths = (uint64_t)npix * spp;
  return CT_OK;
}

int ct_accumulate_pass(ct_tracer* T, const ct_scene* s, const double* frame16, uint32_t width,
                       uint32_t height, uint32_t frameIndex, const ct_trace_opts* opts,
                       float* accum, void* stream) {
  if (!opts) return set_error(CT_EINVAL, "ct_accumulate_pass: null opts");
  TraceOpts o{opts->maxBounces, opts->russianRoulette != 0, opts->nee != 0};
  return render_common(T, s, frame16, width, height, 1, frameIndex, o, accum, false,
                       (hipStream_t)stream);
}

int ct_render_image(ct_tracer* T, const ct_scene* s, const double* frame16,
                    const ct_render_opts* opts, float* out, void* stream) {
  if (!opts) return set_error(CT_EINVAL, "ct_render_image: null opts");
  TraceOpts o{opts->maxBounces, opts->russianRoulette != 0, opts->nee != 0};
  uint32_t base = opts->seed * 7919u;  // frameIndex = s + seed*7919 (uint32 wrap)
  return render_common(T, s, frame16, opts->width, opts->height, opts->spp, base, o, out, true,
                       (hipStream_t)stream);
}

int ct_tracer_stats(ct_tracer* T, ct_render_stats* out) {
  if (!T || !out) return set_error(CT_EINVAL, "ct_tracer_stats: null argument");
  memset(out, 0, sizeof *out);
  CT_HIP(hipStreamSynchronize(T->lastStream));
  out->extensionRays = T->lastExtension;
  out->shadowRays = T->lastShadow;
  out->paths = T->lastPaths;
  if (T->profiling) {
    // events come in (begin, end) pairs in launch order
    for (size_t i = 0; i + 1 < T->events.size(); i += 2) {
      float ms = 0;
      CT_HIP(hipEventElapsedTime(&ms, T->events[i], T->events[i + 1]));
      int kind = T->eventKind[i / 2];
      if (kind == 0) {
        out->extendMs += ms;
        out->extendLaunches++;
      } else if (kind == 1) {
        out->shadeMs += ms;
      } else {
        out->shadowMs += ms;
      }
    }
  } else {
    out->extendLaunches = (uint32_t)T->lastBounces;
  }
  if (T->counting && T->buf) {
    unsigned long long h[8];
    CT_HIP(hipMemcpy(h, T->P.ctr64, sizeof h, hipMemcpyDeviceToHost));
    out->stackFallbacks = h[6];
    out->shadowStackFallbacks = h[7];
    out->nodeVisits = h[0];
    out->triTests = h[1];
    out->sphereTests = h[2];
    out->shadowNodeVisits = h[3];
    out->shadowTriTests = h[4];
    out->shadowSphereTests = h[5];
  }
  return CT_OK;
}

}  // extern "C"
