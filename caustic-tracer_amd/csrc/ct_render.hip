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
// A launch round traces one or more jobs at once.  A job is (scene, camera,
// image size, trace options, frame indices of its passes, float32 sum); a
// path is (job, pass, pixel) and carries its job id, so rays of different
// scenes, sizes and depths share every launch (MULTI instantiation).  A
// single-job round reads its job through uniform (scalar) loads.  Each job's
// finalize adds its passes in pass order, so every float32 sum is the one
// accumulatePass produces pass after pass, whatever else shared the launches.
// The NEE term of bounce b is added by the shadow kernel before bounce b+1 is
// shaded, keeping the reference's order of additions into L.  Live paths are
// compacted with a wave64 ballot + one atomic per wave; traversal kernels are
// persistent and refill lanes from the queue as rays finish.
#include <string.h>

#include <algorithm>
#include <limits>
#include <type_traits>
#include <map>
#include <mutex>
#include <vector>

#include "ct_engine.h"
#include "ct_qwalk.h"
#include "ct_jsmath.h"
#include "ct_wavefront.h"

namespace ct {

__global__ void k_clear(uint32_t* ctr, int n, unsigned long long* ctr64, int n64) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
  if (i < n64) ctr64[i] = 0;
}

// A path's state between bounces (the locals of radianceInto, tracer.ts:199-208).
struct PathState {
  double ox, oy, oz, dx, dy, dz;
  double br, bg, bb;
  double Lr, Lg, Lb;
  double prevPdf;
  bool specular;
  Rng rng;
};

// Path `local` of job J at its start: the caller's ray and RNG state for a
// radianceInto job, else the camera ray of its pixel and pass
// (camera.ts:62-72 generateRay with the pixel/lens draws of tracer.ts:373-384).
__device__ __forceinline__ PathState start_path(const JobDev& J, size_t local) {
  const Frame& F = J.F;
  const uint32_t width = J.width, height = J.height, npix = J.npix;
  uint32_t pass = (uint32_t)(local / npix);
  uint32_t pixel = (uint32_t)(local - (size_t)pass * npix);
  PathState p;
  p.br = 1; p.bg = 1; p.bb = 1;
  p.Lr = 0; p.Lg = 0; p.Lb = 0;
  p.prevPdf = 0;
  p.specular = true;
  if (J.rays) {  // radianceInto: the caller's ray and RNG state (tracer.ts:191-198)
    const double* r = J.rays + 6 * (size_t)pixel;
    p.ox = r[0]; p.oy = r[1]; p.oz = r[2];
    p.dx = r[3]; p.dy = r[4]; p.dz = r[5];
    p.rng.s = J.rngIO[pixel];
    return p;
  }
  uint32_t y = pixel / width, x = pixel - y * width;
  Rng rng;
  rng.s = seed_for(pixel, J.frameBase + pass);
  double sx = (x + rng.next()) / width;
  double sy = 1 - (y + rng.next()) / height;
  double lx = 0, ly = 0;
  if (F.lensRadius > 0) {
    double r = sqrt(rng.next());
    double theta = TWO_PI * rng.next();
    double st, ct;
    jsm::sincos(theta, st, ct);
    lx = r * ct;
    ly = r * st;
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
  p.ox = o[0]; p.oy = o[1]; p.oz = o[2];
  p.dx = d[0]; p.dy = d[1]; p.dz = d[2];
  p.rng = rng;
  return p;
}

template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_generate(PathSoA P, const JobDev* __restrict__ jobs, int njobs, size_t base, size_t count) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const size_t pid = base + i;
    int j = job_of<MULTI, false>(jobs, njobs, pid);
    const PathState p = start_path(jobs[j], pid - jobs[j].pathBase);
    P.ox[pid] = p.ox; P.oy[pid] = p.oy; P.oz[pid] = p.oz;
    P.dx[pid] = p.dx; P.dy[pid] = p.dy; P.dz[pid] = p.dz;
    P.br[pid] = p.br; P.bg[pid] = p.bg; P.bb[pid] = p.bb;
    P.Lr[pid] = p.Lr; P.Lg[pid] = p.Lg; P.Lb[pid] = p.Lb;
    P.prevPdf[pid] = p.prevPdf;
    P.specular[pid] = p.specular ? 1 : 0;
    P.rng[pid] = p.rng.s;
    if (MULTI) P.jobOf[pid] = (uint32_t)j;
    P.q0[i] = (uint32_t)pid;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.ctr[0] = (uint32_t)count;
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
  double sp, cp;
  jsm::sincos(phi, sp, cp);
  out[0] = r * cp;
  out[1] = r * sp;
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
  int prim = gload(S.lights + rng.pick(lightCount));
  const LightGeom g = load_light(S, prim);
  float4 em = gload(S.mats + (size_t)g.obj * 4 + 3);
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
      double sp, cp;
      jsm::sincos(phi, sp, cp);
      double sx = sinT * cp, sy = sinT * sp;
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
  const float4 nrm = g.n;
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

// The NEE ray a bounce queues: direction, segment end and the term added to
// the path's radiance when nothing blocks it (tracer.ts:262-282).
struct NeeRay {
  double w[3], tMax, c[3];
};

// One bounce of radianceInto (tracer.ts:209-312) at the closest hit (prim,
// tHit) of the path's ray (prim -1: a miss): emission or sky added to p's radiance,
// the NEE ray set in `nee` (neeOn), the next direction, throughput, pdf and
// specular flag drawn into p, p's origin moved to the hit point.  Returns
// whether the path goes on (before the maxBounces test).
__device__ __forceinline__ bool shade_bounce(const DevScene& S, const TraceOpts& opts,
                                             PathState& ps, int prim, double tHit, int bounce,
                                             bool& neeOn, NeeRay& nee) {
  const bool useNee = opts.nee && S.lightCount > 0;
  double ox = ps.ox, oy = ps.oy, oz = ps.oz;
  double dx = ps.dx, dy = ps.dy, dz = ps.dz;
  double br = ps.br, bg = ps.bg, bb = ps.bb;
  double Lr = ps.Lr, Lg = ps.Lg, Lb = ps.Lb;
  double prevPdf = ps.prevPdf;
  bool specularBounce = ps.specular;
  Rng rng = ps.rng;
  neeOn = false;
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
    px = ox + dx * tHit;
    py = oy + dy * tHit;
    pz = oz + dz * tHit;
    const HitGeom g = load_hit(S, prim);
    double nx, ny, nz;
    if (g.a.w == 0.0f) {
      double invR = 1 / (double)g.b.x;
      nx = (px - (double)g.a.x) * invR;
      ny = (py - (double)g.a.y) * invR;
      nz = (pz - (double)g.a.z) * invR;
    } else {
      nx = g.n.x;
      ny = g.n.y;
      nz = g.n.z;
    }
    bool frontFace = dx * nx + dy * ny + dz * nz < 0;
    double sx = frontFace ? nx : -nx;
    double sy = frontFace ? ny : -ny;
    double sz = frontFace ? nz : -nz;
    size_t m = (size_t)g.obj * 4;
    float4 m0 = gload(S.mats + m), m1 = gload(S.mats + m + 1), m3 = gload(S.mats + m + 3);
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
        // ((sum % 2) + 2) % 2 != 0 (materials.ts checker parity): sum is an
        // integer, so sum - 2*floor(sum/2) is exactly 1 when odd, 0 when even
        // (NaN for NaN or infinities, as the remainders give)
        if (sum - 2.0 * floor(0.5 * sum) != 0) {  // !checkerPrimary
          float4 m2 = gload(S.mats + m + 2);
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
          neeOn = true;
          nee.w[0] = ls.w[0];
          nee.w[1] = ls.w[1];
          nee.w[2] = ls.w[2];
          nee.tMax = ls.dist * SHADOW_SHORTEN;
          nee.c[0] = br * ar * ls.e[0] * k;
          nee.c[1] = bg * ag * ls.e[1] * k;
          nee.c[2] = bb * ab * ls.e[2] * k;
        }
      }
      // cosineHemisphereInto (materials.ts:47-64)
      double u1 = rng.next();
      double u2 = rng.next();
      double r = sqrt(u1);
      double phi = TWO_PI * u2;
      double sp, cp;
      jsm::sincos(phi, sp, cp);
      double lx = r * cp;
      double ly = r * sp;
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
        double rad = jsm::cbrt(u3) * rough;
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
  ps.Lr = Lr;
  ps.Lg = Lg;
  ps.Lb = Lb;
  // the NEE ray starts at the hit point, which is also the next origin
  ps.ox = px;
  ps.oy = py;
  ps.oz = pz;
  ps.rng = rng;
  if (cont) {
    ps.dx = ndx;
    ps.dy = ndy;
    ps.dz = ndz;
    ps.br = br;
    ps.bg = bg;
    ps.bb = bb;
    ps.prevPdf = prevPdf;
    ps.specular = specularBounce;
  }
  return cont;
}

// One bounce of radianceInto for every live path.  (137 VGPRs, three waves
// per SIMD; a 128-VGPR cap for four waves spills and measured the same on the
// 1080p soup frame.)
template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_shade(PathSoA P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ qin,
        const uint32_t* __restrict__ countIn, uint32_t* __restrict__ qout,
        uint32_t* __restrict__ countOut, uint32_t* __restrict__ shadowCount, int bounce) {
  const uint32_t n = *countIn;
  uint32_t stride = gridDim.x * blockDim.x;
  // Every lane of a wave runs the same number of iterations so the wave-wide
  // ballots below see the whole wave.
  uint32_t iters = (n + stride - 1) / stride;
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t it = 0; it < iters; it++, slot += stride) {
    bool valid = slot < n;
    uint32_t pid = valid ? qin[slot] : 0;
    bool alive = false, shadow = false;
    NeeRay nee;
    if (valid) {
      const JobDev& J = jobs[MULTI ? P.jobOf[pid] : 0];
      PathState ps;
      ps.ox = P.ox[pid]; ps.oy = P.oy[pid]; ps.oz = P.oz[pid];
      ps.dx = P.dx[pid]; ps.dy = P.dy[pid]; ps.dz = P.dz[pid];
      ps.br = P.br[pid]; ps.bg = P.bg[pid]; ps.bb = P.bb[pid];
      ps.Lr = P.Lr[pid]; ps.Lg = P.Lg[pid]; ps.Lb = P.Lb[pid];
      ps.prevPdf = P.prevPdf[pid];
      ps.specular = P.specular[pid] != 0;
      ps.rng.s = P.rng[pid];
      const int prim = P.hitPrim[pid];
      const double t = prim >= 0 ? P.hitT[pid] : 0.0;
      const bool cont = shade_bounce(J.S, J.opts, ps, prim, t, bounce, shadow, nee);
      P.Lr[pid] = ps.Lr;
      P.Lg[pid] = ps.Lg;
      P.Lb[pid] = ps.Lb;
      P.ox[pid] = ps.ox;
      P.oy[pid] = ps.oy;
      P.oz[pid] = ps.oz;
      // kept for ended paths too: radianceInto hands the state back
      P.rng[pid] = ps.rng.s;
      if (cont && bounce < J.opts.maxBounces) {
        alive = true;
        P.dx[pid] = ps.dx;
        P.dy[pid] = ps.dy;
        P.dz[pid] = ps.dz;
        P.br[pid] = ps.br;
        P.bg[pid] = ps.bg;
        P.bb[pid] = ps.bb;
        P.prevPdf[pid] = ps.prevPdf;
        P.specular[pid] = ps.specular ? 1 : 0;
      }
    }
    wave_push(alive, pid, qout, countOut);
    uint32_t sslot = wave_reserve(shadow, shadowCount);
    if (shadow) {
      P.sPath[sslot] = pid;
      P.sdx[sslot] = nee.w[0];
      P.sdy[sslot] = nee.w[1];
      P.sdz[sslot] = nee.w[2];
      P.sTmax[sslot] = nee.tMax;
      P.sCr[sslot] = nee.c[0];
      P.sCg[sslot] = nee.c[1];
      P.sCb[sslot] = nee.c[2];
    }
  }
}

// One interior step of the ordered walks: the quantized 4-wide node (FP32
// planes with the binary64 ray's slack) or the binary node (binary64 slabs).
template <bool WIDE, bool PUSH_ENTRY, int SD, bool COUNT, int TOP, typename EntryT>
__device__ __forceinline__ bool walk_step(const DevScene& S, const qw::QRay& Q, double ix,
                                          double iy, double iz, double oix, double oiy, double oiz,
                                          double tMax, int& cur, int& sp, int* __restrict__ sNode,
                                          EntryT* __restrict__ sEntry, int lanes,
                                          TraversalCounts& cnt, const float4* __restrict__ top) {
  if constexpr (WIDE)
    return qw::qwide_step<PUSH_ENTRY, SD, COUNT, TOP>(S, Q, qw::cull_bound(tMax), cur, sp, sNode,
                                                      sEntry, lanes, cnt, top);
  else
    return interior_step<PUSH_ENTRY, SD, COUNT>(S, ix, iy, iz, oix, oiy, oiz, tMax, cur, sp,
                                                sNode, sEntry, lanes, cnt);
}

// Pops deferred work whose entry bound is not beyond the closest hit.
template <bool WIDE, typename EntryT>
__device__ __forceinline__ void walk_pop(int& cur, int& sp, const int* __restrict__ sNode,
                                         const EntryT* __restrict__ sEntry, int lanes,
                                         double closest) {
  if constexpr (WIDE) {
    qw::qpop(cur, sp, sNode, sEntry, lanes, qw::cull_bound(closest));
  } else {
    while (sp > 0) {
      sp--;
      if (entry_from_bf16(sEntry[sp * lanes]) < closest) {
        cur = sNode[sp * lanes];
        return;
      }
    }
  }
}

// Persistent closest hit.  WALK: CT_WALK_STACKLESS (the reference's skip-link
// loop, exact mirror), CT_WALK_ORDERED (binary nodes, binary64 slab tests) or
// CT_WALK_WIDE (the quantized 4-wide nodes walked in FP32 with the binary64
// ray's rounding folded into the plane slack); the ordered forms park leaves
// (up to PARK_SLOTS64) until every lane of the wave holds PARK_LEAVES of them
// and keep one-byte or bf16 lower bounds of deferred entries.  Primitive tests are binary64 in every walk, so
// the closest hit is the reference's.
// The 4-wide walk's LDS stack (4-byte code + 1-byte entry per slot, 30 KB per
// 256-lane workgroup) leaves room for five waves per SIMD; the register cap
// lets the VGPR budget match.
template <int WALK, int STACK, bool COUNT, bool MULTI, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    __attribute__((amdgpu_waves_per_eu(WALK == CT_WALK_WIDE ? 5 : 1)))
    k_extend(PathSoA P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ q,
         const uint32_t* __restrict__ count, uint32_t* __restrict__ fetch) {
  constexpr bool WIDE = WALK == CT_WALK_WIDE;
  constexpr int SD = WALK == CT_WALK_STACKLESS ? 1 : STACK;
  using EntryT = typename std::conditional<WIDE, qw::Entry, uint16_t>::type;
  // single-scene launches read the top levels of the 4-wide tree from LDS
  constexpr int TOP = (WIDE && !MULTI) ? TOP_NODES : 0;
  __shared__ int sNodeAll[SD * BLOCK];
  __shared__ EntryT sEntryAll[SD * BLOCK];
  int* sNode = sNodeAll + threadIdx.x;
  EntryT* sEntry = sEntryAll + threadIdx.x;
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
  double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0;
  double oix = 0, oiy = 0, oiz = 0;
  qw::QRay Q = qw::qray(0, 0, 0, 1, 1, 1);
  Hit h;
  h.t = 0; h.u = 0; h.v = 0; h.prim = -1; h.tie = false;
  int cur = -1, sp = 0;
  Parked<PARK_SLOTS64> pk;
  pk.n = 0;
  for (;;) {
    if (!exhausted) {
      uint32_t slot = refill(!have, n, fetch, exhausted);
      if (slot != 0xffffffffu) {
        pid = q[slot];
        if (MULTI) S = jobs[P.jobOf[pid]].S;
        ox = P.ox[pid]; oy = P.oy[pid]; oz = P.oz[pid];
        dx = P.dx[pid]; dy = P.dy[pid]; dz = P.dz[pid];
        h.t = __builtin_inf(); h.u = 0; h.v = 0; h.prim = -1; h.tie = false;
        have = true;
        if (WALK == CT_WALK_STACKLESS) {
          closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
          cur = -1;
        } else if (WIDE) {
          Q = qw::qray64(S, ox, oy, oz, dx, dy, dz);
          sp = 0;
          cur = 0;
        } else {
          ix = safe_inv(dx); iy = safe_inv(dy); iz = safe_inv(dz);
          oix = ox * ix; oiy = oy * iy; oiz = oz * iz;
          sp = 0;
          cur = walk_root(S, ix, iy, iz, oix, oiy, oiz, h.t, cnt, COUNT);
        }
      }
    }
    if (!__any(have)) break;
    if (have) {
      bool overflow = false;
      if (WALK != CT_WALK_STACKLESS) {
        while (cur != -1) {
          if (cur <= -2) {
            if (pk.n == PARK_SLOTS64) break;
            park(pk, -2 - cur);
            cur = -1;
          } else if (!walk_step<WIDE, true, SD, COUNT, TOP>(S, Q, ix, iy, iz, oix, oiy, oiz, h.t,
                                                            cur, sp, sNode, sEntry, lanes, cnt,
                                                            top)) {
            overflow = true;
            cur = -1;
            break;
          }
          if (cur == -1) walk_pop<WIDE>(cur, sp, sNode, sEntry, lanes, h.t);
          if (__all(pk.n >= PARK_LEAVES || cur == -1)) break;
        }
#pragma unroll
        for (int k = 0; k < PARK_SLOTS64; k++)  // constant slots: the words stay in registers
          if (k < pk.n) test_leaf<COUNT>(S, pk.w[k], ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
        pk.n = 0;
        if (overflow) {
          if (COUNT) cnt.fallbacks++;
          closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
          cur = -1;
          sp = 0;
        }
      }
      if (cur == -1) {
        settle_ties<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, __builtin_inf(), h, cnt);
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

template <bool UNIFORM = false>
__device__ __forceinline__ bool leaf_blocks(const DevScene& S, int leafWord, double ox, double oy,
                                            double oz, double dx, double dy, double dz,
                                            double tMax, TraversalCounts& cnt, bool count) {
  int c = leafWord & 7, start = leafWord >> 3;
  for (int p = start; p < start + c; p++) {
    PrimGeom g = load_prim<UNIFORM>(S, p);
    if (g.a.w == 0.0f) {
      if (count) cnt.spheres++;
      if (ray_sphere(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, T_MIN, tMax) > 0)
        return true;
    } else {
      if (count) cnt.tris++;
      double t, u, v;
      if (ray_triangle(ox, oy, oz, dx, dy, dz, g.a.x, g.a.y, g.a.z, g.b.x, g.b.y, g.b.z, g.c.x,
                       g.c.y, g.c.z, T_MIN, tMax, t, u, v))
        return true;
    }
  }
  return false;
}

// Persistent any hit of the NEE rays; adds the pending term when unoccluded
// (tracer.ts:268-282).  Whether (tMin, tMax) is blocked does not depend on
// the visiting order, so every walk returns the reference's answer.
template <int WALK, int STACK, bool COUNT, bool MULTI, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_shadow(PathSoA P, const JobDev* __restrict__ jobs, const uint32_t* __restrict__ count,
         uint32_t* __restrict__ fetch) {
  constexpr bool WIDE = WALK == CT_WALK_WIDE;
  constexpr int SD = WALK == CT_WALK_STACKLESS ? 1 : STACK;
  using EntryT = typename std::conditional<WIDE, qw::Entry, uint16_t>::type;
  constexpr int TOP = (WIDE && !MULTI) ? TOP_NODES : 0;
  __shared__ int sNodeAll[SD * BLOCK];
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
  double ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0, tMax = 0;
  double oix = 0, oiy = 0, oiz = 0;
  qw::QRay Q = qw::qray(0, 0, 0, 1, 1, 1);
  int cur = -1, sp = 0;
  Parked<SHADOW_PARK_SLOTS> pk;
  pk.n = 0;
  for (;;) {
    if (!exhausted) {
      uint32_t slot = refill(!have, n, fetch, exhausted);
      if (slot != 0xffffffffu) {
        slotCur = slot;
        pid = P.sPath[slot];
        if (MULTI) S = jobs[P.jobOf[pid]].S;
        ox = P.ox[pid]; oy = P.oy[pid]; oz = P.oz[pid];
        dx = P.sdx[slot]; dy = P.sdy[slot]; dz = P.sdz[slot];
        tMax = P.sTmax[slot];
        have = true;
        blocked = false;
        if (WALK == CT_WALK_STACKLESS) {
          blocked = occluded<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, tMax, cnt);
          cur = -1;
        } else if (WIDE) {
          Q = qw::qray64(S, ox, oy, oz, dx, dy, dz);
          sp = 0;
          cur = 0;
        } else {
          ix = safe_inv(dx); iy = safe_inv(dy); iz = safe_inv(dz);
          oix = ox * ix; oiy = oy * iy; oiz = oz * iz;
          sp = 0;
          cur = walk_root(S, ix, iy, iz, oix, oiy, oiz, tMax, cnt, COUNT);
        }
      }
    }
    if (!__any(have)) break;
    if (have) {
      if (WALK != CT_WALK_STACKLESS) {
        bool overflow = false;
        while (cur != -1) {
          if (cur <= -2) {
            if (pk.n == SHADOW_PARK_SLOTS) break;
            park(pk, -2 - cur);
            cur = -1;
          } else if (!walk_step<WIDE, false, SD, COUNT, TOP>(S, Q, ix, iy, iz, oix, oiy, oiz,
                                                             tMax, cur, sp, sNode,
                                                             (EntryT*)nullptr, lanes, cnt, top)) {
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
        for (int k = 0; k < SHADOW_PARK_SLOTS; k++)
          if (k < pk.n && !blocked)
            blocked = leaf_blocks(S, pk.w[k], ox, oy, oz, dx, dy, dz, tMax, cnt, COUNT);
        pk.n = 0;
        if (overflow && !blocked) {
          if (COUNT) cnt.fallbacks++;
          blocked = occluded<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, tMax, cnt);
        }
        if (blocked || overflow) {
          cur = -1;
          sp = 0;
        }
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

// One lane's closest hit over the quantized 4-wide nodes (ct_qwalk.h), each
// leaf tested when the walk reaches it; a stack overflow continues with the
// skip-link walk under the current bound, which is exact.  Exact-t ties are
// marked (TIES_DETECT) for settle_ties.
template <int STACK, bool COUNT>
__device__ __forceinline__ void closest_wide_lane(const DevScene& S, double ox, double oy,
                                                  double oz, double dx, double dy, double dz,
                                                  Hit& h, int* __restrict__ sNode,
                                                  qw::Entry* __restrict__ sEntry, int lanes,
                                                  TraversalCounts& cnt) {
  const qw::QRay Q = qw::qray64(S, ox, oy, oz, dx, dy, dz);
  int cur = 0, sp = 0;
  while (cur != -1) {
    if (cur <= -2) {
      test_leaf<COUNT>(S, -2 - cur, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
      cur = -1;
    } else if (!qw::qwide_step<true, STACK, COUNT>(S, Q, qw::cull_bound(h.t), cur, sp, sNode,
                                                   sEntry, lanes, cnt)) {
      if (COUNT) cnt.fallbacks++;
      closest_stackless<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, h, cnt);
      return;
    }
    if (cur == -1) qw::qpop(cur, sp, sNode, sEntry, lanes, qw::cull_bound(h.t));
  }
}

// One lane's any hit on (T_MIN, tMax) over the quantized 4-wide nodes; a
// stack overflow answers with the skip-link any-hit walk.
template <int STACK, bool COUNT>
__device__ __forceinline__ bool occluded_wide_lane(const DevScene& S, double ox, double oy,
                                                   double oz, double dx, double dy, double dz,
                                                   double tMax, int* __restrict__ sNode,
                                                   int lanes, TraversalCounts& cnt) {
  const qw::QRay Q = qw::qray64(S, ox, oy, oz, dx, dy, dz);
  int cur = 0, sp = 0;
  while (cur != -1) {
    if (cur <= -2) {
      if (leaf_blocks(S, -2 - cur, ox, oy, oz, dx, dy, dz, tMax, cnt, COUNT)) return true;
      cur = -1;
    } else if (!qw::qwide_step<false, STACK, COUNT>(S, Q, qw::cull_bound(tMax), cur, sp, sNode,
                                                    (qw::Entry*)nullptr, lanes, cnt)) {
      if (COUNT) cnt.fallbacks++;
      return occluded<COUNT>(S, ox, oy, oz, dx, dy, dz, T_MIN, tMax, cnt);
    }
    if (cur == -1 && sp > 0) {
      sp--;
      cur = sNode[sp * lanes];
    }
  }
  return false;
}

// Scenes of at most PATHS_SCAN_MAX primitives: the per-path launch tests
// every primitive in packed order (runs of up to seven through the leaf
// tests) instead of walking.  The walks' contract -- every box on the way to
// a primitive whose binary64 test could return a nearer hit is entered -- makes
// the closest t the minimum over all primitives, which the scan computes with
// the same tests; exact-t ties are marked the same way (TIES_DETECT) and
// settled by settle_ties, and whether (tMin, tMax) is blocked does not depend
// on the order.  Lanes on one scene then run the same tests on the same
// records: no divergence between walks of different lengths, and a wave's
// record loads are scalar loads of one address.  Cornell box (14 primitives)
// 256x256x16: 0.99 -> 0.93 ms; the box with 8 to 50 more spheres (22-64
// primitives) 1.14-1.29x; 128 primitives 0.88x.
#ifndef CT_PATHS_SCAN_MAX
#define CT_PATHS_SCAN_MAX 64
#endif
constexpr int PATHS_SCAN_MAX = CT_PATHS_SCAN_MAX;

template <bool COUNT>
__device__ __forceinline__ void closest_scan_lane(const DevScene& S, double ox, double oy,
                                                  double oz, double dx, double dy, double dz,
                                                  Hit& h, TraversalCounts& cnt) {
  const int n = S.primCount;
  for (int p = 0; p < n; p += 7)
    test_leaf<COUNT, TIES_DETECT, true>(S, (p << 3) | min(7, n - p), ox, oy, oz, dx, dy, dz,
                                        T_MIN, h, cnt);
}

__device__ __forceinline__ bool occluded_scan_lane(const DevScene& S, double ox, double oy,
                                                   double oz, double dx, double dy, double dz,
                                                   double tMax, TraversalCounts& cnt,
                                                   bool count) {
  const int n = S.primCount;
  for (int p = 0; p < n; p += 7)
    if (leaf_blocks<true>(S, (p << 3) | min(7, n - p), ox, oy, oz, dx, dy, dz, tMax, cnt,
                          count))
      return true;
  return false;
}

// LDS stack entries of the per-path walks.
constexpr int PATHS_STACK = 24;
// Waves per SIMD of k_paths: four (127 VGPRs, 4 spilled) against three
// uncapped (139): 1.19x instead of 1.05x the wavefront on a 16k-triangle soup
// at 1920x1080x4, 1.00x instead of 0.88x on the mirror room.
#ifndef CT_PATHS_WAVES
#define CT_PATHS_WAVES 4
#endif
// The scan-only build (below) as well: shading holds the registers, so it
// needs 128 VGPRs like the others; at five waves (96 VGPRs, 32 spilled) the
// Cornell frame takes 3% longer, at six (80, 118 spilled) 8%.
#ifndef CT_PATHS_SCAN_WAVES
#define CT_PATHS_SCAN_WAVES 4
#endif

// Closest-hit and any-hit form of a k_paths build: the skip-link walk (the
// tracer's walk set to CT_WALK_STACKLESS), the 4-wide walk with the scan for
// lanes on scenes of at most PATHS_SCAN_MAX primitives, or the scan alone
// for rounds whose scenes all have at most PATHS_SCAN_MAX primitives (no
// walk code: 1% faster on the Cornell frame and the boxes of 22-64
// primitives).
enum PathsWalk : int { PATHS_SKIP = 0, PATHS_WIDE = 1, PATHS_SCAN = 2 };

// Whole paths, one lane at a time: a lane takes the next path of the group
// (one atomic per wave), starts it (start_path), then per bounce runs the
// closest hit (the scan, the 4-wide walk or the skip-link walk, as WALK says;
// exact-t ties settled in intersectClosest's order), one bounce of shading
// (shade_bounce) and the NEE ray's any-hit test, in the reference's order --
// the same arithmetic as the wavefront launches, so the same radiance bit for
// bit -- and when the path ends leaves its radiance and RNG state in the
// path's slot for k_finalize and takes the next path.  Lanes of a wave are at
// different bounces; each bounce's extension and NEE rays are counted in LDS
// and added to the group's counters once per workgroup.  A persistent grid
// (the workgroups resident at once).  For rounds over small scenes, where the
// wavefront's per-bounce launches (three per bounce and path group) cost more
// than the work they carry.
template <bool MULTI, bool COUNT, int WALK>
__global__ void __launch_bounds__(RBLOCK) __attribute__((
    amdgpu_waves_per_eu(WALK == PATHS_SCAN ? CT_PATHS_SCAN_WAVES : CT_PATHS_WAVES)))
k_paths(PathSoA P, const JobDev* __restrict__ jobs, int njobs, size_t base, size_t count,
        int nb) {
  constexpr bool WIDE = WALK == PATHS_WIDE;
  constexpr int SD = WIDE ? PATHS_STACK : 1;
  __shared__ int sNodeAll[SD * RBLOCK];
  __shared__ qw::Entry sEntryAll[SD * RBLOCK];
  __shared__ uint32_t sRays[2 * MAX_BOUNCE_SLOTS];  // [b] extension, [MAX_BOUNCE_SLOTS + b] NEE
  for (int k = threadIdx.x; k < 2 * MAX_BOUNCE_SLOTS; k += RBLOCK) sRays[k] = 0;
  __syncthreads();
  int* sNode = sNodeAll + threadIdx.x;
  qw::Entry* sEntry = sEntryAll + threadIdx.x;
  uint32_t* fetch = P.ctr + 2 * MAX_BOUNCE_SLOTS;  // the group's path cursor
  const uint32_t n = (uint32_t)count;
  bool have = false, exhausted = false;
  size_t pid = 0;
  const JobDev* J = jobs;
  PathState ps;
  int b = 0;
  TraversalCounts cnt = {0, 0, 0, 0}, scnt = {0, 0, 0, 0};
  for (;;) {
    if (!exhausted) {
      const uint32_t slot = refill(!have, n, fetch, exhausted);
      if (slot != 0xffffffffu) {
        pid = base + slot;
        J = jobs + job_of<MULTI, false>(jobs, njobs, pid);
        ps = start_path(*J, pid - J->pathBase);
        b = 0;
        have = true;
      }
    }
    if (!__any(have)) break;
    if (have) {
      const DevScene& S = J->S;
      atomicAdd(sRays + b, 1u);
      Hit h;
      h.t = __builtin_inf(); h.u = 0; h.v = 0; h.prim = -1; h.tie = false;
      const bool scan = WALK == PATHS_SCAN || (WIDE && S.primCount <= PATHS_SCAN_MAX);
      if (scan)
        closest_scan_lane<COUNT>(S, ps.ox, ps.oy, ps.oz, ps.dx, ps.dy, ps.dz, h, cnt);
      else if (WIDE)
        closest_wide_lane<SD, COUNT>(S, ps.ox, ps.oy, ps.oz, ps.dx, ps.dy, ps.dz, h, sNode,
                                     sEntry, RBLOCK, cnt);
      else
        closest_stackless<COUNT>(S, ps.ox, ps.oy, ps.oz, ps.dx, ps.dy, ps.dz, T_MIN, h, cnt);
      settle_ties<COUNT>(S, ps.ox, ps.oy, ps.oz, ps.dx, ps.dy, ps.dz, T_MIN, __builtin_inf(), h,
                         cnt);
      bool neeOn = false;
      NeeRay nee;
      const bool cont = shade_bounce(S, J->opts, ps, h.prim, h.t, b, neeOn, nee);
      if (neeOn) {
        atomicAdd(sRays + MAX_BOUNCE_SLOTS + b, 1u);
        const bool blocked =
            scan   ? occluded_scan_lane(S, ps.ox, ps.oy, ps.oz, nee.w[0], nee.w[1], nee.w[2],
                                        nee.tMax, scnt, COUNT)
            : WIDE ? occluded_wide_lane<SD, COUNT>(S, ps.ox, ps.oy, ps.oz, nee.w[0], nee.w[1],
                                                   nee.w[2], nee.tMax, sNode, RBLOCK, scnt)
                   : occluded<COUNT>(S, ps.ox, ps.oy, ps.oz, nee.w[0], nee.w[1], nee.w[2],
                                     T_MIN, nee.tMax, scnt);
        if (!blocked) {
          ps.Lr += nee.c[0];
          ps.Lg += nee.c[1];
          ps.Lb += nee.c[2];
        }
      }
      if (cont && b < J->opts.maxBounces) {
        b++;
      } else {
        P.Lr[pid] = ps.Lr;
        P.Lg[pid] = ps.Lg;
        P.Lb[pid] = ps.Lb;
        P.rng[pid] = ps.rng.s;
        have = false;
      }
    }
  }
  if (COUNT) {
    wave_add64(P.ctr64 + 0, cnt.nodes);
    wave_add64(P.ctr64 + 1, cnt.tris);
    wave_add64(P.ctr64 + 2, cnt.spheres);
    wave_add64(P.ctr64 + 3, scnt.nodes);
    wave_add64(P.ctr64 + 4, scnt.tris);
    wave_add64(P.ctr64 + 5, scnt.spheres);
    wave_add64(P.ctr64 + 6, cnt.fallbacks);
    wave_add64(P.ctr64 + 7, scnt.fallbacks);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nb; k += RBLOCK) {
    const uint32_t ext = sRays[k], nee = sRays[MAX_BOUNCE_SLOTS + k];
    if (ext) atomicAdd(P.ctr + k, ext);
    if (nee) atomicAdd(P.ctr + MAX_BOUNCE_SLOTS + k, nee);
  }
}

// Adds each pass's sample in pass order, dropping non-finite samples
// (tracer.ts:386-392); optional final scale (tracer.ts:352-353).
template <bool MULTI>
__global__ void __launch_bounds__(RBLOCK)
k_finalize(PathSoA P, const JobDev* __restrict__ jobs, int njobs, size_t totalPix) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gi < totalPix; gi += stride) {
    const JobDev& J = jobs[job_of<MULTI, true>(jobs, njobs, gi)];
    uint32_t pix = (uint32_t)(gi - J.pixelBase);
    if (J.rays) {  // radianceInto: the path's radiance as it is, and its RNG state
      const size_t pid = J.pathBase + pix;
      J.radiance[3 * (size_t)pix] = P.Lr[pid];
      J.radiance[3 * (size_t)pix + 1] = P.Lg[pid];
      J.radiance[3 * (size_t)pix + 2] = P.Lb[pid];
      J.rngIO[pix] = P.rng[pid];
      continue;
    }
    float* acc = J.accum + 3 * (size_t)pix;
    float a0 = acc[0], a1 = acc[1], a2 = acc[2];
    for (uint32_t s = 0; s < J.passes; s++) {
      size_t pid = J.pathBase + (size_t)s * J.npix + pix;
      double cr = P.Lr[pid], cg = P.Lg[pid], cb = P.Lb[pid];
      if (isfinite(cr) && isfinite(cg) && isfinite(cb)) {
        a0 = (float)((double)a0 + cr);
        a1 = (float)((double)a1 + cg);
        a2 = (float)((double)a2 + cb);
      }
    }
    if (J.doScale) {
      a0 = (float)((double)a0 * J.scale);
      a1 = (float)((double)a1 * J.scale);
      a2 = (float)((double)a2 * J.scale);
    }
    acc[0] = a0;
    acc[1] = a1;
    acc[2] = a2;
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace ct

using namespace ct;

struct ct_tracer {
  void* buf = nullptr;
  size_t capPaths = 0;
  PathSoA P{};
  PathSoA32 P32{};            // FP32 view of the same allocation (precision form)
  int precision = CT_PRECISION_F64;
  float* compBuf = nullptr;   // per-call compensation terms (precision form)
  size_t capComp = 0;
  JobDev* jobsDev = nullptr;  // device job table
  JobDev* jobsHost = nullptr; // pinned staging for the table
  size_t capJobs = 0;
  bool profiling = false;
  bool counting = false;
  int walk = CT_WALK_WIDE;           // closest-hit walk of extension rays
  int shadowWalk = CT_WALK_WIDE;     // any-hit walk of NEE rays
  int stack = 24;                    // LDS stack entries of the ordered walks
  // workgroups of each traversal launch (0: one lane per ray); up to four run
  // at once (two path groups, each with its NEE launch beside its extension
  // launch), so each takes a share of the chip.  Until a caller sets it, the
  // count follows the precision (traversal_blocks).
  unsigned persistBlocks = 768;
  bool blocksAuto = true;
  int block = 256;                   // lanes per workgroup of the traversal launches
  // paths in flight per round (~190 B of state each, allocated as rounds need
  // them): one round of a 1024^2 x 64 spp image is 2.4% (binary64) to 5.7%
  // (FP32) faster than two
  size_t maxPaths = size_t(1) << 26;
  bool graph = false;                // capture rounds into a hipGraph and replay them
  hipStream_t ownStream = nullptr;   // capture stream
  bool overlap = true;               // NEE launches on side streams (SideStream)
  int groups = 2;                    // path groups per round, each on its own stream
  int pathKernel = CT_PATHS_AUTO;    // wavefront launches or one per-path launch (k_paths)
  bool roundScan = false;  // every scene of the current round has at most PATHS_SCAN_MAX prims
  bool groupPaths[2] = {false, false};  // path group g of the current round runs k_paths
  SideStream side, sideB, groupB;    // streams and events, created on first use
  hipGraphExec_t graphExec[2] = {nullptr, nullptr};  // one per path group
  std::vector<char> graphKey;
  // stats of the last call
  int lastBounces = 0;
  uint64_t lastPaths = 0;
  uint64_t lastExtension = 0;
  uint64_t lastShadow = 0;
  uint32_t lastRounds = 0;
  uint32_t lastExtendLaunches = 0;
  uint32_t lastPathRounds = 0;  // rounds of the last call traced by k_paths
  std::vector<hipEvent_t> events;  // (begin, end) pairs per launch when profiling
  std::vector<int> eventKind;      // 0 extend, 1 shade, 2 shadow
  std::vector<hipEvent_t> eventPool;  // recorded before, reused by later calls
  hipStream_t lastStream = nullptr;
};

// Workgroups of a traversal launch: the caller's setting, or by default 640
// for the binary64 kernels and 768 for the FP32 ones (measured best on the
// 1080p soup frame: binary64 1147 Mrays/s at 640 against 1130 at 768 and 1134
// at 576; FP32 1237 at 768 against 1214 at 640).
static unsigned traversal_blocks(const ct_tracer* T) {
  if (!T->blocksAuto) return T->persistBlocks;
  return T->precision == CT_PRECISION_F32 ? 768u : 640u;
}

static void drop_graphs(ct_tracer* T) {
  for (hipGraphExec_t& g : T->graphExec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  T->graphKey.clear();
}

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
  size_t total = 21 * d + 6 * u32 + u8 + al(2 * GROUP_COUNTERS * 4) + al(16 * 8);
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
  P.jobOf = takeU();
  P.q0 = takeU(); P.q1 = takeU(); P.sPath = takeU();
  P.specular = (uint8_t*)p; p += u8;
  P.ctr = (uint32_t*)p; p += al(2 * GROUP_COUNTERS * 4);
  P.ctr64 = (unsigned long long*)p;
  // the FP32 state aliases the start of the same allocation (one precision
  // per round): 21 float arrays, 6 uint32 arrays and the flags fit in the
  // 21 binary64 arrays
  {
    char* f = (char*)T->buf;
    size_t f4 = al(n * 4);
    auto takeF = [&]() { float* r = (float*)f; f += f4; return r; };
    auto takeU32 = [&]() { uint32_t* r = (uint32_t*)f; f += f4; return r; };
    PathSoA32& Q = T->P32;
    Q.ox = takeF(); Q.oy = takeF(); Q.oz = takeF();
    Q.dx = takeF(); Q.dy = takeF(); Q.dz = takeF();
    Q.br = takeF(); Q.bg = takeF(); Q.bb = takeF();
    Q.Lr = takeF(); Q.Lg = takeF(); Q.Lb = takeF();
    Q.prevPdf = takeF(); Q.hitT = takeF();
    Q.sdx = takeF(); Q.sdy = takeF(); Q.sdz = takeF(); Q.sTmax = takeF();
    Q.sCr = takeF(); Q.sCg = takeF(); Q.sCb = takeF();
    Q.rng = takeU32();
    Q.hitPrim = (int32_t*)takeU32();
    Q.jobOf = takeU32();
    Q.q0 = takeU32(); Q.q1 = takeU32(); Q.sPath = takeU32();
    Q.specular = (uint8_t*)f; f += u8;
    Q.ctr = P.ctr;
    Q.ctr64 = P.ctr64;
  }
  T->capPaths = paths;
  // path buffers moved: a captured graph refers to the old ones
  drop_graphs(T);
  return CT_OK;
}

static int jobs_reserve(ct_tracer* T, size_t n) {
  if (n <= T->capJobs) return CT_OK;
  if (T->jobsDev) CT_HIP(hipFree(T->jobsDev));
  if (T->jobsHost) CT_HIP(hipHostFree(T->jobsHost));
  T->jobsDev = nullptr;
  T->jobsHost = nullptr;
  size_t cap = std::max<size_t>(n, 64);
  CT_HIP(hipMalloc((void**)&T->jobsDev, cap * sizeof(JobDev)));
  CT_HIP(hipHostMalloc((void**)&T->jobsHost, cap * sizeof(JobDev), hipHostMallocDefault));
  T->capJobs = cap;
  drop_graphs(T);
  return CT_OK;
}

Frame ct::make_frame(const double* f) {
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

// Returns the last call's events to the pool: a profiled launch then costs
// one hipEventRecord on each side, not an event creation (the timed steps of a
// benchmark run profiled).  A call's rounds end with a stream synchronisation,
// so its events are complete before the next call records them again.
static void clear_events(ct_tracer* T) {
  T->eventPool.insert(T->eventPool.end(), T->events.begin(), T->events.end());
  T->events.clear();
  T->eventKind.clear();
}

static void mark(ct_tracer* T, hipStream_t st, int kind, bool begin) {
  if (!T->profiling || T->graph) return;
  hipEvent_t e;
  if (!T->eventPool.empty()) {
    e = T->eventPool.back();
    T->eventPool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, st);
  T->events.push_back(e);
  if (begin) T->eventKind.push_back(kind);
}

// Instantiated stack depths of the ordered walks (LDS per 256-lane block:
// 6 B/entry for closest hit, 4 B/entry for any hit).
static bool stack_supported(int d) { return d == 8 || d == 16 || d == 24; }

template <int WALK, int STACK, bool COUNT, bool MULTI, int BLOCK>
static void launch_one(unsigned blocks, size_t paths, hipStream_t st, const PathSoA& P,
                       const JobDev* jobs, const uint32_t* q, const uint32_t* c, uint32_t* f,
                       const uint32_t* sc, uint32_t* sf, bool shadowPhase) {
  unsigned grid = grid_for(paths, BLOCK, blocks ? blocks : 1u << 30);
  if (!shadowPhase)
    hipLaunchKernelGGL((k_extend<WALK, STACK, COUNT, MULTI, BLOCK>), dim3(grid), dim3(BLOCK), 0,
                       st, P, jobs, q, c, f);
  else
    hipLaunchKernelGGL((k_shadow<WALK, STACK, COUNT, MULTI, BLOCK>), dim3(grid), dim3(BLOCK), 0,
                       st, P, jobs, sc, sf);
}

// Work-counting launches always use 256-lane workgroups (the counts do not
// depend on the workgroup shape).
template <int WALK, int STACK, bool MULTI>
static void launch_trace(bool counting, int block, unsigned blocks, size_t paths, hipStream_t st,
                         const PathSoA& P, const JobDev* jobs, const uint32_t* q,
                         const uint32_t* c, uint32_t* f, const uint32_t* sc, uint32_t* sf,
                         bool shadowPhase) {
  if (counting)
    launch_one<WALK, STACK, true, MULTI, 256>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else if (block == 64)
    launch_one<WALK, STACK, false, MULTI, 64>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else if (block == 128)
    launch_one<WALK, STACK, false, MULTI, 128>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
  else
    launch_one<WALK, STACK, false, MULTI, 256>(blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase);
}

template <bool MULTI>
static void dispatch_trace(int walk, int stack, bool counting, int block, unsigned blocks,
                           size_t paths, hipStream_t st, const PathSoA& P, const JobDev* jobs,
                           const uint32_t* q, const uint32_t* c, uint32_t* f, const uint32_t* sc,
                           uint32_t* sf, bool shadowPhase) {
#define CT_TRACE(W, D) \
  launch_trace<W, D, MULTI>(counting, block, blocks, paths, st, P, jobs, q, c, f, sc, sf, shadowPhase)
  if (walk == CT_WALK_STACKLESS) {
    CT_TRACE(CT_WALK_STACKLESS, 1);
  } else if (walk == CT_WALK_WIDE) {
    if (stack == 8) CT_TRACE(CT_WALK_WIDE, 8);
    else if (stack == 16) CT_TRACE(CT_WALK_WIDE, 16);
    else CT_TRACE(CT_WALK_WIDE, 24);
  } else {
    if (stack == 8) CT_TRACE(CT_WALK_ORDERED, 8);
    else if (stack == 16) CT_TRACE(CT_WALK_ORDERED, 16);
    else CT_TRACE(CT_WALK_ORDERED, 24);
  }
#undef CT_TRACE
}

// One job's share of a launch round.
struct RoundJob {
  const ct_scene* scene;
  Frame F;
  TraceOpts opts;
  uint32_t width, height;
  uint32_t frameBase;
  uint32_t passes;
  float* accum;
  int doScale;
  double scale;
  float* comp = nullptr;  // FP32 compensation terms
  int fold = 0;
  const double* rays = nullptr;  // radianceInto jobs (JobDev)
  uint32_t* rngIO = nullptr;
  double* radiance = nullptr;
};

static void mark_cb(void* ctx, hipStream_t st, int kind, bool begin) {
  mark((ct_tracer*)ctx, st, kind, begin);
}

static int make_side(SideStream& ss) {
  if (ss.side) return CT_OK;
  CT_HIP(hipStreamCreateWithFlags(&ss.side, hipStreamNonBlocking));
  CT_HIP(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
  CT_HIP(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
  return CT_OK;
}

int ct::tracer_streams(ct_tracer* T) {
  int rc = make_side(T->groupB);
  if (!rc) rc = make_side(T->side);
  if (!rc) rc = make_side(T->sideB);
  return rc;
}

// Smallest round traced as two path groups.
constexpr size_t SPLIT_MIN_PATHS = size_t(1) << 20;

// Rounds CT_PATHS_AUTO traces with k_paths: every scene of at most
// PATHS_AUTO_MAX_PRIMS primitives, at most PATHS_AUTO_MAX_PATHS paths.
// Measured against the wavefront (ms per frame, 256x256x16 / 1920x1080x4,
// sparse soups over the 4-wide tree with primitive boxes): 2k triangles
// 3.4x / 1.4x faster, 8k 2.9x / 1.6x, 16k 2.3x / 1.5x, 32k 1.8x / 1.1x, 64k
// 1.4x / 0.87x; the Cornell variant 2.9x at 256x256x16, 2.5x at 1024x1024x4;
// the mirror room (deep specular paths, 2k triangles) 1.27x / 1.00x.  With
// the limit at 32k instead of 16k the batching form's 64 jobs rendered one by
// one take 7% less time and the serving stream gains 1-2% (64k: 10% and 0%).
// The wavefront's per-bounce compaction wins back huge rounds: the 29 jobs on
// scenes of 1k-16k triangles of the 64-scene batch (95M paths in one round)
// take 46.7 ms in the wavefront, 49.8 ms per path.
#ifndef CT_PATHS_AUTO_MAX_PRIMS
#define CT_PATHS_AUTO_MAX_PRIMS 32768
#endif
constexpr int PATHS_AUTO_MAX_PRIMS = CT_PATHS_AUTO_MAX_PRIMS;
constexpr size_t PATHS_AUTO_MAX_PATHS = size_t(1) << 24;

// The streams of a round of `paths` paths.  Profiling events are recorded on
// each launch's own stream, so they bracket that launch as it runs beside the
// others.
static int round_streams(ct_tracer* T, size_t paths, bool perPath, RoundStreams& rs) {
  rs = RoundStreams();
  int rc;
  // a per-path launch is persistent and holds the whole chip by itself (one
  // group: Cornell 256x256x16 1.02 ms against 1.12 ms as two)
  if (T->groups > 1 && paths >= SPLIT_MIN_PATHS && !perPath) {
    if ((rc = make_side(T->groupB))) return rc;
    rs.groups = 2;
    rs.split = T->groupB;
  }
  // NEE side streams pay in eager rounds; inside a captured graph the side
  // branch replays behind its group's chain on this ROCm and only adds the
  // fork/join nodes (64-scene batch: 179.5 vs 176.6 ms)
  if (T->overlap && !T->graph) {
    if ((rc = make_side(T->side))) return rc;
    rs.shadow[0] = T->side;
    if (rs.groups == 2) {
      if ((rc = make_side(T->sideB))) return rc;
      rs.shadow[1] = T->sideB;
    }
  }
  return CT_OK;
}


// Workgroups of a persistent per-path launch of `count` paths: as many as the
// device holds at once (the occupancy of `kernel`, asked once per kernel and
// device), fewer when the paths do not need them.
static unsigned paths_grid(const void* kernel, size_t count) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, unsigned> resident;
  int dev = 0;
  (void)hipGetDevice(&dev);
  unsigned r;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = resident.find({kernel, dev});
    if (it == resident.end()) {
      int perCU = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, RBLOCK, 0) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          perCU < 1 || cus < 1) {
        perCU = 1;
        cus = 256;
      }
      it = resident.emplace(std::make_pair(kernel, dev), (unsigned)(perCU * cus)).first;
    }
    r = it->second;
  }
  const size_t need = (count + RBLOCK - 1) / RBLOCK;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(need, r));
}

// Kernel launches of one path group: generate, then per bounce extend, shade
// and the NEE launch (on the group's side stream when it has one).
template <bool MULTI>
static int launch_group_body(ct_tracer* T, const PathSoA& P, int njobs, size_t base,
                             size_t count, int nb, hipStream_t st, const SideStream& side,
                             bool perPath);

// A group's launches; its NEE side stream is joined back to `st` on every
// path out, so no later work on `st` can overlap a launch still on the side
// stream.
template <bool MULTI>
static int launch_group(ct_tracer* T, const PathSoA& P, int njobs, size_t base, size_t count,
                        int nb, hipStream_t st, const SideStream& side, bool perPath) {
  int rc = launch_group_body<MULTI>(T, P, njobs, base, count, nb, st, side, perPath);
  if (side.side) {
    int rj = check_hip(side_join(side, st), "side_join");
    if (!rc) rc = rj;
  }
  return rc;
}

template <bool MULTI>
static int launch_group_body(ct_tracer* T, const PathSoA& P, int njobs, size_t base,
                             size_t count, int nb, hipStream_t st, const SideStream& side,
                             bool perPath) {
  const JobDev* jobs = T->jobsDev;
  if (perPath) {  // the whole group in one persistent launch
    mark(T, st, 0, true);
    const int walk = T->walk == CT_WALK_STACKLESS ? PATHS_SKIP
                     : T->roundScan               ? PATHS_SCAN
                                                  : PATHS_WIDE;
#define CT_PATHS(C, W)                                                                          \
  hipLaunchKernelGGL((k_paths<MULTI, C, W>), dim3(paths_grid((const void*)k_paths<MULTI, C, W>, \
                                                             count)),                           \
                     dim3(RBLOCK), 0, st, P, jobs, njobs, base, count, nb)
    if (T->counting) {
      if (walk == PATHS_SCAN) CT_PATHS(true, PATHS_SCAN);
      else if (walk == PATHS_WIDE) CT_PATHS(true, PATHS_WIDE);
      else CT_PATHS(true, PATHS_SKIP);
    } else {
      if (walk == PATHS_SCAN) CT_PATHS(false, PATHS_SCAN);
      else if (walk == PATHS_WIDE) CT_PATHS(false, PATHS_WIDE);
      else CT_PATHS(false, PATHS_SKIP);
    }
#undef CT_PATHS
    CT_LAUNCH_CHECK("k_paths");
    mark(T, st, 0, false);
    return CT_OK;
  }
  unsigned gridPaths = grid_for(count, RBLOCK, 256 * 16);
  hipLaunchKernelGGL((k_generate<MULTI>), dim3(gridPaths), dim3(RBLOCK), 0, st, P, jobs, njobs,
                     base, count);
  CT_LAUNCH_CHECK("k_generate");
  for (int b = 0; b < nb; b++) {
    uint32_t* qin = (b & 1) ? P.q1 : P.q0;
    uint32_t* qout = (b & 1) ? P.q0 : P.q1;
    uint32_t* cin = P.ctr + b;
    uint32_t* cout = P.ctr + b + 1;
    uint32_t* scount = P.ctr + MAX_BOUNCE_SLOTS + b;
    uint32_t* fExt = P.ctr + 2 * MAX_BOUNCE_SLOTS + b;
    uint32_t* fSh = P.ctr + 3 * MAX_BOUNCE_SLOTS + b;
    mark(T, st, 0, true);
    dispatch_trace<MULTI>(T->walk, T->stack, T->counting, T->block, traversal_blocks(T), count, st,
                          P, jobs, qin, cin, fExt, scount, fSh, false);
    CT_LAUNCH_CHECK("k_extend");
    mark(T, st, 0, false);
    if (side.side && b > 0) CT_HIP(side_join(side, st));  // NEE rays of bounce b-1 done
    mark(T, st, 1, true);
    hipLaunchKernelGGL((k_shade<MULTI>), dim3(gridPaths), dim3(RBLOCK), 0, st, P, jobs, qin,
                       cin, qout, cout, scount, b);
    CT_LAUNCH_CHECK("k_shade");
    mark(T, st, 1, false);
    hipStream_t sst = st;
    if (side.side) {
      CT_HIP(side_fork(side, st));
      sst = side.side;
    }
    mark(T, sst, 2, true);
    dispatch_trace<MULTI>(T->shadowWalk, T->stack, T->counting, T->block, traversal_blocks(T),
                          count, sst, P, jobs, qin, cin, fExt, scount, fSh, true);
    CT_LAUNCH_CHECK("k_shadow");
    mark(T, sst, 2, false);
  }
  return CT_OK;
}

// One part of a round's launches (RoundPart), graph-capturable.
template <bool MULTI>
static int launch_part(ct_tracer* T, const RoundStreams& rs, int njobs, size_t paths, size_t pix,
                       int nb, int part, hipStream_t st) {
  if (T->precision == CT_PRECISION_F32) {
    RoundLaunch R{T->P32, T->jobsDev, njobs, paths, pix, nb, T->stack, T->block,
                  traversal_blocks(T), T->counting, mark_cb, T, rs};
    return launch_part32(R, MULTI, part, st);
  }
  const PathSoA& P = T->P;
  if (part == ROUND_CLEAR) {
    hipLaunchKernelGGL(k_clear, dim3(8), dim3(256), 0, st, P.ctr, 2 * GROUP_COUNTERS, P.ctr64,
                       0);
    CT_LAUNCH_CHECK("k_clear");
  } else if (part == ROUND_FINAL) {
    unsigned gridPix = grid_for(pix, RBLOCK, 256 * 16);
    hipLaunchKernelGGL((k_finalize<MULTI>), dim3(gridPix), dim3(RBLOCK), 0, st, P, T->jobsDev,
                       njobs, pix);
    CT_LAUNCH_CHECK("k_finalize");
  } else {
    const int g = part - ROUND_GROUP0;
    size_t base, count;
    group_range(rs, paths, g, base, count);
    return launch_group<MULTI>(T, group_view(P, base, g), njobs, base, count, nb, st,
                               rs.shadow[g], T->groupPaths[g]);
  }
  return CT_OK;
}

// The parts of a round around its path groups: counters cleared on `st`, the
// groups on their streams (forked from and joined back to `st`), the sums
// finalized on `st`.  `group` launches group g on its stream.
template <typename GroupFn>
static int round_around_groups(const RoundStreams& rs, hipStream_t st, GroupFn group) {
  if (rs.groups == 2) CT_HIP(side_fork(rs.split, st));
  int rc = group(0, st);
  if (!rc && rs.groups == 2) rc = group(1, rs.split.side);
  if (rs.groups == 2) {  // joined on every path out
    int rj = check_hip(side_join(rs.split, st), "side_join");
    if (!rc) rc = rj;
  }
  return rc;
}

template <bool MULTI>
static int launch_round(ct_tracer* T, const RoundStreams& rs, int njobs, size_t paths,
                        size_t pix, int nb, hipStream_t st) {
  int rc = launch_part<MULTI>(T, rs, njobs, paths, pix, nb, ROUND_CLEAR, st);
  if (!rc)
    rc = round_around_groups(rs, st, [&](int g, hipStream_t gs) {
      return launch_part<MULTI>(T, rs, njobs, paths, pix, nb, ROUND_GROUP0 + g, gs);
    });
  if (!rc) rc = launch_part<MULTI>(T, rs, njobs, paths, pix, nb, ROUND_FINAL, st);
  return rc;
}

struct BatchTotals {
  uint64_t ext = 0, shadow = 0;
};

// Whether job j's paths would run k_paths instead of the wavefront launches:
// asked for, or (CT_PATHS_AUTO) when its scene has at most PATHS_AUTO_MAX_PRIMS
// primitives, where a bounce's launches and tails cost more than the
// per-bounce compaction of the wavefront saves.  k_paths is the binary64
// arithmetic; the FP32 precision form always runs its wavefront.
static bool scene_uses_paths(const ct_tracer* T, const ct_scene* scene) {
  if (T->precision != CT_PRECISION_F64 || T->pathKernel == CT_PATHS_WAVEFRONT) return false;
  return T->pathKernel == CT_PATHS_PER_PATH || scene->dev.primCount <= PATHS_AUTO_MAX_PRIMS;
}
static bool job_uses_paths(const ct_tracer* T, const RoundJob& j) {
  return scene_uses_paths(T, j.scene);
}

// Traces one round: every job's passes, all jobs' paths sharing the launches.
// The round runs k_paths when every job of it would (job_uses_paths): within
// a round that also holds larger scenes the wavefront launches amortize
// their costs over all paths, and a persistent per-path launch beside them
// would take the CU slots their path groups overlap in (measured: the
// 64-scene batch 160 -> 204 ms, the serving stream 1424 -> 1290 spp*Mpix/s
// with the small scenes' paths split off into a per-path group).
static int run_round(ct_tracer* T, const std::vector<RoundJob>& R, hipStream_t st,
                     BatchTotals& tot) {
  bool allPaths = !R.empty();
  size_t roundPaths = 0;
  for (const RoundJob& j : R) {
    allPaths = allPaths && job_uses_paths(T, j);
    roundPaths += (size_t)j.width * j.height * j.passes;
  }
  if (T->pathKernel == CT_PATHS_AUTO && roundPaths > PATHS_AUTO_MAX_PATHS) allPaths = false;
  size_t paths = 0, pix = 0;
  int maxB = 0;
  for (const RoundJob& j : R) {
    paths += (size_t)j.width * j.height * j.passes;
    pix += (size_t)j.width * j.height;
    maxB = std::max(maxB, j.opts.maxBounces);
  }
  if (paths == 0) return CT_OK;
  if (paths > 0xffffffffu) return set_error(CT_ELIMIT, "render: more than 2^32 paths in a round");
  int rc = tracer_reserve(T, paths);
  if (rc) return rc;
  rc = jobs_reserve(T, R.size());
  if (rc) return rc;
  // the previous round's table copy must be done before the staging is reused
  CT_HIP(hipStreamSynchronize(st));
  size_t pb = 0, xb = 0;
  for (size_t k = 0; k < R.size(); k++) {
    const RoundJob& j = R[k];
    JobDev& d = T->jobsHost[k];
    d.S = j.scene->dev;
    d.F = j.F;
    d.opts = j.opts;
    d.width = j.width;
    d.height = j.height;
    d.npix = j.width * j.height;
    d.passes = j.passes;
    d.frameBase = j.frameBase;
    d.pathBase = (uint32_t)pb;
    d.pixelBase = (uint32_t)xb;
    d.doScale = j.doScale;
    d.scale = j.scale;
    d.accum = j.accum;
    d.comp = j.comp;
    d.fold = j.fold;
    d.rays = j.rays;
    d.rngIO = j.rngIO;
    d.radiance = j.radiance;
    pb += (size_t)d.npix * j.passes;
    xb += d.npix;
  }
  CT_HIP(hipMemcpyAsync(T->jobsDev, T->jobsHost, R.size() * sizeof(JobDev),
                        hipMemcpyHostToDevice, st));
  int nb = maxB + 1;
  bool multi = R.size() > 1;
  RoundStreams rs;
  rc = round_streams(T, paths, allPaths, rs);
  if (rc) return rc;
  T->groupPaths[0] = T->groupPaths[1] = allPaths;
  T->roundScan = allPaths;
  for (const RoundJob& j : R)
    T->roundScan = T->roundScan && j.scene->dev.primCount <= PATHS_SCAN_MAX;
  if (T->graph) {
    // each path group's launches are one captured graph, replayed on the
    // group's stream; they depend only on the round's shape and the tracer's
    // settings, and read the job table from device memory at replay
    std::vector<char> key(sizeof(size_t) * 2 + sizeof(int) * 11);
    size_t kk[2] = {paths, pix};
    int ki[11] = {nb, (int)R.size(), T->walk, T->shadowWalk, T->stack, (int)traversal_blocks(T),
                  T->counting, multi, T->block,
                  T->precision * 64 + (T->overlap ? 32 : 0) + rs.groups,
                  (T->groupPaths[0] ? 1 : 0) + (T->groupPaths[1] ? 2 : 0) +
                      (T->roundScan ? 4 : 0)};
    memcpy(key.data(), kk, sizeof kk);
    memcpy(key.data() + sizeof kk, ki, sizeof ki);
    if (!T->graphExec[0] || key != T->graphKey) {
      drop_graphs(T);
      if (!T->ownStream) CT_HIP(hipStreamCreateWithFlags(&T->ownStream, hipStreamNonBlocking));
      CT_HIP(hipStreamSynchronize(st));
      for (int g = 0; g < rs.groups; g++) {
        hipStream_t cs = g == 0 ? T->ownStream : rs.split.side;
        CT_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        rc = multi ? launch_part<true>(T, rs, (int)R.size(), paths, pix, nb, ROUND_GROUP0 + g, cs)
                   : launch_part<false>(T, rs, (int)R.size(), paths, pix, nb, ROUND_GROUP0 + g,
                                        cs);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &graph);
        if (rc) {
          if (graph) (void)hipGraphDestroy(graph);
          drop_graphs(T);
          return rc;
        }
        if (e != hipSuccess) {
          drop_graphs(T);
          return check_hip(e, "hipStreamEndCapture");
        }
        e = hipGraphInstantiate(&T->graphExec[g], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
          drop_graphs(T);
          return check_hip(e, "hipGraphInstantiate");
        }
      }
      T->graphKey = key;
    }
    rc = multi ? launch_part<true>(T, rs, (int)R.size(), paths, pix, nb, ROUND_CLEAR, st)
               : launch_part<false>(T, rs, (int)R.size(), paths, pix, nb, ROUND_CLEAR, st);
    if (!rc)
      rc = round_around_groups(rs, st, [&](int g, hipStream_t gs) {
        return check_hip(hipGraphLaunch(T->graphExec[g], gs), "hipGraphLaunch");
      });
    if (!rc)
      rc = multi ? launch_part<true>(T, rs, (int)R.size(), paths, pix, nb, ROUND_FINAL, st)
                 : launch_part<false>(T, rs, (int)R.size(), paths, pix, nb, ROUND_FINAL, st);
    if (rc) return rc;
  } else {
    rc = multi ? launch_round<true>(T, rs, (int)R.size(), paths, pix, nb, st)
               : launch_round<false>(T, rs, (int)R.size(), paths, pix, nb, st);
    if (rc) return rc;
  }
  std::vector<uint32_t> h(2 * GROUP_COUNTERS);
  CT_HIP(hipMemcpyAsync(h.data(), T->P.ctr, h.size() * 4, hipMemcpyDeviceToHost, st));
  CT_HIP(hipStreamSynchronize(st));
  for (int g = 0; g < 2; g++)
    for (int b = 0; b < nb; b++) {
      tot.ext += h[g * GROUP_COUNTERS + b];
      tot.shadow += h[g * GROUP_COUNTERS + MAX_BOUNCE_SLOTS + b];
    }
  T->lastBounces = std::max(T->lastBounces, nb);
  T->lastRounds++;
  for (int g = 0; g < rs.groups; g++) T->lastExtendLaunches += T->groupPaths[g] ? 1u : (uint32_t)nb;
  T->lastPathRounds += T->groupPaths[0] ? 1 : 0;
  return CT_OK;
}

// Packs jobs into rounds of at most maxPaths paths (whole passes per job) and
// traces them; returns the totals.
int ct::run_plan(ct_tracer* T, std::vector<PlanJob> jobs, hipStream_t st) {
  clear_events(T);
  T->lastStream = st;
  T->lastBounces = 0;
  T->lastRounds = 0;
  T->lastExtendLaunches = 0;
  T->lastPathRounds = 0;
  BatchTotals tot;
  uint64_t paths = 0;
  for (const PlanJob& j : jobs) {
    size_t npix = (size_t)j.width * j.height;
    if (j.count && npix > T->maxPaths)
      return set_error(CT_ELIMIT, "render: one pass of a job exceeds the tracer's path capacity");
    paths += (uint64_t)npix * j.count;
  }
  if (T->counting && paths) {
    // work counters sum over the rounds of one call
    int rc = tracer_reserve(T, std::min<uint64_t>(paths, T->maxPaths));
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, T->P.ctr, 0, T->P.ctr64, 16);
    CT_LAUNCH_CHECK("k_clear");
  }
  // FP32 precision form: compensation terms of jobs without their own live in
  // a per-call scratch buffer, zeroed here and folded into the sums at the end
  std::vector<float*> comp(jobs.size(), nullptr);
  if (T->precision == CT_PRECISION_F32) {
    size_t need = 0;
    for (const PlanJob& j : jobs)
      if (!j.comp) need += (size_t)j.width * j.height * 3;
    if (need > T->capComp) {
      if (T->compBuf) CT_HIP(hipFree(T->compBuf));
      T->compBuf = nullptr;
      T->capComp = 0;
      CT_HIP(hipMalloc((void**)&T->compBuf, need * sizeof(float)));
      T->capComp = need;
    }
    int rc = zero_floats(T->compBuf, need, st);
    if (rc) return rc;
    size_t off = 0;
    for (size_t k = 0; k < jobs.size(); k++) {
      if (jobs[k].comp) {
        comp[k] = jobs[k].comp;
      } else {
        comp[k] = T->compBuf + off;
        off += (size_t)jobs[k].width * jobs[k].height * 3;
      }
    }
  }
  std::vector<uint32_t> done(jobs.size(), 0);
  // Jobs whose paths run k_paths (job_uses_paths) are packed into rounds of
  // their own, traced before the others' wavefront rounds, so that neither
  // kind waits on the other's launches -- when they fit the per-path rounds'
  // size limit.  A tracer in graph mode keeps one round shape (it replays the
  // graph of the last one) and packs all jobs together.
  std::vector<int> kind(jobs.size(), 0);
  bool split = false;
  if (!T->graph) {
    bool anyWave = false;
    uint64_t smallPaths = 0;
    for (size_t k = 0; k < jobs.size(); k++) {
      kind[k] = scene_uses_paths(T, jobs[k].scene) ? 0 : 1;
      if (kind[k] == 0) smallPaths += (uint64_t)jobs[k].width * jobs[k].height * jobs[k].count;
      anyWave = anyWave || kind[k] == 1;
    }
    split = smallPaths > 0 && anyWave &&
            (T->pathKernel != CT_PATHS_AUTO || smallPaths <= PATHS_AUTO_MAX_PATHS);
  }
  for (int pass = split ? 0 : 1; pass < 2;) {
    std::vector<RoundJob> R;
    size_t budget = T->maxPaths;
    for (size_t k = 0; k < jobs.size(); k++) {
      const PlanJob& j = jobs[k];
      if (split && kind[k] != pass) continue;
      uint32_t left = j.count - done[k];
      size_t npix = (size_t)j.width * j.height;
      if (left == 0 || npix == 0) continue;
      size_t fit = budget / npix;
      if (fit == 0) continue;
      uint32_t p = (uint32_t)std::min<size_t>(left, fit);
      RoundJob r;
      r.scene = j.scene;
      r.F = j.F;
      r.opts = j.opts;
      r.width = j.width;
      r.height = j.height;
      r.frameBase = j.frameBase + j.first + done[k];
      r.passes = p;
      r.accum = j.accum;
      r.doScale = (j.finish && done[k] + p == j.count) ? 1 : 0;
      r.scale = j.scale;
      r.comp = comp[k];
      r.fold = (comp[k] && !j.comp && done[k] + p == j.count) ? 1 : 0;
      r.rays = j.rays;
      r.rngIO = j.rngIO;
      r.radiance = j.radiance;
      R.push_back(r);
      done[k] += p;
      budget -= p * npix;
    }
    if (R.empty()) {
      pass++;
      continue;
    }
    int rc = run_round(T, R, st, tot);
    if (rc) return rc;
  }
  T->lastExtension = tot.ext;
  T->lastShadow = tot.shadow;
  T->lastPaths = paths;
  return CT_OK;
}

int ct::zero_floats(float* p, size_t n, hipStream_t st) {
  if (n == 0) return CT_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, RBLOCK, 256 * 16)), dim3(RBLOCK), 0, st, p, n, 0.0f);
  CT_LAUNCH_CHECK("k_fill");
  return CT_OK;
}

size_t ct::tracer_capacity(const ct_tracer* T) { return T->maxPaths; }

// Value a render's image starts from: zero, or with no samples NaN, which is
// what renderImage leaves (tracer.ts:351-353: the zero sums times 1/0).
static float image_start(uint32_t spp) {
  return spp ? 0.0f : std::numeric_limits<float>::quiet_NaN();
}

static int check_job(const ct_scene* s, const double* frame, uint32_t w, uint32_t h,
                     int32_t maxBounces, const void* out) {
  if (!s || !frame || !out) return set_error(CT_EINVAL, "render: null argument");
  if (maxBounces < 0 || maxBounces + 2 > MAX_BOUNCE_SLOTS)
    return set_error(CT_ELIMIT, "render: maxBounces must be in [0, 254]");
  if ((size_t)w * h > (size_t)1 << 26) return set_error(CT_ELIMIT, "render: more than 2^26 pixels");
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
  for (hipEvent_t e : T->eventPool) (void)hipEventDestroy(e);
  T->eventPool.clear();
  hipError_t e = hipSuccess;
  drop_graphs(T);
  if (T->buf) e = hipFree(T->buf);
  if (T->jobsDev) (void)hipFree(T->jobsDev);
  if (T->jobsHost) (void)hipHostFree(T->jobsHost);
  if (T->compBuf) (void)hipFree(T->compBuf);
  if (T->ownStream) (void)hipStreamDestroy(T->ownStream);
  for (SideStream* ss : {&T->side, &T->sideB, &T->groupB}) {
    if (ss->side) (void)hipStreamDestroy(ss->side);
    if (ss->fork) (void)hipEventDestroy(ss->fork);
    if (ss->join) (void)hipEventDestroy(ss->join);
  }
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
    return set_error(CT_EINVAL, "ct_tracer_set_traversal: stack depth must be 8, 16 or 24");
  T->shadowWalk = shadowWalk;
  T->stack = stackDepth;
  return CT_OK;
}

int ct_tracer_set_schedule(ct_tracer* T, int32_t persistent, int32_t blocks) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (persistent && (blocks < 1 || blocks > 65535))
    return set_error(CT_EINVAL, "ct_tracer_set_schedule: blocks must be in [1, 65535]");
  T->persistBlocks = persistent ? (unsigned)blocks : 0u;
  T->blocksAuto = false;
  return CT_OK;
}

int ct_tracer_set_precision(ct_tracer* T, int32_t precision) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (precision != CT_PRECISION_F64 && precision != CT_PRECISION_F32)
    return set_error(CT_EINVAL, "ct_tracer_set_precision: unknown precision");
  T->precision = precision;
  return CT_OK;
}

int ct_tracer_set_workgroup(ct_tracer* T, int32_t lanes) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (lanes != 64 && lanes != 128 && lanes != 256)
    return set_error(CT_EINVAL, "ct_tracer_set_workgroup: lanes must be 64, 128 or 256");
  T->block = lanes;
  return CT_OK;
}

int ct_tracer_get_config(const ct_tracer* T, ct_tracer_config* out) {
  if (!T || !out) return set_error(CT_EINVAL, "ct_tracer_get_config: null argument");
  out->maxPaths = T->maxPaths;
  out->walk = T->walk;
  out->shadowWalk = T->shadowWalk;
  out->stackDepth = T->stack;
  const unsigned blocks = traversal_blocks(T);
  out->persistent = blocks ? 1 : 0;
  out->blocks = blocks ? (int32_t)blocks : (T->precision == CT_PRECISION_F32 ? 768 : 640);
  out->workgroup = T->block;
  out->pathGroups = T->groups;
  return CT_OK;
}

int ct_tracer_set_config(ct_tracer* T, const ct_tracer_config* c) {
  if (!T || !c) return set_error(CT_EINVAL, "ct_tracer_set_config: null argument");
  ct_tracer_config old;
  ct_tracer_get_config(T, &old);
  const bool oldAuto = T->blocksAuto;
  int rc = ct_tracer_set_walk(T, c->walk);
  if (!rc) rc = ct_tracer_set_traversal(T, c->shadowWalk, c->stackDepth);
  if (!rc) rc = ct_tracer_set_schedule(T, c->persistent, c->blocks);
  if (!rc) rc = ct_tracer_set_workgroup(T, c->workgroup);
  if (!rc) rc = ct_tracer_set_capacity(T, c->maxPaths);
  if (!rc && (c->pathGroups < 1 || c->pathGroups > 2))
    rc = set_error(CT_EINVAL, "ct_tracer_set_config: pathGroups must be 1 or 2");
  if (!rc) T->groups = c->pathGroups;
  if (rc) {  // all or nothing
    std::string msg = ct_last_error();
    T->walk = old.walk;
    T->shadowWalk = old.shadowWalk;
    T->stack = old.stackDepth;
    T->persistBlocks = old.persistent ? (unsigned)old.blocks : 0u;
    T->blocksAuto = oldAuto;
    T->block = old.workgroup;
    T->maxPaths = old.maxPaths;
    T->groups = old.pathGroups;
    return set_error(rc, msg);
  }
  return CT_OK;
}

int ct_tracer_set_capacity(ct_tracer* T, uint64_t maxPaths) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (maxPaths < 1024 || maxPaths > ((uint64_t)1 << 31))
    return set_error(CT_EINVAL, "ct_tracer_set_capacity: maxPaths must be in [1024, 2^31]");
  T->maxPaths = (size_t)maxPaths;
  return CT_OK;
}

int ct_tracer_set_graph(ct_tracer* T, int32_t enable) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  T->graph = enable != 0;
  if (!T->graph) drop_graphs(T);
  return CT_OK;
}

int ct_tracer_set_overlap(ct_tracer* T, int32_t enable) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  T->overlap = enable != 0;
  drop_graphs(T);  // a captured round keeps the streams it was captured with
  return CT_OK;
}

int ct_tracer_set_path_kernel(ct_tracer* T, int32_t mode) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (mode != CT_PATHS_AUTO && mode != CT_PATHS_WAVEFRONT && mode != CT_PATHS_PER_PATH)
    return set_error(CT_EINVAL, "ct_tracer_set_path_kernel: unknown mode");
  T->pathKernel = mode;
  return CT_OK;
}

int ct_tracer_get_path_kernel(const ct_tracer* T, int32_t* mode, uint32_t* perPathRounds) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (mode) *mode = T->pathKernel;
  if (perPathRounds) *perPathRounds = T->lastPathRounds;
  return CT_OK;
}

int ct_tracer_set_walk(ct_tracer* T, int32_t walk) {
  if (!T) return set_error(CT_EINVAL, "null tracer");
  if (walk != CT_WALK_ORDERED && walk != CT_WALK_STACKLESS && walk != CT_WALK_WIDE)
    return set_error(CT_EINVAL, "ct_tracer_set_walk: unknown walk");
  T->walk = walk;
  return CT_OK;
}

int ct_accumulate_pass(ct_tracer* T, const ct_scene* s, const double* frame16, uint32_t width,
                       uint32_t height, uint32_t frameIndex, const ct_trace_opts* opts,
                       float* accum, void* stream) {
  if (!T || !opts) return set_error(CT_EINVAL, "ct_accumulate_pass: null argument");
  int rc = check_job(s, frame16, width, height, opts->maxBounces, accum);
  if (rc) return rc;
  PlanJob j{s, make_frame(frame16), {opts->maxBounces, opts->russianRoulette != 0, opts->nee != 0},
            width, height, frameIndex, 0, 1, accum, false, 1.0};
  return run_plan(T, {j}, (hipStream_t)stream);
}

int ct_render_image(ct_tracer* T, const ct_scene* s, const double* frame16,
                    const ct_render_opts* opts, float* out, void* stream) {
  if (!T || !opts) return set_error(CT_EINVAL, "ct_render_image: null argument");
  int rc = check_job(s, frame16, opts->width, opts->height, opts->maxBounces, out);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  size_t n = (size_t)opts->width * opts->height * 3;
  if (n) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, RBLOCK, 256 * 16)), dim3(RBLOCK), 0, st, out, n,
                       image_start(opts->spp));
    CT_LAUNCH_CHECK("k_fill");
  }
  PlanJob j{s, make_frame(frame16),
            {opts->maxBounces, opts->russianRoulette != 0, opts->nee != 0},
            opts->width, opts->height, opts->seed * 7919u, 0, opts->spp, out, true,
            opts->spp ? 1.0 / opts->spp : 1.0};
  return run_plan(T, {j}, st);
}

int ct_radiance(ct_tracer* T, const ct_scene* s, const double* rays, uint32_t n,
                uint32_t* rngStates, const ct_trace_opts* opts, double* out, void* stream) {
  if (!T || !s || !opts || (n && (!rays || !rngStates || !out)))
    return set_error(CT_EINVAL, "ct_radiance: null argument");
  if (opts->maxBounces < 0 || opts->maxBounces + 2 > MAX_BOUNCE_SLOTS)
    return set_error(CT_ELIMIT, "ct_radiance: maxBounces must be in [0, 254]");
  if (T->precision != CT_PRECISION_F64)
    return set_error(CT_EINVAL, "ct_radiance: radianceInto is traced in binary64 only");
  if (n == 0) return CT_OK;
  // rays beyond the path capacity go in further jobs of the same plan
  std::vector<PlanJob> plan;
  const size_t cap = T->maxPaths;
  for (size_t k = 0; k < n; k += cap) {
    uint32_t m = (uint32_t)std::min<size_t>(cap, n - k);
    PlanJob j{s, Frame{}, {opts->maxBounces, opts->russianRoulette != 0, opts->nee != 0},
              m, 1, 0, 0, 1, nullptr, false, 1.0};
    j.rays = rays + 6 * k;
    j.rngIO = rngStates + k;
    j.radiance = out + 3 * k;
    plan.push_back(j);
  }
  return run_plan(T, plan, (hipStream_t)stream);
}

int ct_render_batch(ct_tracer* T, const ct_job* jobs, uint32_t n, void* stream) {
  if (!T || (n && !jobs)) return set_error(CT_EINVAL, "ct_render_batch: null argument");
  hipStream_t st = (hipStream_t)stream;
  std::vector<PlanJob> plan;
  for (uint32_t k = 0; k < n; k++) {
    const ct_job& J = jobs[k];
    int rc = check_job(J.scene, J.frame, J.width, J.height, J.maxBounces, J.d_out);
    if (rc) return rc;
    size_t m = (size_t)J.width * J.height * 3;
    if (m) {
      hipLaunchKernelGGL(k_fill, dim3(grid_for(m, RBLOCK, 256 * 16)), dim3(RBLOCK), 0, st, J.d_out,
                         m, image_start(J.spp));
      CT_LAUNCH_CHECK("k_fill");
    }
    plan.push_back({J.scene, make_frame(J.frame),
                    {J.maxBounces, J.russianRoulette != 0, J.nee != 0}, J.width, J.height,
                    J.seed * 7919u, 0, J.spp, J.d_out, true, J.spp ? 1.0 / J.spp : 1.0});
  }
  return run_plan(T, plan, st);
}

int ct_accumulate_batch(ct_tracer* T, const ct_pass_job* jobs, uint32_t n, void* stream) {
  if (!T || (n && !jobs)) return set_error(CT_EINVAL, "ct_accumulate_batch: null argument");
  std::vector<PlanJob> plan;
  for (uint32_t k = 0; k < n; k++) {
    const ct_pass_job& J = jobs[k];
    int rc = check_job(J.scene, J.frame, J.width, J.height, J.maxBounces, J.d_accum);
    if (rc) return rc;
    plan.push_back({J.scene, make_frame(J.frame),
                    {J.maxBounces, J.russianRoulette != 0, J.nee != 0}, J.width, J.height,
                    J.frameIndex, 0, J.passes, J.d_accum, false, 1.0});
  }
  return run_plan(T, plan, (hipStream_t)stream);
}

int ct_tracer_stats(ct_tracer* T, ct_render_stats* out) {
  if (!T || !out) return set_error(CT_EINVAL, "ct_tracer_stats: null argument");
  memset(out, 0, sizeof *out);
  CT_HIP(hipStreamSynchronize(T->lastStream));
  out->extensionRays = T->lastExtension;
  out->shadowRays = T->lastShadow;
  out->paths = T->lastPaths;
  out->rounds = T->lastRounds;
  if (T->profiling && !T->graph) {
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
    out->extendLaunches = T->lastExtendLaunches;
  }
  if (T->counting && T->buf) {
    unsigned long long h[8];
    CT_HIP(hipMemcpy(h, T->P.ctr64, sizeof h, hipMemcpyDeviceToHost));
    out->nodeVisits = h[0];
    out->triTests = h[1];
    out->sphereTests = h[2];
    out->shadowNodeVisits = h[3];
    out->shadowTriTests = h[4];
    out->shadowSphereTests = h[5];
    out->stackFallbacks = h[6];
    out->shadowStackFallbacks = h[7];
  }
  return CT_OK;
}

}  // extern "C"
