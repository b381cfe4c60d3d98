// Math.sin, Math.cos and Math.cbrt with the JS engine's results, bit for bit,
// for the binary64 shading kernels.
//
// The reference draws directions with Math.cos/Math.sin(2*pi*u) and rough
// reflections with Math.cbrt(u) (materials.ts:50-51,71-72,115, tracer.ts:117,
// 380-382).  V8 computes these with its ieee754 routines, restatements of
// fdlibm 5.3 (argument reduction e_rem_pio2, kernels k_sin / k_cos) and of
// FreeBSD msun's s_cbrt.  The device library's sin/cos/cbrt are accurate to an
// ulp or so but not these functions: a last-ulp difference in a direction
// can send a path of several mirror bounces past a triangle edge it grazes,
// and the pixel then differs from the reference.  So the shading kernels use
// the same published algorithms, in binary64 with every operation rounded
// once (-ffp-contract=off): sin/cos for |x| up to 2^19 * pi/2 (the kernels pass
// 2*pi*u, u in [0,1)), cbrt for every finite x.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ct {
namespace jsm {

// High words of the doubles nearest n*pi/2, n = 1..32: an argument whose high
// word equals one of them may cancel in the first reduction step.
static __constant__ const uint32_t NPIO2_HW[32] = {
    0x3FF921FBu, 0x400921FBu, 0x4012D97Cu, 0x401921FBu, 0x401F6A7Au, 0x4022D97Cu,
    0x4025FDBBu, 0x402921FBu, 0x402C463Au, 0x402F6A7Au, 0x4031475Cu, 0x4032D97Cu,
    0x40346B9Cu, 0x4035FDBBu, 0x40378FDBu, 0x403921FBu, 0x403AB41Bu, 0x403C463Au,
    0x403DD85Au, 0x403F6A7Au, 0x40407E4Cu, 0x4041475Cu, 0x4042106Cu, 0x4042D97Cu,
    0x4043A28Cu, 0x40446B9Cu, 0x404534ACu, 0x4045FDBBu, 0x4046C6CBu, 0x40478FDBu,
    0x404858EBu, 0x404921FBu};

__device__ __forceinline__ uint32_t hi_word(double x) {
  return (uint32_t)((unsigned long long)__double_as_longlong(x) >> 32);
}
__device__ __forceinline__ double from_words(uint32_t hi, uint32_t lo) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// sin on [-pi/4, pi/4] of x + y (tail y; iy == 0: y is zero)
__device__ __forceinline__ double kernel_sin(double x, double y, int iy) {
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
               S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
               S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix < 0x3e400000u && (int)x == 0) return x;
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// cos on [-pi/4, pi/4] of x + y
__device__ __forceinline__ double kernel_cos(double x, double y) {
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
               C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
               C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix < 0x3e400000u && (int)x == 0) return 1.0;
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3fd33333u) return 1.0 - (0.5 * z - (z * r - x * y));
  double qx = ix > 0x3fe90000u ? 0.28125 : from_words(ix - 0x00200000u, 0);
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// x - n*pi/2 = y0 + y1 for |x| <= 2^19 * pi/2; returns n.  The high words of
// n*pi/2 (n = 1..32) flag arguments near a multiple, where cancellation calls
// for the longer splits of pi/2.
__device__ __forceinline__ int rem_pio2(double x, double& y0, double& y1) {
  const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
               pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
               pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
               pio2_3t = 8.47842766036889956997e-32;
  uint32_t hx = hi_word(x);
  uint32_t ix = hx & 0x7fffffffu;
  bool neg = (hx >> 31) != 0;
  if (ix <= 0x3fe921fbu) {
    y0 = x;
    y1 = 0;
    return 0;
  }
  if (ix < 0x4002d97cu) {
    if (!neg) {
      double z = x - pio2_1;
      if (ix != 0x3ff921fbu) {
        y0 = z - pio2_1t;
        y1 = (z - y0) - pio2_1t;
      } else {
        z -= pio2_2;
        y0 = z - pio2_2t;
        y1 = (z - y0) - pio2_2t;
      }
      return 1;
    }
    double z = x + pio2_1;
    if (ix != 0x3ff921fbu) {
      y0 = z + pio2_1t;
      y1 = (z - y0) + pio2_1t;
    } else {
      z += pio2_2;
      y0 = z + pio2_2t;
      y1 = (z - y0) + pio2_2t;
    }
    return -1;
  }
  double t = fabs(x);
  int n = (int)(t * invpio2 + 0.5);
  double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  uint32_t nhw = n < 32 ? NPIO2_HW[n - 1] : 0;
  if (n < 32 && ix != nhw) {
    y0 = r - w;
  } else {
    int j = (int)(ix >> 20);
    y0 = r - w;
    int i = j - (int)((hi_word(y0) >> 20) & 0x7ff);
    if (i > 16) {
      t = r;
      w = fn * pio2_2;
      r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      y0 = r - w;
      i = j - (int)((hi_word(y0) >> 20) & 0x7ff);
      if (i > 49) {
        t = r;
        w = fn * pio2_3;
        r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        y0 = r - w;
      }
    }
  }
  y1 = (r - y0) - w;
  if (neg) {
    y0 = -y0;
    y1 = -y1;
    return -n;
  }
  return n;
}

__device__ __forceinline__ double sin(double x) {
  uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix <= 0x3fe921fbu) return kernel_sin(x, 0.0, 0);
  if (ix >= 0x7ff00000u) return x - x;
  double y0, y1;
  int n = rem_pio2(x, y0, y1);
  switch (n & 3) {
    case 0: return kernel_sin(y0, y1, 1);
    case 1: return kernel_cos(y0, y1);
    case 2: return -kernel_sin(y0, y1, 1);
    default: return -kernel_cos(y0, y1);
  }
}

__device__ __forceinline__ double cos(double x) {
  uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix <= 0x3fe921fbu) return kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000u) return x - x;
  double y0, y1;
  int n = rem_pio2(x, y0, y1);
  switch (n & 3) {
    case 0: return kernel_cos(y0, y1);
    case 1: return -kernel_sin(y0, y1, 1);
    case 2: return -kernel_cos(y0, y1);
    default: return kernel_sin(y0, y1, 1);
  }
}

// sin(x) and cos(x) from one argument reduction: per quadrant the same kernel
// calls as sin and cos above, so the same two results bit for bit.
__device__ __forceinline__ void sincos(double x, double& s, double& c) {
  uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix <= 0x3fe921fbu) {
    s = kernel_sin(x, 0.0, 0);
    c = kernel_cos(x, 0.0);
    return;
  }
  if (ix >= 0x7ff00000u) {
    s = c = x - x;
    return;
  }
  double y0, y1;
  int n = rem_pio2(x, y0, y1);
  const double ks = kernel_sin(y0, y1, 1), kc = kernel_cos(y0, y1);
  switch (n & 3) {
    case 0: s = ks; c = kc; break;
    case 1: s = kc; c = -ks; break;
    case 2: s = -ks; c = -kc; break;
    default: s = -kc; c = ks; break;
  }
}

// cube root: a 5-bit estimate from the exponent, a polynomial to 23 bits, a
// rounding to 23 bits away from zero, one Newton step to 53 bits
__device__ __forceinline__ double cbrt(double x) {
  const uint32_t B1 = 715094163u, B2 = 696219795u;
  const double P0 = 1.87595182427177009643, P1 = -1.88497979543377169875,
               P2 = 1.621429720105354466140, P3 = -0.758397934778766047437,
               P4 = 0.145996192886612446982;
  unsigned long long bx = (unsigned long long)__double_as_longlong(x);
  uint32_t hx = (uint32_t)(bx >> 32), low = (uint32_t)bx;
  uint32_t sign = hx & 0x80000000u;
  hx ^= sign;
  if (hx >= 0x7ff00000u) return x + x;
  double t;
  if (hx < 0x00100000u) {
    if ((hx | low) == 0) return x;
    t = from_words(0x43500000u, 0) * x;
    t = from_words(sign | ((hi_word(t) & 0x7fffffffu) / 3 + B2), 0);
  } else {
    t = from_words(sign | (hx / 3 + B1), 0);
  }
  double r = (t * t) * (t / x);
  t = t * ((P0 + r * (P1 + r * P2)) + ((r * r) * r) * (P3 + r * P4));
  unsigned long long bt = (unsigned long long)__double_as_longlong(t);
  bt = (bt + 0x80000000ull) & 0xffffffffc0000000ull;
  t = __longlong_as_double((long long)bt);
  double s = t * t;
  r = x / s;
  double w = t + t;
  r = (r - t) / (w + r);
  return t + t * r;
}

}  // namespace jsm
}  // namespace ct
