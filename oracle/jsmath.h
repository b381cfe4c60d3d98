/*
 * Math.sin, Math.cos and Math.cbrt as the JS engine computes them: V8's
 * ieee754 routines, which restate fdlibm 5.3 (Sun Microsystems; e_rem_pio2.c,
 * k_sin.c, k_cos.c, s_sin.c, s_cos.c) and FreeBSD msun's s_cbrt.c (B. Evans).
 * Written here from those published algorithms; the coefficients are theirs
 * (minimax polynomials, 33/33/53-bit splits of pi/2).  Only the paths the
 * tracer reaches are kept: arguments of sin/cos below 2^19 * pi/2 (it passes
 * 2*pi*u with u in [0, 1)), every finite cbrt argument.
 *
 * Test infrastructure: included by o64.c only.  Pinned bit for bit against
 * Node's Math functions by tests/golden/jsmath_vectors.json.
 */
#ifndef CAUSTIC_JSMATH_H
#define CAUSTIC_JSMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t jsm_hi(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return (uint32_t)(b >> 32);
}
static inline double jsm_from_words(uint32_t hi, uint32_t lo) {
  uint64_t b = ((uint64_t)hi << 32) | lo;
  double x;
  memcpy(&x, &b, 8);
  return x;
}

/* k_sin.c: sin on [-pi/4, pi/4], x + y the argument with tail y (iy = 0: y = 0). */
static inline double jsm_kernel_sin(double x, double y, int iy) {
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
               S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
               S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  uint32_t ix = jsm_hi(x) & 0x7fffffffu;
  if (ix < 0x3e400000u && (int)x == 0) return x; /* |x| < 2^-27 */
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

/* k_cos.c: cos on [-pi/4, pi/4] with tail y. */
static inline double jsm_kernel_cos(double x, double y) {
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
               C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
               C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  uint32_t ix = jsm_hi(x) & 0x7fffffffu;
  if (ix < 0x3e400000u && (int)x == 0) return 1.0; /* |x| < 2^-27 */
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3fd33333u) return 1.0 - (0.5 * z - (z * r - x * y)); /* |x| < 0.3 */
  double qx = ix > 0x3fe90000u ? 0.28125 : jsm_from_words(ix - 0x00200000u, 0); /* x/4 */
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

/* e_rem_pio2.c for |x| <= 2^19 * pi/2: x - n*pi/2 = y[0] + y[1], returns n. */
static inline int jsm_rem_pio2(double x, double *y) {
  static const uint32_t npio2_hw[32] = {
      0x3FF921FBu, 0x400921FBu, 0x4012D97Cu, 0x401921FBu, 0x401F6A7Au, 0x4022D97Cu,
      0x4025FDBBu, 0x402921FBu, 0x402C463Au, 0x402F6A7Au, 0x4031475Cu, 0x4032D97Cu,
      0x40346B9Cu, 0x4035FDBBu, 0x40378FDBu, 0x403921FBu, 0x403AB41Bu, 0x403C463Au,
      0x403DD85Au, 0x403F6A7Au, 0x40407E4Cu, 0x4041475Cu, 0x4042106Cu, 0x4042D97Cu,
      0x4043A28Cu, 0x40446B9Cu, 0x404534ACu, 0x4045FDBBu, 0x4046C6CBu, 0x40478FDBu,
      0x404858EBu, 0x404921FBu};
  const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
               pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
               pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
               pio2_3t = 8.47842766036889956997e-32;
  uint32_t hx = jsm_hi(x);
  uint32_t ix = hx & 0x7fffffffu;
  int neg = (hx >> 31) != 0;
  if (ix <= 0x3fe921fbu) { /* |x| <= pi/4 */
    y[0] = x;
    y[1] = 0;
    return 0;
  }
  if (ix < 0x4002d97cu) { /* |x| < 3pi/4: n = +-1 */
    if (!neg) {
      double z = x - pio2_1;
      if (ix != 0x3ff921fbu) {
        y[0] = z - pio2_1t;
        y[1] = (z - y[0]) - pio2_1t;
      } else { /* near pi/2: 33+33+53 bits of pi */
        z -= pio2_2;
        y[0] = z - pio2_2t;
        y[1] = (z - y[0]) - pio2_2t;
      }
      return 1;
    }
    double z = x + pio2_1;
    if (ix != 0x3ff921fbu) {
      y[0] = z + pio2_1t;
      y[1] = (z - y[0]) + pio2_1t;
    } else {
      z += pio2_2;
      y[0] = z + pio2_2t;
      y[1] = (z - y[0]) + pio2_2t;
    }
    return -1;
  }
  /* medium: |x| <= 2^19 * pi/2 (the only other range the tracer reaches) */
  double t = fabs(x);
  int n = (int)(t * invpio2 + 0.5);
  double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t; /* first round: 85 bits */
  if (n < 32 && ix != npio2_hw[n - 1]) {
    y[0] = r - w; /* no cancellation */
  } else {
    int j = (int)(ix >> 20);
    y[0] = r - w;
    int i = j - (int)((jsm_hi(y[0]) >> 20) & 0x7ff);
    if (i > 16) { /* second iteration: 118 bits */
      t = r;
      w = fn * pio2_2;
      r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      y[0] = r - w;
      i = j - (int)((jsm_hi(y[0]) >> 20) & 0x7ff);
      if (i > 49) { /* third iteration: 151 bits */
        t = r;
        w = fn * pio2_3;
        r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        y[0] = r - w;
      }
    }
  }
  y[1] = (r - y[0]) - w;
  if (neg) {
    y[0] = -y[0];
    y[1] = -y[1];
    return -n;
  }
  return n;
}

/* s_sin.c */
static inline double js_sin(double x) {
  uint32_t ix = jsm_hi(x) & 0x7fffffffu;
  if (ix <= 0x3fe921fbu) return jsm_kernel_sin(x, 0.0, 0);
  if (ix >= 0x7ff00000u) return x - x; /* NaN, +-Inf */
  double y[2];
  int n = jsm_rem_pio2(x, y);
  switch (n & 3) {
    case 0: return jsm_kernel_sin(y[0], y[1], 1);
    case 1: return jsm_kernel_cos(y[0], y[1]);
    case 2: return -jsm_kernel_sin(y[0], y[1], 1);
    default: return -jsm_kernel_cos(y[0], y[1]);
  }
}

/* s_cos.c */
static inline double js_cos(double x) {
  uint32_t ix = jsm_hi(x) & 0x7fffffffu;
  if (ix <= 0x3fe921fbu) return jsm_kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000u) return x - x;
  double y[2];
  int n = jsm_rem_pio2(x, y);
  switch (n & 3) {
    case 0: return jsm_kernel_cos(y[0], y[1]);
    case 1: return -jsm_kernel_sin(y[0], y[1], 1);
    case 2: return -jsm_kernel_cos(y[0], y[1]);
    default: return jsm_kernel_sin(y[0], y[1], 1);
  }
}

/* s_cbrt.c: a 5-bit estimate from the exponent, a polynomial to 23 bits, a
 * rounding to 23 bits away from zero, one Newton step to 53 bits. */
static inline double js_cbrt(double x) {
  const uint32_t B1 = 715094163u, B2 = 696219795u;
  const double P0 = 1.87595182427177009643, P1 = -1.88497979543377169875,
               P2 = 1.621429720105354466140, P3 = -0.758397934778766047437,
               P4 = 0.145996192886612446982;
  uint64_t bx;
  memcpy(&bx, &x, 8);
  uint32_t hx = (uint32_t)(bx >> 32), low = (uint32_t)bx;
  uint32_t sign = hx & 0x80000000u;
  hx ^= sign;
  if (hx >= 0x7ff00000u) return x + x; /* NaN, +-Inf */
  double t;
  if (hx < 0x00100000u) { /* zero or subnormal */
    if ((hx | low) == 0) return x;
    t = jsm_from_words(0x43500000u, 0) * x; /* x * 2^54 */
    t = jsm_from_words(sign | ((jsm_hi(t) & 0x7fffffffu) / 3 + B2), 0);
  } else {
    t = jsm_from_words(sign | (hx / 3 + B1), 0);
  }
  double r = (t * t) * (t / x);
  t = t * ((P0 + r * (P1 + r * P2)) + ((r * r) * r) * (P3 + r * P4));
  uint64_t bt;
  memcpy(&bt, &t, 8);
  bt = (bt + 0x80000000ull) & 0xffffffffc0000000ull;
  memcpy(&t, &bt, 8);
  double s = t * t;
  r = x / s;
  double w = t + t;
  r = (r - t) / (w + r);
  return t + t * r;
}

#endif
