/*
 * O64 — CPU oracle (test infrastructure only; see o64.h for what pins each
 * part).  Each function cites the reference file:line it restates.  All
 * arithmetic is binary64, evaluated left to right exactly as the JS source
 * writes it; every Float32Array store is a (float) conversion.
 */
#include "o64.h"
#include "jsmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PRIM_STRIDE 16
#define MAT_STRIDE 16
#define NODE_STRIDE 8
#define MAX_LEAF_PRIMS 4
#define T_MIN 1e-4
#define ORDERED_STACK 512

static const double PI_D = 3.141592653589793;       /* Math.PI */
#define TWO_PI (2 * PI_D)                            /* materials.ts:11, tracer.ts:37 */
#define INV_PI (1 / PI_D)                            /* materials.ts:10 */
#define SHADOW_SHORTEN (1 - 1e-3)                    /* tracer.ts:36 */

/* Math.sin/cos/cbrt in draws (materials.ts, tracer.ts) are the JS engine's:
 * js_sin/js_cos/js_cbrt (jsmath.h).  The camera frame (camera.ts:25-51, host
 * side on both the oracle and the product) keeps the C library's. */

/* Planted mistakes, for the tests that show the oracle's pins have teeth: a
 * build with -DO64_PLANT=k carries mistake k (the numbers are listed in
 * o64.h); the default build (0) carries none. */
#ifndef O64_PLANT
#define O64_PLANT 0
#endif
#define PLANTED(k) (O64_PLANT == (k))

/* ======================================================================
 * rng.ts
 * ====================================================================== */

/* rng.ts:10-15 — PCG RXS-M-XS over one LCG step, uint32 arithmetic. */
uint32_t o64_pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t shifted = (state >> ((state >> 28) + 4u)) ^ state;
  uint32_t word = shifted * 277803737u;
  return (word >> 22) ^ word;
}

/* rng.ts:18-23 */
uint32_t o64_seed_for(uint32_t pixel, uint32_t frame) {
  uint32_t a = o64_pcg_hash(pixel);
  uint32_t b = o64_pcg_hash(frame ^ 0x9e3779b9u);
  uint32_t s = o64_pcg_hash(a ^ (b * 0x85ebca6bu));
  return s == 0 ? 0x1234567u : s;
}

/* rng.ts:28-31 */
void o64_rng_init(O64Rng *r, uint32_t seed) {
  uint32_t s = o64_pcg_hash(seed);
  r->state = s == 0 ? 0x9e3779b9u : s;
}

/* rng.ts:41-50 — xorshift32 (13, 17, 5). */
uint32_t o64_rng_u32(O64Rng *r) {
  uint32_t x = r->state;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  r->state = x;
  return x;
}

/* rng.ts:52-54 — u32 / 2^32 as a double (exact). */
double o64_rng_next(O64Rng *r) { return (double)o64_rng_u32(r) / 4294967296.0; }

/* rng.ts:57-59 */
double o64_rng_range(O64Rng *r, double lo, double hi) { return lo + (hi - lo) * o64_rng_next(r); }

/* rng.ts:62-64 */
int o64_rng_int(O64Rng *r, int n) {
  double f = floor(o64_rng_next(r) * n);
  double m = (double)(n - 1);
  return (int)(m < f ? m : f);
}

void o64_rng_stream(uint32_t seed, int n, double *out) {
  O64Rng r;
  o64_rng_init(&r, seed);
  for (int i = 0; i < n; i++) out[i] = o64_rng_next(&r);
}

void o64_rng_stream_u32(uint32_t seed, int n, uint32_t *out) {
  O64Rng r;
  o64_rng_init(&r, seed);
  for (int i = 0; i < n; i++) out[i] = o64_rng_u32(&r);
}

/* ======================================================================
 * pack.ts / scene.ts / vec3.ts helpers
 * ====================================================================== */

/* Math.min / Math.max with JS semantics (-0 < +0, NaN propagates). */
static double js_min2(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}
static double js_max2(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}
static double js_min3(double a, double b, double c) { return js_min2(js_min2(a, b), c); }
static double js_max3(double a, double b, double c) { return js_max2(js_max2(a, b), c); }

/* PrimBounds as pack.ts:131-192 computes them; setPrimBounds bvh.ts:46-58. */
static void set_bounds(double *bmin, double *bmax, double *bcen, int i, double x0, double y0,
                       double z0, double x1, double y1, double z1) {
  int o = i * 3;
  bmin[o] = x0; bmin[o + 1] = y0; bmin[o + 2] = z0;
  bmax[o] = x1; bmax[o + 1] = y1; bmax[o + 2] = z1;
  bcen[o] = 0.5 * (x0 + x1);
  bcen[o + 1] = 0.5 * (y0 + y1);
  bcen[o + 2] = 0.5 * (z0 + z1);
}

void o64_prim_bounds(int n, const int32_t *kind, const double *geom, double *bmin, double *bmax,
                     double *bcen) {
  for (int p = 0; p < n; p++) {
    const double *g = geom + 9 * p;
    if (kind[p] == O64_PRIM_SPHERE) {
      double r = g[3];
      double rp = r + 1e-6 + fabs(r) * 1e-6; /* pack.ts:153 */
      set_bounds(bmin, bmax, bcen, p, g[0] - rp, g[1] - rp, g[2] - rp, g[0] + rp, g[1] + rp,
                 g[2] + rp);
    } else {
      const double pad = 1e-6; /* pack.ts:177 */
      set_bounds(bmin, bmax, bcen, p, js_min3(g[0], g[3], g[6]) - pad,
                 js_min3(g[1], g[4], g[7]) - pad, js_min3(g[2], g[5], g[8]) - pad,
                 js_max3(g[0], g[3], g[6]) + pad, js_max3(g[1], g[4], g[7]) + pad,
                 js_max3(g[2], g[5], g[8]) + pad);
    }
  }
}

/* pack.ts:131-192 — unordered prim records. scene.ts:120-127 normal/area. */
static void write_prims(int n, const int32_t *kind, const double *geom, const int32_t *obj,
                        float *prims) {
  for (int p = 0; p < n; p++) {
    const double *g = geom + 9 * p;
    float *o = prims + PRIM_STRIDE * p;
    memset(o, 0, sizeof(float) * PRIM_STRIDE);
    if (kind[p] == O64_PRIM_SPHERE) {
      double r = g[3];
      o[0] = (float)g[0];
      o[1] = (float)g[1];
      o[2] = (float)g[2];
      o[3] = (float)O64_PRIM_SPHERE;
      o[4] = (float)r;
      o[7] = (float)obj[p];
      o[11] = (float)(4 * PI_D * r * r);
    } else {
      double e1x = g[3] - g[0], e1y = g[4] - g[1], e1z = g[5] - g[2];
      double e2x = g[6] - g[0], e2y = g[7] - g[1], e2z = g[8] - g[2];
      /* vec3.ts:22 cross */
      double cx = e1y * e2z - e1z * e2y;
      double cy = e1z * e2x - e1x * e2z;
      double cz = e1x * e2y - e1y * e2x;
      /* vec3.ts:28-31 normalize: scale by 1/len */
      double len = sqrt(cx * cx + cy * cy + cz * cz);
      double nx = 0, ny = 0, nz = 0;
      if (len > 0) {
        double s = 1 / len;
        nx = cx * s;
        ny = cy * s;
        nz = cz * s;
      }
      /* scene.ts:124-127 area */
      double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
      o[0] = (float)g[0]; o[1] = (float)g[1]; o[2] = (float)g[2];
      o[3] = (float)O64_PRIM_TRIANGLE;
      o[4] = (float)g[3]; o[5] = (float)g[4]; o[6] = (float)g[5];
      o[7] = (float)obj[p];
      o[8] = (float)g[6]; o[9] = (float)g[7]; o[10] = (float)g[8];
      o[11] = (float)area;
      o[12] = (float)nx; o[13] = (float)ny; o[14] = (float)nz;
    }
  }
}

/* ======================================================================
 * bvh.ts — binned SAH build
 * ====================================================================== */

typedef struct {
  const double *min, *max, *cen;
  uint32_t *order;
  float *nodes;
  uint32_t *words; /* exact leaf words, or NULL */
  int nodeCount;
  int bins;
  /* scratch (bvh.ts:79-83) */
  int32_t *bucketCount;
  double *bucketMin, *bucketMax, *rightArea;
  int32_t *rightCount;
  uint32_t *sortTmp;
} Builder;

static double surface_area(double dx, double dy, double dz) {
  return 2 * (dx * dy + dy * dz + dz * dx); /* bvh.ts:70 */
}

/* bvh.ts:168-281 */
static int sah_split(Builder *B, int start, int end, double cMinX, double cMinY, double cMinZ,
                     double cMaxX, double cMaxY, double cMaxZ) {
  const int NB = B->bins;
  const double *min = B->min, *max = B->max, *centroid = B->cen;
  uint32_t *order = B->order;
  double bestCost = INFINITY;
  int bestAxis = -1, bestSplit = -1;
  double bestOrigin = 0, bestScale = 0;
  for (int axisStep = 0; axisStep < 3; axisStep++) {
    int axis = PLANTED(O64_PLANT_AXIS_REVERSED) ? 2 - axisStep : axisStep;
    double cmin = axis == 0 ? cMinX : axis == 1 ? cMinY : cMinZ;
    double cmax = axis == 0 ? cMaxX : axis == 1 ? cMaxY : cMaxZ;
    double extent = cmax - cmin;
    if (!(extent > 1e-12)) continue;
    double toBucket = NB / extent;
    for (int b = 0; b < NB; b++) {
      B->bucketCount[b] = 0;
      for (int k = 0; k < 3; k++) {
        B->bucketMin[b * 3 + k] = INFINITY;
        B->bucketMax[b * 3 + k] = -INFINITY;
      }
    }
    for (int i = start; i < end; i++) {
      int o = (int)order[i] * 3;
      double fb = floor((centroid[o + axis] - cmin) * toBucket);
      int b;
      if (fb >= NB) b = NB - 1;
      else if (fb < 0) b = 0;
      else b = (int)fb;
      B->bucketCount[b]++;
      int bo = b * 3;
      for (int k = 0; k < 3; k++) {
        if (min[o + k] < B->bucketMin[bo + k]) B->bucketMin[bo + k] = min[o + k];
        if (max[o + k] > B->bucketMax[bo + k]) B->bucketMax[bo + k] = max[o + k];
      }
    }
    double rx0 = INFINITY, ry0 = INFINITY, rz0 = INFINITY;
    double rx1 = -INFINITY, ry1 = -INFINITY, rz1 = -INFINITY;
    int rc = 0;
    for (int s = NB - 1; s >= 1; s--) {
      int bo = s * 3;
      if (B->bucketCount[s] > 0) {
        if (B->bucketMin[bo] < rx0) rx0 = B->bucketMin[bo];
        if (B->bucketMin[bo + 1] < ry0) ry0 = B->bucketMin[bo + 1];
        if (B->bucketMin[bo + 2] < rz0) rz0 = B->bucketMin[bo + 2];
        if (B->bucketMax[bo] > rx1) rx1 = B->bucketMax[bo];
        if (B->bucketMax[bo + 1] > ry1) ry1 = B->bucketMax[bo + 1];
        if (B->bucketMax[bo + 2] > rz1) rz1 = B->bucketMax[bo + 2];
        rc += B->bucketCount[s];
      }
      int slot = PLANTED(O64_PLANT_SWEEP_SHIFT) ? (s < NB - 1 ? s : s - 1) : s - 1;
      B->rightCount[slot] = rc;
      B->rightArea[slot] = rc > 0 ? surface_area(rx1 - rx0, ry1 - ry0, rz1 - rz0) : 0;
    }
    double lx0 = INFINITY, ly0 = INFINITY, lz0 = INFINITY;
    double lx1 = -INFINITY, ly1 = -INFINITY, lz1 = -INFINITY;
    int lc = 0;
    for (int s = 0; s < NB - 1; s++) {
      int bo = s * 3;
      if (B->bucketCount[s] > 0) {
        if (B->bucketMin[bo] < lx0) lx0 = B->bucketMin[bo];
        if (B->bucketMin[bo + 1] < ly0) ly0 = B->bucketMin[bo + 1];
        if (B->bucketMin[bo + 2] < lz0) lz0 = B->bucketMin[bo + 2];
        if (B->bucketMax[bo] > lx1) lx1 = B->bucketMax[bo];
        if (B->bucketMax[bo + 1] > ly1) ly1 = B->bucketMax[bo + 1];
        if (B->bucketMax[bo + 2] > lz1) lz1 = B->bucketMax[bo + 2];
        lc += B->bucketCount[s];
      }
      int rcs = B->rightCount[s];
      if (lc == 0 || rcs == 0) continue;
      double cost = lc * surface_area(lx1 - lx0, ly1 - ly0, lz1 - lz0) + rcs * B->rightArea[s];
      if (PLANTED(O64_PLANT_SAH_LE) ? cost <= bestCost : cost < bestCost) {
        bestCost = cost;
        bestAxis = axis;
        bestSplit = s;
        bestOrigin = cmin;
        bestScale = toBucket;
      }
    }
  }
  if (bestAxis < 0) return -1;
  /* bvh.ts:262-278 in-place two-pointer partition */
  int lo = start, hi = end - 1;
  while (lo <= hi) {
    int o = (int)order[lo] * 3;
    double fb = floor((centroid[o + bestAxis] - bestOrigin) * bestScale);
    int b;
    if (fb >= NB) b = NB - 1;
    else if (fb < 0) b = 0;
    else b = (int)fb;
    if (PLANTED(O64_PLANT_PARTITION_LT) ? b < bestSplit : b <= bestSplit) {
      lo++;
    } else {
      uint32_t tmp = order[lo];
      order[lo] = order[hi];
      order[hi] = tmp;
      hi--;
    }
  }
  if (lo == start || lo == end) return -1;
  return lo;
}

/* Stable merge sort of order[start,end) by centroid[axis] (V8's sort is stable). */
static void stable_sort_range(Builder *B, int start, int end, int axis) {
  int n = end - start;
  uint32_t *a = B->order + start;
  uint32_t *tmp = B->sortTmp;
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      int mid = lo + width < n ? lo + width : n;
      int hi = lo + 2 * width < n ? lo + 2 * width : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        double ki = B->cen[a[i] * 3 + axis], kj = B->cen[a[j] * 3 + axis];
        /* comparator a-b: take right only when strictly smaller */
        if (PLANTED(O64_PLANT_MEDIAN_UNSTABLE) ? kj - ki <= 0 : kj - ki < 0) tmp[k++] = a[j++];
        else tmp[k++] = a[i++];
      }
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    memcpy(a, tmp, sizeof(uint32_t) * n);
  }
}

/* bvh.ts:284-296 */
static int median_split(Builder *B, int start, int end, double cMinX, double cMinY, double cMinZ,
                        double cMaxX, double cMaxY, double cMaxZ) {
  double ex = cMaxX - cMinX, ey = cMaxY - cMinY, ez = cMaxZ - cMinZ;
  int axis = ex >= ey && ex >= ez ? 0 : ey >= ez ? 1 : 2;
  if (PLANTED(O64_PLANT_MEDIAN_AXIS_LAST)) axis = ez >= ey && ez >= ex ? 2 : ey >= ex ? 1 : 0;
  stable_sort_range(B, start, end, axis);
  return (start + end) >> 1;
}

/* bvh.ts:111-165 */
static void build_node(Builder *B, int start, int end) {
  int nodeIndex = B->nodeCount++;
  float *nd = B->nodes + (size_t)nodeIndex * NODE_STRIDE;
  double bMinX = INFINITY, bMinY = INFINITY, bMinZ = INFINITY;
  double bMaxX = -INFINITY, bMaxY = -INFINITY, bMaxZ = -INFINITY;
  double cMinX = INFINITY, cMinY = INFINITY, cMinZ = INFINITY;
  double cMaxX = -INFINITY, cMaxY = -INFINITY, cMaxZ = -INFINITY;
  for (int i = start; i < end; i++) {
    int o = (int)B->order[i] * 3;
    double x0 = B->min[o], y0 = B->min[o + 1], z0 = B->min[o + 2];
    double x1 = B->max[o], y1 = B->max[o + 1], z1 = B->max[o + 2];
    if (x0 < bMinX) bMinX = x0;
    if (y0 < bMinY) bMinY = y0;
    if (z0 < bMinZ) bMinZ = z0;
    if (x1 > bMaxX) bMaxX = x1;
    if (y1 > bMaxY) bMaxY = y1;
    if (z1 > bMaxZ) bMaxZ = z1;
    double cx = B->cen[o], cy = B->cen[o + 1], cz = B->cen[o + 2];
    if (cx < cMinX) cMinX = cx;
    if (cy < cMinY) cMinY = cy;
    if (cz < cMinZ) cMinZ = cz;
    if (cx > cMaxX) cMaxX = cx;
    if (cy > cMaxY) cMaxY = cy;
    if (cz > cMaxZ) cMaxZ = cz;
  }
  const double nudgeSign = PLANTED(O64_PLANT_NUDGE_INWARD) ? -1 : 1;
  nd[0] = (float)(bMinX - nudgeSign * (fabs(bMinX) * 1.2e-7) - nudgeSign * 1e-9);
  nd[1] = (float)(bMinY - nudgeSign * (fabs(bMinY) * 1.2e-7) - nudgeSign * 1e-9);
  nd[2] = (float)(bMinZ - nudgeSign * (fabs(bMinZ) * 1.2e-7) - nudgeSign * 1e-9);
  nd[4] = (float)(bMaxX + nudgeSign * (fabs(bMaxX) * 1.2e-7) + nudgeSign * 1e-9);
  nd[5] = (float)(bMaxY + nudgeSign * (fabs(bMaxY) * 1.2e-7) + nudgeSign * 1e-9);
  nd[6] = (float)(bMaxZ + nudgeSign * (fabs(bMaxZ) * 1.2e-7) + nudgeSign * 1e-9);
  int count = end - start;
  if (count <= (PLANTED(O64_PLANT_LEAF3) ? 3 : MAX_LEAF_PRIMS)) {
    nd[3] = (float)(nodeIndex + 1);
    nd[7] = (float)((double)start * 8 + count); /* encodeLeaf bvh.ts:60-62 */
    if (B->words) B->words[nodeIndex] = (uint32_t)start * 8u + (uint32_t)count;
    return;
  }
  int mid = sah_split(B, start, end, cMinX, cMinY, cMinZ, cMaxX, cMaxY, cMaxZ);
  if (mid <= start || mid >= end)
    mid = median_split(B, start, end, cMinX, cMinY, cMinZ, cMaxX, cMaxY, cMaxZ);
  nd[7] = -1;
  if (B->words) B->words[nodeIndex] = 0xffffffffu;
  build_node(B, start, mid);
  build_node(B, mid, end);
  B->nodes[(size_t)nodeIndex * NODE_STRIDE + 3] = (float)B->nodeCount;
}

/* bvh.ts:92-109, 299-301 */
int o64_build_bvh_ex(int n, const double *bmin, const double *bmax, const double *bcen, int bins,
                     float *nodes, uint32_t *order, uint32_t *leafWords) {
  Builder B;
  memset(&B, 0, sizeof B);
  B.min = bmin; B.max = bmax; B.cen = bcen;
  B.order = order;
  B.nodes = nodes;
  B.words = leafWords;
  B.bins = bins;
  if (n == 0 && leafWords) leafWords[0] = 0;
  int maxNodes = n > 0 ? 2 * n - 1 : 1;
  memset(nodes, 0, sizeof(float) * (size_t)maxNodes * NODE_STRIDE);
  for (int i = 0; i < n; i++) order[i] = (uint32_t)i;
  if (n == 0) {
    nodes[3] = 1;
    nodes[7] = 0;
    return 1;
  }
  B.bucketCount = (int32_t *)calloc((size_t)bins, sizeof(int32_t));
  B.rightCount = (int32_t *)calloc((size_t)bins, sizeof(int32_t));
  B.bucketMin = (double *)calloc((size_t)bins * 3, sizeof(double));
  B.bucketMax = (double *)calloc((size_t)bins * 3, sizeof(double));
  B.rightArea = (double *)calloc((size_t)bins, sizeof(double));
  B.sortTmp = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
  build_node(&B, 0, n);
  free(B.bucketCount); free(B.rightCount); free(B.bucketMin); free(B.bucketMax);
  free(B.rightArea); free(B.sortTmp);
  return B.nodeCount;
}

int o64_build_bvh(int n, const double *bmin, const double *bmax, const double *bcen, int bins,
                  float *nodes, uint32_t *order) {
  return o64_build_bvh_ex(n, bmin, bmax, bcen, bins, nodes, order, NULL);
}

int o64_partition_flags(int n, const uint8_t *good, int32_t *perm) {
  /* perm[i] = original index found at position i after bvh.ts:262-278 */
  for (int i = 0; i < n; i++) perm[i] = i;
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    if (good[perm[lo]]) {
      lo++;
    } else {
      int32_t t = perm[lo];
      perm[lo] = perm[hi];
      perm[hi] = t;
      hi--;
    }
  }
  return lo;
}

/* pack.ts:195-223; leafWords (or NULL) receives the exact leaf words. */
int o64_pack_scene_ex(int n, const int32_t *kind, const double *geom, const int32_t *obj,
                      const uint8_t *emits, int bins, float *prims, uint32_t *primObject,
                      float *nodes, uint32_t *order, int32_t *lights, int *lightCount,
                      uint32_t *leafWords) {
  int cap = n > 0 ? n : 1;
  float *raw = (float *)calloc((size_t)cap * PRIM_STRIDE, sizeof(float));
  double *bmin = (double *)calloc((size_t)cap * 3, sizeof(double));
  double *bmax = (double *)calloc((size_t)cap * 3, sizeof(double));
  double *bcen = (double *)calloc((size_t)cap * 3, sizeof(double));
  write_prims(n, kind, geom, obj, raw);
  o64_prim_bounds(n, kind, geom, bmin, bmax, bcen);
  int nodeCount = o64_build_bvh_ex(n, bmin, bmax, bcen, bins, nodes, order, leafWords);
  memset(prims, 0, sizeof(float) * (size_t)cap * PRIM_STRIDE);
  memset(primObject, 0, sizeof(uint32_t) * (size_t)cap);
  for (int i = 0; i < n; i++) {
    uint32_t src = order[i];
    memcpy(prims + (size_t)i * PRIM_STRIDE, raw + (size_t)src * PRIM_STRIDE,
           sizeof(float) * PRIM_STRIDE);
    primObject[i] = (uint32_t)obj[src];
  }
  int nl = 0;
  for (int i = 0; i < n; i++)
    if (emits[primObject[i]]) lights[nl++] = i;
  *lightCount = nl;
  free(raw); free(bmin); free(bmax); free(bcen);
  return nodeCount;
}

int o64_pack_scene(int n, const int32_t *kind, const double *geom, const int32_t *obj,
                   const uint8_t *emits, int bins, float *prims, uint32_t *primObject,
                   float *nodes, uint32_t *order, int32_t *lights, int *lightCount) {
  return o64_pack_scene_ex(n, kind, geom, obj, emits, bins, prims, primObject, nodes, order,
                           lights, lightCount, NULL);
}

/* bvh.ts:319-381 */
typedef struct {
  const float *n;
  const uint32_t *words; /* exact leaf words, or NULL */
  int nodeCount;
  const uint32_t *order;
  int primCount;
  uint8_t *covered;
  int errors, maxLeaf, leafCount, depth;
} VState;

static int vwalk(VState *S, int index, int d) {
  if (d > S->depth) S->depth = d;
  if (index >= S->nodeCount) {
    S->errors++;
    return index + 1;
  }
  const float *nb = S->n + (size_t)index * NODE_STRIDE;
  double skip = nb[3];
  double leaf = S->words ? (S->words[index] == 0xffffffffu ? -1.0 : (double)S->words[index])
                         : nb[7];
  if (leaf >= 0) {
    double start = floor(leaf / 8);
    double count = leaf - floor(leaf / 8) * 8;
    S->leafCount++;
    if (count > S->maxLeaf) S->maxLeaf = (int)count;
    if (count > MAX_LEAF_PRIMS) S->errors++;
    for (double i = start; i < start + count; i++) {
      if (i >= S->primCount) {
        S->errors++;
        continue;
      }
      uint32_t orig = S->order[(int)i];
      if (S->covered[orig]) S->errors++;
      S->covered[orig] = 1;
    }
    if (skip != index + 1) S->errors++;
    return index + 1;
  }
  int afterLeft = vwalk(S, index + 1, d + 1);
  int afterRight = vwalk(S, afterLeft, d + 1);
  if (skip != afterRight) S->errors++;
  int children[2] = {index + 1, afterLeft};
  for (int c = 0; c < 2; c++) {
    const float *cb = S->n + (size_t)children[c] * NODE_STRIDE;
    if (children[c] >= S->nodeCount) continue;
    for (int k = 0; k < 3; k++) {
      if ((double)cb[k] < (double)nb[k] - 1e-6 || (double)cb[4 + k] > (double)nb[4 + k] + 1e-6) {
        S->errors++;
        break;
      }
    }
  }
  return afterRight;
}

int o64_validate_bvh_ex(const float *nodes, int nodeCount, const uint32_t *order, int primCount,
                        const uint32_t *leafWords, int32_t *stats) {
  VState S;
  memset(&S, 0, sizeof S);
  S.n = nodes;
  S.words = leafWords;
  S.nodeCount = nodeCount;
  S.order = order;
  S.primCount = primCount;
  S.covered = (uint8_t *)calloc((size_t)(primCount > 0 ? primCount : 1), 1);
  int end = vwalk(&S, 0, 1);
  if (end != nodeCount) S.errors++;
  for (int i = 0; i < primCount; i++)
    if (!S.covered[i]) {
      S.errors++;
      break;
    }
  free(S.covered);
  if (stats) {
    stats[0] = S.maxLeaf;
    stats[1] = S.leafCount;
    stats[2] = S.depth;
  }
  return S.errors;
}

int o64_validate_bvh(const float *nodes, int nodeCount, const uint32_t *order, int primCount,
                     int32_t *stats) {
  return o64_validate_bvh_ex(nodes, nodeCount, order, primCount, NULL, stats);
}

/* ======================================================================
 * intersect.ts
 * ====================================================================== */

/* intersect.ts:20-39 */
double o64_ray_sphere(double ox, double oy, double oz, double dx, double dy, double dz,
                      double cx, double cy, double cz, double r, double tMin, double tMax) {
  double lx = ox - cx, ly = oy - cy, lz = oz - cz;
  double a = dx * dx + dy * dy + dz * dz;
  double halfB = lx * dx + ly * dy + lz * dz;
  double c = lx * lx + ly * ly + lz * lz - r * r;
  double disc = halfB * halfB - a * c;
  if (disc < 0) return -1;
  double s = sqrt(disc);
  double t = (-halfB - s) / a;
  if (t <= tMin || t >= tMax) {
    t = (-halfB + s) / a;
    if (t <= tMin || t >= tMax) return -1;
  }
  return t;
}

/* intersect.ts:52-85 */
int o64_ray_triangle(double ox, double oy, double oz, double dx, double dy, double dz, double ax,
                     double ay, double az, double bx, double by, double bz, double cx, double cy,
                     double cz, double tMin, double tMax, double *out3) {
  double e1x = bx - ax, e1y = by - ay, e1z = bz - az;
  double e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
  double px = dy * e2z - dz * e2y;
  double py = dz * e2x - dx * e2z;
  double pz = dx * e2y - dy * e2x;
  double det = e1x * px + e1y * py + e1z * pz;
  if (det > -1e-12 && det < 1e-12) return 0;
  double invDet = 1 / det;
  double tx = ox - ax, ty = oy - ay, tz = oz - az;
  double u = (tx * px + ty * py + tz * pz) * invDet;
  if (u < 0 || u > 1) return 0;
  double qx = ty * e1z - tz * e1y;
  double qy = tz * e1x - tx * e1z;
  double qz = tx * e1y - ty * e1x;
  double v = (dx * qx + dy * qy + dz * qz) * invDet;
  if (v < 0 || u + v > 1) return 0;
  double t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
  if (t <= tMin || t >= tMax) return 0;
  out3[0] = t;
  out3[1] = u;
  out3[2] = v;
  return 1;
}

/* intersect.ts:92-123 */
double o64_ray_aabb(double ox, double oy, double oz, double idx, double idy, double idz,
                    double minX, double minY, double minZ, double maxX, double maxY, double maxZ,
                    double tMax) {
  double t0 = (minX - ox) * idx;
  double t1 = (maxX - ox) * idx;
  double tNear = t0 < t1 ? t0 : t1;
  double tFar = t0 < t1 ? t1 : t0;
  t0 = (minY - oy) * idy;
  t1 = (maxY - oy) * idy;
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0;
    if (t1 < tFar) tFar = t1;
  } else {
    if (t1 > tNear) tNear = t1;
    if (t0 < tFar) tFar = t0;
  }
  t0 = (minZ - oz) * idz;
  t1 = (maxZ - oz) * idz;
  if (t0 < t1) {
    if (t0 > tNear) tNear = t0;
    if (t1 < tFar) tFar = t1;
  } else {
    if (t1 > tNear) tNear = t1;
    if (t0 < tFar) tFar = t0;
  }
  if (tNear > tFar || tFar < 0 || tNear > tMax) return INFINITY;
  return tNear < 0 ? 0 : tNear;
}

/* intersect.ts:142 */
static double safe_inv(double d) { return 1 / (d == 0 ? 1e-20 : d); }

#define NODEF(i, k) ((double)nodes[(size_t)(i) * NODE_STRIDE + (k)])
#define PRIMF(p, k) ((double)prims[(size_t)(p) * PRIM_STRIDE + (k)])

/* A node's leaf word: the float32 one of the reference layout, or the exact
 * one of the layout extension (interior nodes: -1 either way). */
static double node_leaf(const O64Scene *s, int node) {
  if (s->leafWords)
    return s->leafWords[node] == 0xffffffffu ? -1.0 : (double)s->leafWords[node];
  return (double)s->nodes[(size_t)node * NODE_STRIDE + 7];
}

/* Shared leaf test: intersect.ts:155-193 (ordered) and 328-350 (stackless). */
static double test_leaf(const float *prims, double leaf, const double *o, const double *d,
                        double tMin, double closest, int32_t *bestPrim, double *bestU,
                        double *bestV, O64Counts *cnt) {
  int li = (int)leaf; /* (leaf | 0): leaf words are small non-negative integers */
  int count = li & 7;
  int start = li >> 3;
  double tri[3];
  for (int p = start; p < start + count; p++) {
    if (PRIMF(p, 3) == O64_PRIM_SPHERE) {
      if (cnt) cnt->spheres++;
      double t = o64_ray_sphere(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                                PRIMF(p, 2), PRIMF(p, 4), tMin, closest);
      if (t > 0) {
        closest = t;
        *bestPrim = p;
      }
    } else {
      if (cnt) cnt->tris++;
      if (o64_ray_triangle(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                           PRIMF(p, 2), PRIMF(p, 4), PRIMF(p, 5), PRIMF(p, 6), PRIMF(p, 8),
                           PRIMF(p, 9), PRIMF(p, 10), tMin, closest, tri)) {
        closest = tri[0];
        *bestPrim = p;
        *bestU = tri[1];
        *bestV = tri[2];
      }
    }
  }
  return closest;
}

static void set_miss(O64Hit *h) {
  h->t = INFINITY;
  h->prim = -1;
}

/* intersect.ts:202-282 — ordered front-to-back walk with a stack. */
int o64_intersect_closest(const O64Scene *s, const double *o, const double *d, double tMin,
                          double tMax, O64Hit *hit, O64Counts *cnt) {
  const float *nodes = s->nodes, *prims = s->prims;
  if (s->nodeCount == 0) {
    set_miss(hit);
    return 0;
  }
  double idx = safe_inv(d[0]), idy = safe_inv(d[1]), idz = safe_inv(d[2]);
  double closest = tMax;
  int32_t bestPrim = -1;
  double bestU = 0, bestV = 0;
  if (cnt) cnt->nodes++;
  if (o64_ray_aabb(o[0], o[1], o[2], idx, idy, idz, NODEF(0, 0), NODEF(0, 1), NODEF(0, 2),
                   NODEF(0, 4), NODEF(0, 5), NODEF(0, 6), closest) == INFINITY) {
    set_miss(hit);
    return 0;
  }
  int32_t stackNode[ORDERED_STACK];
  double stackEntry[ORDERED_STACK];
  int sp = 0;
  int node = 0;
  for (;;) {
    double leaf = node_leaf(s, node);
    if (leaf >= 0) {
      if (cnt) cnt->leaves++;
      closest = test_leaf(prims, leaf, o, d, tMin, closest, &bestPrim, &bestU, &bestV, cnt);
    } else {
      int left = node + 1;
      int right = (int)NODEF(left, 3);
      if (cnt) cnt->nodes += 2;
      double tl = o64_ray_aabb(o[0], o[1], o[2], idx, idy, idz, NODEF(left, 0), NODEF(left, 1),
                               NODEF(left, 2), NODEF(left, 4), NODEF(left, 5), NODEF(left, 6),
                               closest);
      double tr = o64_ray_aabb(o[0], o[1], o[2], idx, idy, idz, NODEF(right, 0),
                               NODEF(right, 1), NODEF(right, 2), NODEF(right, 4),
                               NODEF(right, 5), NODEF(right, 6), closest);
      if (tl != INFINITY) {
        if (tr != INFINITY) {
          if (sp >= ORDERED_STACK) abort();
          if (tl <= tr) {
            stackNode[sp] = right;
            stackEntry[sp++] = tr;
            node = left;
          } else {
            stackNode[sp] = left;
            stackEntry[sp++] = tl;
            node = right;
          }
        } else {
          node = left;
        }
        continue;
      }
      if (tr != INFINITY) {
        node = right;
        continue;
      }
    }
    do {
      if (sp == 0) {
        if (bestPrim < 0) {
          set_miss(hit);
          return 0;
        }
        hit->t = closest;
        hit->prim = bestPrim;
        hit->u = bestU;
        hit->v = bestV;
        return 1;
      }
      sp--;
    } while (stackEntry[sp] >= closest);
    node = stackNode[sp];
  }
}

/* intersect.ts:290-365 — stackless skip-link walk. */
int o64_intersect_stackless(const O64Scene *s, const double *o, const double *d, double tMin,
                            double tMax, O64Hit *hit, O64Counts *cnt) {
  const float *nodes = s->nodes, *prims = s->prims;
  int nodeCount = s->nodeCount;
  double idx = safe_inv(d[0]), idy = safe_inv(d[1]), idz = safe_inv(d[2]);
  double closest = tMax;
  int32_t bestPrim = -1;
  double bestU = 0, bestV = 0;
  int node = 0;
  while (node < nodeCount) {
    if (cnt) cnt->nodes++;
    double entry = o64_ray_aabb(o[0], o[1], o[2], idx, idy, idz, NODEF(node, 0), NODEF(node, 1),
                                NODEF(node, 2), NODEF(node, 4), NODEF(node, 5), NODEF(node, 6),
                                closest);
    if (entry == INFINITY) {
      node = (int)NODEF(node, 3);
      continue;
    }
    double leaf = node_leaf(s, node);
    if (leaf >= 0) {
      if (cnt) cnt->leaves++;
      closest = test_leaf(prims, leaf, o, d, tMin, closest, &bestPrim, &bestU, &bestV, cnt);
    }
    node++;
  }
  if (bestPrim < 0) {
    set_miss(hit);
    return 0;
  }
  hit->t = closest;
  hit->prim = bestPrim;
  hit->u = bestU;
  hit->v = bestV;
  return 1;
}

/* intersect.ts:368-420 — stackless any-hit. */
int o64_occluded(const O64Scene *s, const double *o, const double *d, double tMin, double tMax,
                 O64Counts *cnt) {
  const float *nodes = s->nodes, *prims = s->prims;
  int nodeCount = s->nodeCount;
  double idx = safe_inv(d[0]), idy = safe_inv(d[1]), idz = safe_inv(d[2]);
  double tri[3];
  int node = 0;
  while (node < nodeCount) {
    if (cnt) cnt->nodes++;
    double entry = o64_ray_aabb(o[0], o[1], o[2], idx, idy, idz, NODEF(node, 0), NODEF(node, 1),
                                NODEF(node, 2), NODEF(node, 4), NODEF(node, 5), NODEF(node, 6),
                                tMax);
    if (entry == INFINITY) {
      node = (int)NODEF(node, 3);
      continue;
    }
    double leaf = node_leaf(s, node);
    if (leaf >= 0) {
      if (cnt) cnt->leaves++;
      int li = (int)leaf;
      int count = li & 7;
      int start = li >> 3;
      for (int p = start; p < start + count; p++) {
        if (PRIMF(p, 3) == O64_PRIM_SPHERE) {
          if (cnt) cnt->spheres++;
          if (o64_ray_sphere(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                             PRIMF(p, 2), PRIMF(p, 4), tMin, tMax) > 0)
            return 1;
        } else {
          if (cnt) cnt->tris++;
          if (o64_ray_triangle(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                               PRIMF(p, 2), PRIMF(p, 4), PRIMF(p, 5), PRIMF(p, 6), PRIMF(p, 8),
                               PRIMF(p, 9), PRIMF(p, 10), tMin, tMax, tri))
            return 1;
        }
      }
    }
    node++;
  }
  return 0;
}

/* intersect.ts:423-468 */
int o64_intersect_brute(const O64Scene *s, const double *o, const double *d, double tMin,
                        double tMax, O64Hit *hit) {
  const float *prims = s->prims;
  double closest = tMax;
  int32_t bestPrim = -1;
  double bestU = 0, bestV = 0, tri[3];
  for (int p = 0; p < s->primCount; p++) {
    if (PRIMF(p, 3) == O64_PRIM_SPHERE) {
      double t = o64_ray_sphere(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                                PRIMF(p, 2), PRIMF(p, 4), tMin, closest);
      if (t > 0) {
        closest = t;
        bestPrim = p;
      }
    } else if (o64_ray_triangle(o[0], o[1], o[2], d[0], d[1], d[2], PRIMF(p, 0), PRIMF(p, 1),
                                PRIMF(p, 2), PRIMF(p, 4), PRIMF(p, 5), PRIMF(p, 6),
                                PRIMF(p, 8), PRIMF(p, 9), PRIMF(p, 10), tMin, closest, tri)) {
      closest = tri[0];
      bestPrim = p;
      bestU = tri[1];
      bestV = tri[2];
    }
  }
  hit->t = bestPrim < 0 ? INFINITY : closest;
  hit->prim = bestPrim;
  hit->u = bestU;
  hit->v = bestV;
  return bestPrim >= 0;
}

void o64_closest_batch(const O64Scene *s, int n, const double *rays, double tMin,
                       const double *tMax, int walk, int32_t *prim, double *t, double *u,
                       double *v, int64_t *counts) {
  O64Counts c = {0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    O64Hit h = {INFINITY, -1, 0, 0};
    const double *r = rays + 6 * (size_t)i;
    if (walk == 0) o64_intersect_closest(s, r, r + 3, tMin, tMax[i], &h, &c);
    else if (walk == 1) o64_intersect_stackless(s, r, r + 3, tMin, tMax[i], &h, &c);
    else o64_intersect_brute(s, r, r + 3, tMin, tMax[i], &h);
    prim[i] = h.prim;
    t[i] = h.t;
    u[i] = h.u;
    v[i] = h.v;
  }
  if (counts) {
    counts[0] = c.nodes;
    counts[1] = c.tris;
    counts[2] = c.spheres;
    counts[3] = c.leaves;
  }
}

void o64_occluded_batch(const O64Scene *s, int n, const double *rays, double tMin,
                        const double *tMax, uint8_t *out, int64_t *counts) {
  O64Counts c = {0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    const double *r = rays + 6 * (size_t)i;
    out[i] = (uint8_t)o64_occluded(s, r, r + 3, tMin, tMax[i], &c);
  }
  if (counts) {
    counts[0] = c.nodes;
    counts[1] = c.tris;
    counts[2] = c.spheres;
    counts[3] = c.leaves;
  }
}

/* ======================================================================
 * materials.ts
 * ====================================================================== */

/* materials.ts:20-42 */
void o64_basis(double nx, double ny, double nz, double *out) {
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
  out[0] = tx;
  out[1] = ty;
  out[2] = tz;
  out[3] = ny * tz - nz * ty;
  out[4] = nz * tx - nx * tz;
  out[5] = nx * ty - ny * tx;
}

/* materials.ts:47-64 */
void o64_cosine_hemisphere(double nx, double ny, double nz, double u1, double u2, double *out) {
  double r = sqrt(u1);
  double phi = TWO_PI * u2;
  double x = r * js_cos(phi);
  double y = r * js_sin(phi);
  double z = sqrt(u1 < 1 ? 1 - u1 : 0);
  double b[6];
  o64_basis(nx, ny, nz, b);
  double dx = b[0] * x + b[3] * y + nx * z;
  double dy = b[1] * x + b[4] * y + ny * z;
  double dz = b[2] * x + b[5] * y + nz * z;
  double inv = 1 / sqrt(dx * dx + dy * dy + dz * dz);
  out[0] = dx * inv;
  out[1] = dy * inv;
  out[2] = dz * inv;
}

/* materials.ts:67-74 */
void o64_unit_vector(double u1, double u2, double *out) {
  double z = 1 - 2 * u1;
  double r = sqrt(z * z < 1 ? 1 - z * z : 0);
  double phi = TWO_PI * u2;
  out[0] = r * js_cos(phi);
  out[1] = r * js_sin(phi);
  out[2] = z;
}

/* materials.ts:77-81 */
double o64_schlick(double cosTheta, double etaI, double etaT) {
  double q = (etaI - etaT) / (etaI + etaT);
  double r0 = q * q;
  double c = js_max2(0, js_min2(1, cosTheta));
  double m = 1 - c;
  return r0 + (1 - r0) * m * m * m * m * m;
}

/* materials.ts:89-95 */
double o64_dielectric_reflectance(double cosI, double etaI, double etaT) {
  double eta = etaI / etaT;
  double sin2T = eta * eta * (1 - cosI * cosI);
  if (sin2T >= 1) return 1;
  if (etaI > etaT) return o64_schlick(sqrt(1 - sin2T), etaI, etaT);
  return o64_schlick(cosI, etaI, etaT);
}

/* materials.ts:102-124 */
void o64_metal(double dx, double dy, double dz, double nx, double ny, double nz, double roughness,
               double u1, double u2, double u3, double *out) {
  double k = 2 * (dx * nx + dy * ny + dz * nz);
  double rx = dx - k * nx;
  double ry = dy - k * ny;
  double rz = dz - k * nz;
  if (roughness > 0) {
    o64_unit_vector(u1, u2, out);
    double rad = js_cbrt(u3) * roughness;
    rx += out[0] * rad;
    ry += out[1] * rad;
    rz += out[2] * rad;
  }
  double inv = 1 / sqrt(rx * rx + ry * ry + rz * rz);
  out[0] = rx * inv;
  out[1] = ry * inv;
  out[2] = rz * inv;
}

/* materials.ts:131-166 */
int o64_dielectric(double dx, double dy, double dz, double nx, double ny, double nz, double ior,
                   int frontFace, double u, double *out) {
  double etaI = frontFace ? 1 : ior;
  double etaT = frontFace ? ior : 1;
  double eta = etaI / etaT;
  double cosI = -(dx * nx + dy * ny + dz * nz);
  if (cosI > 1) cosI = 1;
  double reflectance = o64_dielectric_reflectance(cosI, etaI, etaT);
  double sin2T = eta * eta * (1 - cosI * cosI);
  if (u < reflectance || sin2T > 1) {
    double k = 2 * (dx * nx + dy * ny + dz * nz);
    out[0] = dx - k * nx;
    out[1] = dy - k * ny;
    out[2] = dz - k * nz;
    return 1;
  }
  double cosT = sqrt(1 - sin2T);
  double s = eta * cosI - cosT;
  double tx = dx * eta + nx * s;
  double ty = dy * eta + ny * s;
  double tz = dz * eta + nz * s;
  double inv = 1 / sqrt(tx * tx + ty * ty + tz * tz);
  out[0] = tx * inv;
  out[1] = ty * inv;
  out[2] = tz * inv;
  return 0;
}

/* materials.ts:169-174 */
double o64_power_heuristic(double a, double b) {
  double a2 = a * a;
  double b2 = b * b;
  double denom = a2 + b2;
  return denom > 0 ? a2 / denom : 0;
}

/* materials.ts:177-181 */
int o64_checker_primary(double px, double py, double pz, double s) {
  double sum = floor(px * s + 0.5) + floor(py * s + 0.5) + floor(pz * s + 0.5);
  return fmod(fmod(sum, 2) + 2, 2) == 0;
}

/* ======================================================================
 * camera.ts
 * ====================================================================== */

static void normalize3(double *v) {
  double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (len > 0) {
    double s = 1 / len;
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
  } else {
    v[0] = v[1] = v[2] = 0;
  }
}

/* camera.ts:25-51 */
void o64_camera_frame(const double *target, double yaw, double pitch, double distance,
                      double fovDeg, double lensRadius, double focusDistance, double aspect,
                      double *f) {
  double cp = cos(pitch);
  double o[3] = {target[0] + cp * sin(yaw) * distance, target[1] + sin(pitch) * distance,
                 target[2] + cp * cos(yaw) * distance};
  double fw[3] = {target[0] - o[0], target[1] - o[1], target[2] - o[2]};
  normalize3(fw);
  /* cross(forward, (0,1,0)) */
  double r[3] = {fw[1] * 0 - fw[2] * 1, fw[2] * 0 - fw[0] * 0, fw[0] * 1 - fw[1] * 0};
  if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < 1e-12) {
    r[0] = 1;
    r[1] = 0;
    r[2] = 0;
  }
  normalize3(r);
  double up[3] = {r[1] * fw[2] - r[2] * fw[1], r[2] * fw[0] - r[0] * fw[2],
                  r[0] * fw[1] - r[1] * fw[0]};
  f[0] = o[0]; f[1] = o[1]; f[2] = o[2];
  f[3] = r[0]; f[4] = r[1]; f[5] = r[2];
  f[6] = up[0]; f[7] = up[1]; f[8] = up[2];
  f[9] = fw[0]; f[10] = fw[1]; f[11] = fw[2];
  f[12] = tan((fovDeg * PI_D) / 360);
  f[13] = aspect;
  f[14] = lensRadius;
  f[15] = js_max2(1e-3, focusDistance);
}

/* camera.ts:62-72 */
void o64_generate_ray(const double *f, double sx, double sy, double lx, double ly, double *o3,
                      double *d3) {
  double px = (2 * sx - 1) * f[12] * f[13];
  double py = (2 * sy - 1) * f[12];
  double lr = f[14], fd = f[15];
  double focus[3], dir[3];
  for (int k = 0; k < 3; k++) {
    focus[k] = f[k] + (f[3 + k] * px + f[6 + k] * py + f[9 + k]) * fd;
    o3[k] = f[k] + (f[3 + k] * (lx * lr) + f[6 + k] * (ly * lr));
  }
  for (int k = 0; k < 3; k++)
    dir[k] = PLANTED(O64_PLANT_LENS_NO_REFOCUS) ? focus[k] - f[k] : focus[k] - o3[k];
  normalize3(dir);
  d3[0] = dir[0];
  d3[1] = dir[1];
  d3[2] = dir[2];
}

/* ======================================================================
 * tracer.ts
 * ====================================================================== */

/* tracer.ts:61-90 */
static double light_pdf(const O64Scene *s, double px, double py, double pz, double hx, double hy,
                        double hz, double hnx, double hny, double hnz, int prim) {
  int lightCount = s->lightCount;
  if (lightCount == 0) return 0;
  const float *prims = s->prims;
  double tx = hx - px, ty = hy - py, tz = hz - pz;
  double dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return 0;
  double inv = 1 / sqrt(dist2);
  double wx = tx * inv, wy = ty * inv, wz = tz * inv;
  if (PRIMF(prim, 3) == O64_PRIM_SPHERE) {
    double cx = PRIMF(prim, 0) - px, cy = PRIMF(prim, 1) - py, cz = PRIMF(prim, 2) - pz;
    double r = PRIMF(prim, 4);
    double d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      double cosMax = sqrt(js_max2(0, 1 - (r * r) / d2));
      double solid = TWO_PI * (1 - cosMax);
      return solid > 0 ? 1 / (solid * lightCount) : 0;
    }
  }
  double cosL = fabs(hnx * wx + hny * wy + hnz * wz);
  if (cosL < 1e-6) return 0;
  return dist2 / (PRIMF(prim, 11) * cosL * lightCount);
}

/* tracer.ts:93-180. out: [dir3, dist, pdf, emission3]. */
static int sample_light(const O64Scene *s, double px, double py, double pz, O64Rng *rng,
                        double *out) {
  int lightCount = s->lightCount;
  if (lightCount == 0) return 0;
  int prim = s->lights[o64_rng_int(rng, lightCount)];
  const float *prims = s->prims;
  size_t m = (size_t)s->primObject[prim] * MAT_STRIDE;
  out[5] = s->mats[m + 12];
  out[6] = s->mats[m + 13];
  out[7] = s->mats[m + 14];
  if (PRIMF(prim, 3) == O64_PRIM_SPHERE) {
    double cx = PRIMF(prim, 0) - px, cy = PRIMF(prim, 1) - py, cz = PRIMF(prim, 2) - pz;
    double r = PRIMF(prim, 4);
    double d2 = cx * cx + cy * cy + cz * cz;
    if (d2 > r * r) {
      double cosMax = PLANTED(O64_PLANT_CONE_NO_SQRT) ? js_max2(0, 1 - (r * r) / d2)
                                                      : sqrt(js_max2(0, 1 - (r * r) / d2));
      double cosT = PLANTED(O64_PLANT_CONE_ANGLE_UNIFORM)
                        ? cos(o64_rng_next(rng) * acos(cosMax))
                        : 1 - o64_rng_next(rng) * (1 - cosMax);
      double sinT = sqrt(js_max2(0, 1 - cosT * cosT));
      double phi = TWO_PI * o64_rng_next(rng);
      double invD = 1 / sqrt(d2);
      double ax = cx * invD, ay = cy * invD, az = cz * invD;
      double b[6];
      o64_basis(ax, ay, az, b);
      double sx = sinT * js_cos(phi), sy = sinT * js_sin(phi);
      double wx = b[0] * sx + b[3] * sy + ax * cosT;
      double wy = b[1] * sx + b[4] * sy + ay * cosT;
      double wz = b[2] * sx + b[5] * sy + az * cosT;
      double invW = 1 / sqrt(wx * wx + wy * wy + wz * wz);
      wx *= invW;
      wy *= invW;
      wz *= invW;
      double halfB = -(cx * wx + cy * wy + cz * wz);
      double disc = halfB * halfB - (d2 - r * r);
      if (disc < 0) return 0;
      double dist = -halfB - sqrt(disc);
      if (dist <= 0) return 0;
      out[0] = wx;
      out[1] = wy;
      out[2] = wz;
      out[3] = dist;
      out[4] = 1 / ((PLANTED(O64_PLANT_CONE_PDF_2X) ? 2 * TWO_PI : TWO_PI) * (1 - cosMax) *
                    lightCount);
      return 1;
    }
    double u1 = o64_rng_next(rng);
    double u2 = o64_rng_next(rng);
    double nv[3];
    o64_unit_vector(u1, u2, nv);
    double tx = cx + nv[0] * r, ty = cy + nv[1] * r, tz = cz + nv[2] * r;
    double dist2 = tx * tx + ty * ty + tz * tz;
    if (dist2 <= 0) return 0;
    double dist = sqrt(dist2);
    double wx = tx / dist, wy = ty / dist, wz = tz / dist;
    double cosL = fabs(nv[0] * wx + nv[1] * wy + nv[2] * wz);
    if (cosL < 1e-6) return 0;
    out[0] = wx;
    out[1] = wy;
    out[2] = wz;
    out[3] = dist;
    out[4] = dist2 / ((PLANTED(O64_PLANT_INSIDE_HALF_AREA) ? 0.5 * PRIMF(prim, 11) : PRIMF(prim, 11)) *
                      cosL * lightCount);
    return 1;
  }
  double u1 = o64_rng_next(rng);
  double u2 = o64_rng_next(rng);
  if (u1 + u2 > 1) {
    u1 = 1 - u1;
    u2 = 1 - u2;
  }
  double ax = PRIMF(prim, 0), ay = PRIMF(prim, 1), az = PRIMF(prim, 2);
  double qx = ax + (PRIMF(prim, 4) - ax) * u1 + (PRIMF(prim, 8) - ax) * u2;
  double qy = ay + (PRIMF(prim, 5) - ay) * u1 + (PRIMF(prim, 9) - ay) * u2;
  double qz = az + (PRIMF(prim, 6) - az) * u1 + (PRIMF(prim, 10) - az) * u2;
  double tx = qx - px, ty = qy - py, tz = qz - pz;
  double dist2 = tx * tx + ty * ty + tz * tz;
  if (dist2 <= 0) return 0;
  double dist = sqrt(dist2);
  double wx = tx / dist, wy = ty / dist, wz = tz / dist;
  double cosL = fabs(PRIMF(prim, 12) * wx + PRIMF(prim, 13) * wy + PRIMF(prim, 14) * wz);
  if (cosL < 1e-6) return 0;
  out[0] = wx;
  out[1] = wy;
  out[2] = wz;
  out[3] = dist;
  out[4] = dist2 / (PRIMF(prim, 11) * cosL * lightCount);
  return 1;
}

/* tracer.ts:191-328 */
void o64_radiance(const O64Scene *s, double ox, double oy, double oz, double dx, double dy,
                  double dz, O64Rng *rng, const O64TraceOpts *opts, double *out,
                  O64RayCounts *rc) {
  double Lr = 0, Lg = 0, Lb = 0;
  double br = 1, bg = 1, bb = 1;
  int specularBounce = 1;
  double prevPdf = 0;
  double prevX = ox, prevY = oy, prevZ = oz;
  const float *prims = s->prims, *mats = s->mats;
  int useNee = opts->nee && s->lightCount > 0;
  double ls[8], dir[3];
  for (int bounce = 0; bounce <= opts->maxBounces; bounce++) {
    O64Hit hit;
    double o[3] = {ox, oy, oz}, d[3] = {dx, dy, dz};
    if (rc) rc->extension++;
    if (!o64_intersect_closest(s, o, d, T_MIN, INFINITY, &hit, NULL)) {
      if (s->skyKind == 1) {
        double t = PLANTED(O64_PLANT_SKY_FLIP) ? 0.5 * (1 - dy) : 0.5 * (dy + 1);
        Lr += br * (s->horizon[0] + (s->zenith[0] - s->horizon[0]) * t) * s->skyStrength;
        Lg += bg * (s->horizon[1] + (s->zenith[1] - s->horizon[1]) * t) * s->skyStrength;
        Lb += bb * (s->horizon[2] + (s->zenith[2] - s->horizon[2]) * t) * s->skyStrength;
      }
      break;
    }
    int prim = hit.prim;
    double t = hit.t;
    double px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t;
    double nx, ny, nz;
    if (PRIMF(prim, 3) == O64_PRIM_SPHERE) {
      double invR = 1 / PRIMF(prim, 4);
      nx = (px - PRIMF(prim, 0)) * invR;
      ny = (py - PRIMF(prim, 1)) * invR;
      nz = (pz - PRIMF(prim, 2)) * invR;
    } else {
      nx = PRIMF(prim, 12);
      ny = PRIMF(prim, 13);
      nz = PRIMF(prim, 14);
    }
    int frontFace = dx * nx + dy * ny + dz * nz < 0;
    double sx = frontFace ? nx : -nx;
    double sy = frontFace ? ny : -ny;
    double sz = frontFace ? nz : -nz;
    size_t m = (size_t)s->primObject[prim] * MAT_STRIDE;
    double type = mats[m + 3];
    double er = mats[m + 12], eg = mats[m + 13], eb = mats[m + 14];
    if (er > 0 || eg > 0 || eb > 0) {
      double w = 1;
      if (useNee && !specularBounce) {
        double pl = light_pdf(s, prevX, prevY, prevZ, px, py, pz, nx, ny, nz, prim);
        w = o64_power_heuristic(prevPdf, pl);
      }
      Lr += br * er * w;
      Lg += bg * eg * w;
      Lb += bb * eb * w;
    }
    if (type == 3) break; /* MAT_EMISSIVE */
    double ndx, ndy, ndz;
    if (type == 0) { /* MAT_LAMBERTIAN */
      double ar = mats[m], ag = mats[m + 1], ab = mats[m + 2];
      if (mats[m + 6] > 0.5 && !o64_checker_primary(px, py, pz, mats[m + 5])) {
        ar = mats[m + 8];
        ag = mats[m + 9];
        ab = mats[m + 10];
      }
      if (useNee && sample_light(s, px, py, pz, rng, ls)) {
        double wx = ls[0], wy = ls[1], wz = ls[2];
        double pdfL = ls[4];
        double cosS = sx * wx + sy * wy + sz * wz;
        if (cosS > 0 && pdfL > 0) {
          double po[3] = {px, py, pz}, pw[3] = {wx, wy, wz};
          if (rc) rc->shadow++;
          if (!o64_occluded(s, po, pw, T_MIN, ls[3] * SHADOW_SHORTEN, NULL)) {
            double pdfB = cosS * INV_PI;
            double k = (cosS * INV_PI * o64_power_heuristic(pdfL, pdfB)) / pdfL;
            Lr += br * ar * ls[5] * k;
            Lg += bg * ag * ls[6] * k;
            Lb += bb * ab * ls[7] * k;
          }
        }
      }
      double u1 = o64_rng_next(rng);
      double u2 = o64_rng_next(rng);
      o64_cosine_hemisphere(sx, sy, sz, u1, u2, dir);
      ndx = dir[0];
      ndy = dir[1];
      ndz = dir[2];
      prevPdf = js_max2(0, sx * ndx + sy * ndy + sz * ndz) * INV_PI;
      br *= ar;
      bg *= ag;
      bb *= ab;
      specularBounce = 0;
    } else if (type == 1) { /* MAT_METAL */
      double u1 = o64_rng_next(rng);
      double u2 = o64_rng_next(rng);
      double u3 = o64_rng_next(rng);
      o64_metal(dx, dy, dz, sx, sy, sz, mats[m + 4], u1, u2, u3, dir);
      ndx = dir[0];
      ndy = dir[1];
      ndz = dir[2];
      if (ndx * sx + ndy * sy + ndz * sz <= 0) break;
      br *= mats[m];
      bg *= mats[m + 1];
      bb *= mats[m + 2];
      specularBounce = 1;
    } else {
      double u = o64_rng_next(rng);
      o64_dielectric(dx, dy, dz, sx, sy, sz, mats[m + 4], frontFace, u, dir);
      ndx = dir[0];
      ndy = dir[1];
      ndz = dir[2];
      br *= mats[m];
      bg *= mats[m + 1];
      bb *= mats[m + 2];
      specularBounce = 1;
    }
    if (opts->russianRoulette && bounce >= 3) {
      double maxB = br > bg ? (br > bb ? br : bb) : bg > bb ? bg : bb;
      double pSurvive = maxB > 1 ? 1 : maxB < 0.05 ? 0.05 : maxB;
      if (o64_rng_next(rng) > pSurvive) break;
      br /= pSurvive;
      bg /= pSurvive;
      bb /= pSurvive;
    }
    prevX = px;
    prevY = py;
    prevZ = pz;
    ox = px;
    oy = py;
    oz = pz;
    dx = ndx;
    dy = ndy;
    dz = ndz;
  }
  out[0] = Lr;
  out[1] = Lg;
  out[2] = Lb;
}

/* The JS engine's Math.sin/cos/cbrt as the oracle computes them (jsmath.h),
 * exported for the tests that pin them against the engine's own outputs. */
double o64_js_sin(double x) { return js_sin(x); }
double o64_js_cos(double x) { return js_cos(x); }
double o64_js_cbrt(double x) { return js_cbrt(x); }

/* The camera ray of one path (tracer.ts:373-383): pixel jitter, then the lens
 * draws when lensRadius > 0, then generateRay.  For the lens and sky pins. */
void o64_camera_sample(const double *frame, int width, int height, int x, int y,
                       uint32_t frameIndex, double *o3, double *d3) {
  uint32_t pixel = (uint32_t)y * (uint32_t)width + (uint32_t)x;
  O64Rng rng;
  rng.state = o64_seed_for(pixel, frameIndex);
  double sx = (x + o64_rng_next(&rng)) / width;
  double sy = 1 - (y + o64_rng_next(&rng)) / height;
  double lx = 0, ly = 0;
  if (frame[14] > 0) {
    double r = PLANTED(O64_PLANT_LENS_LINEAR_R) ? o64_rng_next(&rng) : sqrt(o64_rng_next(&rng));
    double theta = TWO_PI * o64_rng_next(&rng);
    lx = r * js_cos(theta);
    ly = r * js_sin(theta);
  }
  o64_generate_ray(frame, sx, sy, lx, ly, o3, d3);
}

/* sampleLightInto / lightPdfScalar (tracer.ts:61-180) from a given RNG state,
 * for the light-sampling pins.  *state is advanced as the draws consume it. */
int o64_sample_light(const O64Scene *s, double px, double py, double pz, uint32_t *state,
                     double *out8) {
  O64Rng rng;
  rng.state = *state;
  int ok = sample_light(s, px, py, pz, &rng, out8);
  *state = rng.state;
  return ok;
}
double o64_light_pdf(const O64Scene *s, double px, double py, double pz, double hx, double hy,
                     double hz, double hnx, double hny, double hnz, int prim) {
  return light_pdf(s, px, py, pz, hx, hy, hz, hnx, hny, hnz, prim);
}

/* Camera ray + radiance of one path: tracer.ts:373-385 */
static void pixel_sample(const O64Scene *s, const double *frame, int width, int height, int x,
                         int y, uint32_t frameIndex, const O64TraceOpts *opts, double *c,
                         O64RayCounts *rc) {
  uint32_t pixel = (uint32_t)y * (uint32_t)width + (uint32_t)x;
  O64Rng rng;
  rng.state = o64_seed_for(pixel, frameIndex);
  double sx = (x + o64_rng_next(&rng)) / width;
  double sy = 1 - (y + o64_rng_next(&rng)) / height;
  double lx = 0, ly = 0;
  if (frame[14] > 0) {
    double r = PLANTED(O64_PLANT_LENS_LINEAR_R) ? o64_rng_next(&rng) : sqrt(o64_rng_next(&rng));
    double theta = TWO_PI * o64_rng_next(&rng);
    lx = r * js_cos(theta);
    ly = r * js_sin(theta);
  }
  double o[3], d[3];
  o64_generate_ray(frame, sx, sy, lx, ly, o, d);
  o64_radiance(s, o[0], o[1], o[2], d[0], d[1], d[2], &rng, opts, c, rc);
}

/* tracer.ts:361-395 */
void o64_accumulate_pass(const O64Scene *s, const double *frame, int width, int height,
                         uint32_t frameIndex, const O64TraceOpts *opts, float *accum, int y0,
                         int y1, int64_t *rayCounts2) {
  O64RayCounts rc = {0, 0};
  for (int y = y0; y < y1; y++) {
    for (int x = 0; x < width; x++) {
      double c[3];
      pixel_sample(s, frame, width, height, x, y, frameIndex, opts, c, &rc);
      size_t i = ((size_t)y * width + x) * 3;
      if (isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2])) {
        accum[i] = (float)((double)accum[i] + c[0]);
        accum[i + 1] = (float)((double)accum[i + 1] + c[1]);
        accum[i + 2] = (float)((double)accum[i + 2] + c[2]);
      }
    }
  }
  if (rayCounts2) {
    rayCounts2[0] += rc.extension;
    rayCounts2[1] += rc.shadow;
  }
}

void o64_pass_radiance(const O64Scene *s, const double *frame, int width, int height,
                       uint32_t frameIndex, const O64TraceOpts *opts, double *out, int y0,
                       int y1) {
  for (int y = y0; y < y1; y++)
    for (int x = 0; x < width; x++)
      pixel_sample(s, frame, width, height, x, y, frameIndex, opts,
                   out + ((size_t)y * width + x) * 3, NULL);
}

/* renderImage (tracer.ts:348-355) of single pixels: each listed pixel's spp
 * passes (frame index s + 7919*seed) summed in float32 in pass order, then
 * times 1/spp -- the values renderImage leaves there.  out: 3 per pixel. */
void o64_render_pixels(const O64Scene *s, const double *frame, int width, int height,
                       const uint32_t *pixels, int count, int spp, uint32_t seed,
                       const O64TraceOpts *opts, float *out) {
  double inv = 1.0 / spp;
  for (int k = 0; k < count; k++) {
    int x = (int)(pixels[k] % (uint32_t)width), y = (int)(pixels[k] / (uint32_t)width);
    float acc[3] = {0, 0, 0};
    for (int sIdx = 0; sIdx < spp; sIdx++) {
      double c[3];
      pixel_sample(s, frame, width, height, x, y, (uint32_t)sIdx + seed * 7919u, opts, c, NULL);
      if (isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2]))
        for (int j = 0; j < 3; j++) acc[j] = (float)((double)acc[j] + c[j]);
    }
    for (int j = 0; j < 3; j++) out[3 * k + j] = (float)((double)acc[j] * inv);
  }
}

/* tracer.ts:348-355 */
void o64_render_image(const O64Scene *s, const double *frame, int width, int height, int spp,
                      uint32_t seed, const O64TraceOpts *opts, float *out) {
  size_t n = (size_t)width * height * 3;
  memset(out, 0, n * sizeof(float));
  for (int sIdx = 0; sIdx < spp; sIdx++)
    o64_accumulate_pass(s, frame, width, height, (uint32_t)sIdx + seed * 7919u, opts, out, 0,
                        height, NULL);
  double inv = 1.0 / spp;
  for (size_t i = 0; i < n; i++) out[i] = (float)((double)out[i] * inv);
}

/* ======================================================================
 * tonemap.ts
 * ====================================================================== */

/* tonemap.ts:7-16 */
double o64_aces(double x) {
  const double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  double v = (x * (a * x + b)) / (x * (c * x + d) + e);
  return v < 0 ? 0 : v > 1 ? 1 : v;
}

/* tonemap.ts:19-21 */
double o64_reinhard(double x) { return x <= 0 ? 0 : x / (1 + x); }

/* tonemap.ts:28-32 */
double o64_linear_to_srgb(double x) {
  if (x <= 0) return 0;
  if (x >= 1) return 1;
  return x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1 / 2.4) - 0.055;
}

/* Math.round: nearest integer, ties toward +infinity. */
static double js_round(double x) {
  double f = floor(x);
  return (x - f >= 0.5) ? f + 1 : f;
}

static uint8_t clamp_u8(double v) {
  /* Uint8ClampedArray store of an integer-valued double */
  if (!(v > 0)) return 0;
  if (v > 255) return 255;
  return (uint8_t)v;
}

/* tonemap.ts:41-62 */
void o64_tonemap_buffer(const float *accum, int width, int height, double samples,
                        double exposureEv, int mode, uint8_t *rgba) {
  double inv = samples > 0 ? pow(2, exposureEv) / samples : 0;
  for (int i = 0; i < width * height; i++) {
    for (int k = 0; k < 3; k++) {
      double x = (double)accum[i * 3 + k] * inv;
      double tm = mode == 0 ? o64_aces(x) : o64_reinhard(x);
      rgba[i * 4 + k] = clamp_u8(js_round(o64_linear_to_srgb(tm) * 255));
    }
    rgba[i * 4 + 3] = 255;
  }
}
