/*
 * O64 — CPU oracle for the caustic-tracer path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker.
 *
 * O64 restates the reference TypeScript algorithms in plain C with the
 * reference's arithmetic: every computation is IEEE binary64 (JS numbers),
 * every stored array keeps the reference's element type (Float32Array ->
 * float with round-to-nearest-even, Float64Array -> double, Int32Array,
 * Uint32Array).  Build with -ffp-contract=off and no fast-math so that
 * each + - * / sqrt is one correctly rounded operation, as in V8.
 *
 * Pinning (what independently checks each part):
 *   rng            integer-exact restatement of src/tracer/rng.ts; pinned by
 *                  the golden Cornell PPM (any draw out of place changes it)
 *                  and by closed-form statistics in tests/test_oracle_*.py
 *   pack + bvh     pinned by validateBvh invariants, brute-force agreement,
 *                  the prefix-count closed form of the two-pointer partition
 *                  and the golden Cornell PPM
 *   walks          pinned against brute force (bvh.test.ts:87-146) and closed
 *                  forms (coincident spheres t = 7.5)
 *   materials      closed forms of materials.test.ts (Snell, Schlick, mirror,
 *                  cos-weighted moments, power-heuristic sum)
 *   tracer         golden PPM out/cornell.ppm (reference render, 96x64x16)
 *                  and the white-furnace closed form E/(1-a) = 5
 *   tonemap        golden PPM plus reference.test.ts:61-94 closed forms
 * Transcendentals (sin, cos, cbrt, tan, pow) come from the C library; V8's
 * last-ulp results for them are not reproducible here, which the golden PPM
 * bounds at 8-bit precision (parity of those last ulps is unpinned).
 */
#ifndef CAUSTIC_O64_H
#define CAUSTIC_O64_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rng.ts ------------------------------------------------------------ */
typedef struct { uint32_t state; } O64Rng;
uint32_t o64_pcg_hash(uint32_t input);
uint32_t o64_seed_for(uint32_t pixel, uint32_t frame);
void o64_rng_init(O64Rng *r, uint32_t seed);
uint32_t o64_rng_u32(O64Rng *r);
double o64_rng_next(O64Rng *r);
double o64_rng_range(O64Rng *r, double lo, double hi);
int o64_rng_int(O64Rng *r, int n);
/* Bulk helpers for tests: n draws of next() from Rng(seed). */
void o64_rng_stream(uint32_t seed, int n, double *out);
void o64_rng_stream_u32(uint32_t seed, int n, uint32_t *out);

/* ---- pack.ts + bvh.ts ------------------------------------------------- */
enum { O64_PRIM_SPHERE = 0, O64_PRIM_TRIANGLE = 1 };

/*
 * Build the packed scene from primitives already expanded in scene order
 * (quads and meshes expanded to triangles as scene.ts does).
 *   kind[i]      0 sphere, 1 triangle
 *   geom[9*i..]  sphere: cx,cy,cz,r ; triangle: ax,ay,az,bx,by,bz,cx,cy,cz
 *   obj[i]       scene object index
 *   emits[o]     1 when object o's material emits (materialEmission != null)
 * Outputs (caller-allocated): prims[max(1,n)*16], primObject[max(1,n)],
 * nodes[max(1,2n-1)*8], order[max(1,n)], lights[max(1,n)].
 * Returns nodeCount; *lightCount receives the number of light prims.
 */
int o64_pack_scene(int n, const int32_t *kind, const double *geom, const int32_t *obj,
                   const uint8_t *emits, int bins,
                   float *prims, uint32_t *primObject, float *nodes, uint32_t *order,
                   int32_t *lights, int *lightCount);

/* PrimBounds exactly as enumeratePrimitives computes them (pack.ts:131-192). */
void o64_prim_bounds(int n, const int32_t *kind, const double *geom,
                     double *bmin, double *bmax, double *bcen);

/* buildBvh (bvh.ts:72-301) with a configurable bucket count (12 in the reference). */
int o64_build_bvh(int n, const double *bmin, const double *bmax, const double *bcen, int bins,
                  float *nodes, uint32_t *order);

/* validateBvh (bvh.ts:319-381). Returns number of errors; stats[0]=maxLeaf,
 * stats[1]=leafCount, stats[2]=depth. */
int o64_validate_bvh(const float *nodes, int nodeCount, const uint32_t *order, int primCount,
                     int32_t *stats);

/* o64_pack_scene plus the exact leaf words (see O64Scene). */
int o64_pack_scene_ex(int n, const int32_t *kind, const double *geom, const int32_t *obj,
                      const uint8_t *emits, int bins, float *prims, uint32_t *primObject,
                      float *nodes, uint32_t *order, int32_t *lights, int *lightCount,
                      uint32_t *leafWords);
/* o64_build_bvh plus the exact leaf word of every node (see O64Scene), and
 * validateBvh reading those words instead of the float32 ones. */
int o64_build_bvh_ex(int n, const double *bmin, const double *bmax, const double *bcen, int bins,
                     float *nodes, uint32_t *order, uint32_t *leafWords);
int o64_validate_bvh_ex(const float *nodes, int nodeCount, const uint32_t *order, int primCount,
                        const uint32_t *leafWords, int32_t *stats);

/* Two-pointer partition exactly as bvh.ts:262-278, on a flag vector (tests). */
int o64_partition_flags(int n, const uint8_t *good, int32_t *perm);

/* ---- packed scene view ------------------------------------------------ */
typedef struct {
  const float *prims; int primCount;
  const float *nodes; int nodeCount;
  const float *mats; int matCount;
  const int32_t *lights; int lightCount;
  const uint32_t *primObject;
  int skyKind;                 /* 0 none, 1 gradient */
  double zenith[3], horizon[3], skyStrength;
  /* Layout extension: the exact leaf word start*8+count of every node
   * (0xffffffff on interior nodes), for scenes of 2^21 primitives or more
   * where the float32 words of bvh.ts:60-62 stop being exact.  NULL: the
   * walks read the float32 words, as the reference does. */
  const uint32_t *leafWords;
} O64Scene;

/* ---- intersect.ts ----------------------------------------------------- */
typedef struct { double t; int32_t prim; double u, v; } O64Hit;
typedef struct { int64_t nodes; int64_t tris; int64_t spheres; int64_t leaves; } O64Counts;

double o64_ray_sphere(double ox, double oy, double oz, double dx, double dy, double dz,
                      double cx, double cy, double cz, double r, double tMin, double tMax);
int o64_ray_triangle(double ox, double oy, double oz, double dx, double dy, double dz,
                     double ax, double ay, double az, double bx, double by, double bz,
                     double cx, double cy, double cz, double tMin, double tMax, double *out3);
double o64_ray_aabb(double ox, double oy, double oz, double idx, double idy, double idz,
                    double minX, double minY, double minZ, double maxX, double maxY, double maxZ,
                    double tMax);

int o64_intersect_closest(const O64Scene *s, const double *o, const double *d, double tMin,
                          double tMax, O64Hit *hit, O64Counts *cnt);
int o64_intersect_stackless(const O64Scene *s, const double *o, const double *d, double tMin,
                            double tMax, O64Hit *hit, O64Counts *cnt);
int o64_occluded(const O64Scene *s, const double *o, const double *d, double tMin, double tMax,
                 O64Counts *cnt);
int o64_intersect_brute(const O64Scene *s, const double *o, const double *d, double tMin,
                        double tMax, O64Hit *hit);

/* Batched forms for tests: rays[6*i] = ox,oy,oz,dx,dy,dz; tmax per ray.
 * walk: 0 ordered, 1 stackless, 2 brute force. counts: 4 int64 totals. */
void o64_closest_batch(const O64Scene *s, int n, const double *rays, double tMin,
                       const double *tMax, int walk, int32_t *prim, double *t, double *u,
                       double *v, int64_t *counts);
void o64_occluded_batch(const O64Scene *s, int n, const double *rays, double tMin,
                        const double *tMax, uint8_t *out, int64_t *counts);

/* ---- materials.ts ----------------------------------------------------- */
void o64_basis(double nx, double ny, double nz, double *out6);
void o64_cosine_hemisphere(double nx, double ny, double nz, double u1, double u2, double *out3);
void o64_unit_vector(double u1, double u2, double *out3);
double o64_schlick(double cosTheta, double etaI, double etaT);
double o64_dielectric_reflectance(double cosI, double etaI, double etaT);
void o64_metal(double dx, double dy, double dz, double nx, double ny, double nz, double rough,
               double u1, double u2, double u3, double *out3);
int o64_dielectric(double dx, double dy, double dz, double nx, double ny, double nz, double ior,
                   int frontFace, double u, double *out3);
double o64_power_heuristic(double a, double b);
int o64_checker_primary(double px, double py, double pz, double cells);

/* ---- camera.ts -------------------------------------------------------- */
/* frame[16] = origin3, right3, up3, forward3, tanHalfFov, aspect, lensRadius, focusDistance */
void o64_camera_frame(const double *target, double yaw, double pitch, double distance,
                      double fovDeg, double lensRadius, double focusDistance, double aspect,
                      double *frame16);
void o64_generate_ray(const double *frame16, double sx, double sy, double lx, double ly,
                      double *o3, double *d3);

/* ---- tracer.ts -------------------------------------------------------- */
typedef struct { int maxBounces; int russianRoulette; int nee; } O64TraceOpts;
/* Per-path ray counters: extension rays traced and shadow rays traced. */
typedef struct { int64_t extension; int64_t shadow; } O64RayCounts;

void o64_radiance(const O64Scene *s, double ox, double oy, double oz, double dx, double dy,
                  double dz, O64Rng *rng, const O64TraceOpts *opts, double *out3,
                  O64RayCounts *rc);
/* accumulatePass over rows [y0, y1) (full image: 0, height). */
void o64_accumulate_pass(const O64Scene *s, const double *frame16, int width, int height,
                         uint32_t frameIndex, const O64TraceOpts *opts, float *accum, int y0,
                         int y1, int64_t *rayCounts2);
/* Per-sample radiance of one pass (no accumulation), rows [y0,y1): out[3*pixel]. */
void o64_pass_radiance(const O64Scene *s, const double *frame16, int width, int height,
                       uint32_t frameIndex, const O64TraceOpts *opts, double *out, int y0,
                       int y1);
/* renderImage: spp passes with frameIndex = s + seed*7919, then *= 1/spp. */
void o64_render_image(const O64Scene *s, const double *frame16, int width, int height, int spp,
                      uint32_t seed, const O64TraceOpts *opts, float *out);

/* ---- tonemap.ts ------------------------------------------------------- */
double o64_aces(double x);
double o64_reinhard(double x);
double o64_linear_to_srgb(double x);
/* mode 0 aces, 1 reinhard; rgba out (4 bytes per pixel). */
void o64_tonemap_buffer(const float *accum, int width, int height, double samples,
                        double exposureEv, int mode, uint8_t *rgba);

/* Planted-mistake numbers (-DO64_PLANT=k; tests/test_oracle_pins.py builds
 * each variant and shows that a pin fails on it). */
#define O64_PLANT_SAH_LE 1              /* cost <= best: the last of equal-cost splits */
#define O64_PLANT_AXIS_REVERSED 2       /* axes tried z, y, x */
#define O64_PLANT_NUDGE_INWARD 3        /* node boxes nudged inward */
#define O64_PLANT_PARTITION_LT 4        /* bucket < bestSplit goes left */
#define O64_PLANT_MEDIAN_UNSTABLE 5     /* equal centroids reordered */
#define O64_PLANT_MEDIAN_AXIS_LAST 6    /* median axis ties prefer z */
#define O64_PLANT_SWEEP_SHIFT 7         /* suffix-sweep counts stored one slot off */
#define O64_PLANT_LEAF3 8               /* leaves of at most three primitives */
#define O64_PLANT_CONE_NO_SQRT 20       /* cone half-angle cosine without the square root */
#define O64_PLANT_CONE_ANGLE_UNIFORM 21 /* cone sampled uniform in angle, not in cosine */
#define O64_PLANT_CONE_PDF_2X 22        /* cone pdf over twice the solid angle */
#define O64_PLANT_INSIDE_HALF_AREA 23   /* inside-sphere pdf with half the area */
#define O64_PLANT_LENS_LINEAR_R 30      /* lens radius square root dropped */
#define O64_PLANT_LENS_NO_REFOCUS 31    /* lens rays aimed from the lens centre */
#define O64_PLANT_SKY_FLIP 40           /* sky gradient upside down */

double o64_js_sin(double x);
double o64_js_cos(double x);
double o64_js_cbrt(double x);
/* renderImage's values at a list of pixels (y*width + x). */
void o64_render_pixels(const O64Scene *s, const double *frame, int width, int height,
                       const uint32_t *pixels, int count, int spp, uint32_t seed,
                       const O64TraceOpts *opts, float *out);
/* Test exports: the camera ray of one path (tracer.ts:373-383), and light
 * sampling from a given RNG state (tracer.ts:61-180). */
void o64_camera_sample(const double *frame, int width, int height, int x, int y,
                       uint32_t frameIndex, double *o3, double *d3);
int o64_sample_light(const O64Scene *s, double px, double py, double pz, uint32_t *state,
                     double *out8);
double o64_light_pdf(const O64Scene *s, double px, double py, double pz, double hx, double hy,
                     double hz, double hnx, double hny, double hnz, int prim);

#ifdef __cplusplus
}
#endif
#endif
