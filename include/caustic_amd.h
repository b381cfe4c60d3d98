/*
 * caustic_amd.h — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * Plain pointers and sizes only.  Every array argument uses the packed layout
 * the reference's packer emits (src/tracer/pack.ts:1-17, src/tracer/bvh.ts:1-16):
 *   prims       float32, 16 per primitive   [a.xyz,type | b.xyz,obj | c.xyz,area | n.xyz,0]
 *   nodes       float32,  8 per BVH node    [min.xyz,skip | max.xyz,leaf]
 *   mats        float32, 16 per object      [rgb,type | param,chkScale,chkOn,0 | rgb2,0 | e.rgb,0]
 *   lights      int32, packed indices of emitting primitives
 *   primObject  uint32, packed primitive -> scene object
 * Rays are binary64 like the JS numbers the reference passes: 6 per ray
 * (ox,oy,oz,dx,dy,dz).  A camera frame is 16 binary64 values:
 * origin.xyz, right.xyz, up.xyz, forward.xyz, tanHalfFov, aspect, lensRadius,
 * focusDistance (camera.ts:8-22).
 *
 * Pointers named d_* are device (HBM) pointers; h_* are host pointers.
 * Every call returns 0 on success or a negative CT_E* code; ct_last_error()
 * returns the message of the calling thread's last failure.  The reference's
 * hot path never throws (a miss is prim -1 / t = +inf, intersect.ts:213-226);
 * errors here are only argument, allocation and launch failures.
 * Handles are independent: calls on different handles may run concurrently
 * from different threads; calls sharing one ct_tracer must be serialised.
 */
#ifndef CAUSTIC_AMD_H
#define CAUSTIC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_OK 0
#define CT_EINVAL -1  /* bad argument */
#define CT_ENOMEM -2  /* device allocation failed */
#define CT_EDEVICE -3 /* HIP runtime or launch error */
#define CT_ELIMIT -4  /* input beyond a documented size limit */
#define CT_ETIMEOUT -5 /* a wait ended before the job did */

typedef struct ct_scene ct_scene;
typedef struct ct_tracer ct_tracer;
typedef struct ct_server ct_server;

/* Sky (scene.ts:57-59): kind 0 none, 1 gradient. */
typedef struct {
  int32_t kind;
  double zenith[3];
  double horizon[3];
  double strength;
} ct_sky;

/* TraceOptions (tracer.ts:27-31). */
typedef struct {
  int32_t maxBounces;
  int32_t russianRoulette;
  int32_t nee;
} ct_trace_opts;

/* RenderOptions (tracer.ts:337-342). */
typedef struct {
  uint32_t width;
  uint32_t height;
  uint32_t spp;
  uint32_t seed;
  int32_t maxBounces;
  int32_t russianRoulette;
  int32_t nee;
} ct_render_opts;

/* Which closest-hit walk a query mirrors. */
#define CT_WALK_STACKLESS 0 /* intersect.ts:290-365 */
#define CT_WALK_ORDERED 1   /* intersect.ts:202-282 */
/* Ordered walk over 4-wide nodes derived from the reference BVH at upload
 * (same closest hit as the reference walks). */
#define CT_WALK_WIDE 2
/* Every primitive tested in packed order (intersectBruteForce,
 * intersect.ts:423-468); FP32 queries only. */
#define CT_WALK_BRUTE 3

/* Arithmetic of a tracer's kernels. */
#define CT_PRECISION_F64 0 /* the reference's binary64 over float32 data (default) */
/* Precision form: 8-bit parent-relative quantized 4-wide nodes (outward
 * rounded; conservative traversal), FP32 intersection and shading, per-pixel
 * Neumaier-compensated FP32 sums.  Images agree with the binary64 form in
 * distribution (same sampling), not per sample. */
#define CT_PRECISION_F32 1

/* Counters of the last render on a tracer (not part of the reference API). */
typedef struct {
  uint64_t extensionRays;  /* closest-hit rays traced */
  uint64_t shadowRays;     /* any-hit rays traced */
  uint64_t paths;          /* camera paths started */
  uint32_t extendLaunches; /* closest-hit kernel launches */
  double extendMs;         /* summed device time of those launches (profiling on) */
  double shadeMs;          /* summed device time of shade launches (profiling on) */
  double shadowMs;         /* summed device time of any-hit launches (profiling on) */
  uint64_t nodeVisits;     /* node fetches of closest-hit walks (counting on) */
  uint64_t triTests;       /* triangle tests of closest-hit walks (counting on) */
  uint64_t sphereTests;    /* sphere tests of closest-hit walks (counting on) */
  uint64_t shadowNodeVisits; /* node fetches of any-hit walks (counting on) */
  uint64_t shadowTriTests;   /* triangle tests of any-hit walks (counting on) */
  uint64_t shadowSphereTests;/* sphere tests of any-hit walks (counting on) */
  uint64_t stackFallbacks;   /* closest-hit rays that outgrew the LDS stack (counting on) */
  uint64_t shadowStackFallbacks; /* any-hit rays that outgrew the LDS stack (counting on) */
  uint32_t rounds;           /* launch rounds the call was split into */
} ct_render_stats;

/* One image of a batch render: renderImage (tracer.ts:348-355) of `scene`
 * seen through `frame` into d_out (width*height*3 float32, overwritten). */
typedef struct {
  const ct_scene *scene;
  double frame[16];
  uint32_t width, height, spp, seed;
  int32_t maxBounces, russianRoulette, nee;
  float *d_out;
} ct_job;

/* Passes of one progressive image: `passes` calls of accumulatePass
 * (tracer.ts:361-395) with frame indices frameIndex, frameIndex+1, ...
 * adding into the float32 sum d_accum (width*height*3). */
typedef struct {
  const ct_scene *scene;
  double frame[16];
  uint32_t width, height;
  uint32_t frameIndex;
  uint32_t passes;
  int32_t maxBounces, russianRoulette, nee;
  float *d_accum;
} ct_pass_job;

/* ---- runtime ------------------------------------------------------------ */
int ct_init(int32_t device);
const char *ct_last_error(void);
/* Device memory for hosts without their own allocator (the N-API addon). */
int ct_device_alloc(void **d_ptr, size_t bytes);
int ct_device_free(void *d_ptr);
int ct_copy_h2d(void *d_dst, const void *h_src, size_t bytes);
int ct_copy_d2h(void *h_dst, const void *d_src, size_t bytes);
int ct_stream_sync(void *stream);

/* ---- scene (replaces handing PackedScene to the CPU tracer) --------------
 * Copies the packed arrays (host pointers) into device-resident buffers owned
 * by the handle.  Mirrors PackedScene (pack.ts:46-59). */
int ct_scene_create(const float *h_prims, uint32_t primCount, const float *h_nodes,
                    uint32_t nodeCount, const float *h_mats, uint32_t matCount,
                    const int32_t *h_lights, uint32_t lightCount, const uint32_t *h_primObject,
                    const ct_sky *sky, ct_scene **out);
/* Same from device-resident arrays (copied on the stream). */
int ct_scene_create_device(const float *d_prims, uint32_t primCount, const float *d_nodes,
                           uint32_t nodeCount, const float *d_mats, uint32_t matCount,
                           const int32_t *d_lights, uint32_t lightCount,
                           const uint32_t *d_primObject, const ct_sky *sky, void *stream,
                           ct_scene **out);
/* Layout extension for scenes of 2^21 primitives or more, where the float32
 * leaf words of pack.ts (start*8 + count, bvh.ts:60-62) stop being exact:
 * h_leafWords holds the exact word of every node as uint32 (0xffffffff on
 * interior nodes; ct_bvh_build_ex writes it).  Every walk decodes leaves from
 * these words (the float32 ones only tell leaves from interior nodes), so all
 * walks return the reference's hits on such scenes as well. */
int ct_scene_create_ex(const float *h_prims, uint32_t primCount, const float *h_nodes,
                       uint32_t nodeCount, const float *h_mats, uint32_t matCount,
                       const int32_t *h_lights, uint32_t lightCount,
                       const uint32_t *h_primObject, const ct_sky *sky,
                       const uint32_t *h_leafWords, ct_scene **out);
int ct_scene_destroy(ct_scene *scene);

/* ---- closest hit / any hit ----------------------------------------------
 * Batched intersectClosest / intersectClosestStackless (intersect.ts:202,290):
 * one query per ray with the given tMin and per-ray tMax (d_tMax NULL = +inf).
 * Outputs per ray: prim (-1 on a miss), t (+inf on a miss), u, v. */
int ct_closest_hit(const ct_scene *scene, const double *d_rays, uint32_t n, double tMin,
                   const double *d_tMax, int32_t walk, int32_t *d_prim, double *d_t, double *d_u,
                   double *d_v, void *stream);
/* Batched occluded (intersect.ts:368): out[i] = 1 when (tMin, tMax[i]) is blocked. */
int ct_occluded(const ct_scene *scene, const double *d_rays, uint32_t n, double tMin,
                const double *d_tMax, uint8_t *d_out, void *stream);
/* Traversal work of a closest-hit batch: counts[0] node fetches, [1] triangle
 * tests, [2] sphere tests (the counters of scripts/debug-aabb.ts:27-55).
 * Synchronises the stream. */
int ct_closest_hit_counts(const ct_scene *scene, const double *d_rays, uint32_t n, double tMin,
                          const double *d_tMax, int32_t walk, uint64_t counts[3], void *stream);

/* FP32 queries of the precision form: rays are 6 float32 per ray, tMax per
 * ray (NULL = +inf); walk CT_WALK_WIDE (quantized nodes) or CT_WALK_BRUTE
 * (every primitive; the exhaustive answer the walk must reproduce).  Exact
 * ties in t may resolve to a different primitive than the scan's. */
int ct_closest_hit32(const ct_scene *scene, const float *d_rays, uint32_t n, float tMin,
                     const float *d_tMax, int32_t walk, int32_t *d_prim, float *d_t, float *d_u,
                     float *d_v, void *stream);
int ct_occluded32(const ct_scene *scene, const float *d_rays, uint32_t n, float tMin,
                  const float *d_tMax, int32_t walk, uint8_t *d_out, void *stream);
/* Work of the quantized walk: adds node-slot tests (4 per node), triangle
 * tests, sphere tests and stack fallbacks into d_counts4 (device uint64[4]). */
int ct_closest_hit32_counts(const ct_scene *scene, const float *d_rays, uint32_t n, float tMin,
                            const float *d_tMax, uint64_t *d_counts4, void *stream);
/* Per-channel binary64 sums of a float32 RGB image (npix pixels) computed on
 * the device in a fixed order (the furnace energy balance); synchronises. */
int ct_image_sum_f64(const float *d_img, uint64_t npix, double *h_sum3, void *stream);

/* ---- path tracing ---------------------------------------------------------
 * A tracer owns the wavefront path state (SoA queues in HBM) and grows it on
 * demand. */
int ct_tracer_create(ct_tracer **out);
int ct_tracer_destroy(ct_tracer *tracer);
/* Device-time profiling of each launch (HIP events on the launch stream). */
int ct_tracer_set_profiling(ct_tracer *tracer, int32_t enable);
/* Per-ray node/primitive counters in the traversal kernels (slower; for the
 * roofline's algorithmic byte count, never for timing). */
int ct_tracer_set_counting(ct_tracer *tracer, int32_t enable);
/* Closest-hit walk used by the extension rays (CT_WALK_*). */
int ct_tracer_set_walk(ct_tracer *tracer, int32_t walk);
/* Walk of the NEE any-hit rays (CT_WALK_*; the result does not depend on it)
 * and the LDS stack depth of the ordered walks (8, 16 or 24 entries;
 * deeper rays finish with the stackless walk, same result). */
int ct_tracer_set_traversal(ct_tracer *tracer, int32_t shadowWalk, int32_t stackDepth);
/* Launch schedule of the traversal kernels: persistent = 1 runs `blocks`
 * workgroups that pull rays from the queue as lanes free up;
 * 0 launches one lane per path.  Results do not depend on it.  Until it is
 * set, persistent launches of 640 workgroups (binary64) or 768 (FP32). */
int ct_tracer_set_schedule(ct_tracer *tracer, int32_t persistent, int32_t blocks);
/* accumulatePass (tracer.ts:361-395): adds one sample per pixel into the
 * float32 sum d_accum (width*height*3) with frameIndex as the RNG frame. */
int ct_accumulate_pass(ct_tracer *tracer, const ct_scene *scene, const double *h_frame16,
                       uint32_t width, uint32_t height, uint32_t frameIndex,
                       const ct_trace_opts *opts, float *d_accum, void *stream);
/* renderImage (tracer.ts:348-355): spp passes with frameIndex = s + 7919*seed,
 * then d_out *= 1/spp.  d_out is width*height*3 float32 (overwritten); with
 * spp 0 every value is NaN, as the reference's zero sums times 1/0. */
int ct_render_image(ct_tracer *tracer, const ct_scene *scene, const double *h_frame16,
                    const ct_render_opts *opts, float *d_out, void *stream);
/* radianceInto (tracer.ts:191-328) for n rays at once: ray i (6 doubles at
 * d_rays + 6i: origin, direction) is traced with the xorshift32 state
 * d_rngStates[i] (the Rng passed to radianceInto), through at most
 * opts->maxBounces bounces; its radiance (3 doubles at d_out + 3i, as the
 * reference returns it: no finiteness filter) is written and d_rngStates[i]
 * left as the Rng is after the call.  Binary64 arithmetic only (CT_EINVAL
 * under CT_PRECISION_F32); maxBounces above 254: CT_ELIMIT. */
int ct_radiance(ct_tracer *tracer, const ct_scene *scene, const double *d_rays, uint32_t n,
                uint32_t *d_rngStates, const ct_trace_opts *opts, double *d_out, void *stream);
/* Several scenes / cameras / sizes / depths traced in shared launches: the
 * paths of all jobs share every extend, shade and shadow launch (rounds of at
 * most the tracer's path capacity).  Each output equals what the job alone
 * produces (each pixel's samples are added in pass order per job).  Not part of
 * the reference API: it serves the batching and serving forms. */
int ct_render_batch(ct_tracer *tracer, const ct_job *jobs, uint32_t n, void *stream);
int ct_accumulate_batch(ct_tracer *tracer, const ct_pass_job *jobs, uint32_t n, void *stream);
/* Paths in flight per round (path state is about 190 B per path, allocated as
 * rounds need it; default 2^26). */
int ct_tracer_set_capacity(ct_tracer *tracer, uint64_t maxPaths);
/* Capture each round's launches into a hipGraph and replay it while the
 * round's shape (paths, pixels, jobs, bounces, settings) repeats. */
int ct_tracer_set_graph(ct_tracer *tracer, int32_t enable);
/* NEE launches of each path group on a side stream beside the next bounce's
 * extension launch (1, default) or in order on the group's stream (0).  With
 * one path group and no overlap every launch of a round runs alone, which
 * per-launch profiling uses; results never depend on it. */
int ct_tracer_set_overlap(ct_tracer *tracer, int32_t enable);
/* How a binary64 round is traced: CT_PATHS_WAVEFRONT runs the launches above
 * (per bounce an extension, a shading and an NEE launch per path group);
 * CT_PATHS_PER_PATH runs one launch per path group in which each lane carries
 * one path through all its bounces (the skip-link walks of intersect.ts:
 * 290-420 for the closest hit and the NEE test, the same shading), which
 * suits small scenes, where a bounce's launches cost more than the work they
 * carry; CT_PATHS_AUTO (default) takes the per-path launch for rounds of at
 * most 2^24 paths whose scenes have at most 32768 primitives each (and, not
 * in graph mode, packs such jobs into rounds of their own).  The per-path launch is
 * persistent (a lane takes the next path when its own ends); its walks are
 * the 4-wide ones (LDS stack, leaves tested as reached), or the skip-link
 * walks when the tracer's walk is CT_WALK_STACKLESS.
 * Results are identical either way.
 * The FP32 precision form always runs its wavefront.  With the per-path
 * launch, a round's profiling events and extendLaunches count the per-path
 * launches (no shading or NEE launches are recorded).  No counterpart in the
 * reference, which traces one path at a time (tracer.ts:361-395). */
#define CT_PATHS_AUTO 0
#define CT_PATHS_WAVEFRONT 1
#define CT_PATHS_PER_PATH 2
int ct_tracer_set_path_kernel(ct_tracer *tracer, int32_t mode);
/* The mode set, and how many rounds of the last call ran the per-path launch. */
int ct_tracer_get_path_kernel(const ct_tracer *tracer, int32_t *mode, uint32_t *perPathRounds);
/* CT_PRECISION_F64 or CT_PRECISION_F32 (which always walks the quantized
 * 4-wide nodes; the walk settings then only pick the stack depth). */
int ct_tracer_set_precision(ct_tracer *tracer, int32_t precision);
/* Lanes per workgroup of the traversal launches: 64, 128 or 256 (default). */
int ct_tracer_set_workgroup(ct_tracer *tracer, int32_t lanes);
/* Counters and device times of the last accumulate/render (synchronises). */
int ct_tracer_stats(ct_tracer *tracer, ct_render_stats *out);

/* Every launch setting of a tracer in one record (none changes any result). */
typedef struct ct_tracer_config {
  uint64_t maxPaths;  /* paths in flight per round */
  int32_t walk;       /* closest-hit walk of extension rays (CT_WALK_*) */
  int32_t shadowWalk; /* any-hit walk of NEE rays (CT_WALK_*) */
  int32_t stackDepth; /* LDS stack entries of the ordered walks: 8, 16, 24 */
  int32_t persistent; /* 1: persistent traversal launches pulling from the queue */
  int32_t blocks;     /* workgroups of the persistent launches */
  int32_t workgroup;  /* lanes per traversal workgroup: 64, 128, 256 */
  int32_t pathGroups; /* 1 or 2: a round's paths traced as that many groups, each on its
                         own stream (its NEE launches on a side stream), so that one
                         group's shading overlaps the other's traversal (default 2) */
} ct_tracer_config;
int ct_tracer_get_config(const ct_tracer *tracer, ct_tracer_config *out);
int ct_tracer_set_config(ct_tracer *tracer, const ct_tracer_config *config);

/* ---- autotuner --------------------------------------------------------------
 * Picks the tracer's launch settings for a batch of jobs by timing
 * ct_render_batch on the batch itself: coordinate descent over the queue
 * capacity (paths per round), the traversal policy (ordered walk with an LDS
 * stack of 8/16/24 entries that restarts a deep ray with the stackless walk,
 * or the stackless walk alone; binary or 4-wide nodes), the workgroup size
 * (64/128/256 lanes), the lanes kept in flight by the persistent launches and
 * the number of path groups traced side by side.
 * Every trial renders the jobs, so on return the outputs hold the batch's
 * images (identical under every setting).  The best setting is applied to
 * the tracer.  With cachePath set, a file of earlier results keyed by the
 * batch's signature (device, job count, path and primitive counts, depth) is
 * read first; a hit applies it without trials, a miss appends the result.
 * No counterpart in the reference (it has one configuration). */
typedef struct ct_tune_result {
  ct_tracer_config config; /* chosen setting (applied) */
  double ms;               /* its time for one ct_render_batch of the jobs */
  double initialMs;        /* time of the setting the tracer had on entry */
  uint32_t trials;         /* settings timed (0 when taken from the cache) */
  int32_t fromCache;
} ct_tune_result;
int ct_tracer_autotune(ct_tracer *tracer, const ct_job *jobs, uint32_t n, uint32_t maxTrials,
                       const char *cachePath, ct_tune_result *out, void *stream);

/* ---- serving -------------------------------------------------------------
 * A request loop over resident scenes: render jobs (ct_job, d_out ignored)
 * are submitted as they arrive and traced by the server's own thread with
 * continuous batching -- every launch round takes passes from all admitted
 * jobs, and jobs that arrive meanwhile join at the next round once a slot and
 * HBM for their progressive buffer are free.  The progressive buffer (the
 * job's float32 sum, width*height*3) stays on the device until released;
 * when the last pass is added it holds renderImage's result (the sum times
 * 1/spp), bit-identical to ct_render_image of the job alone.  Scenes must
 * outlive the jobs that use them.  All functions may be called from any
 * thread.  The reference has no serving layer; its boundary (renderImage)
 * is what each job computes. */
#define CT_JOB_QUEUED 0  /* waiting for a slot or HBM */
#define CT_JOB_RUNNING 1 /* admitted; passes in progress */
#define CT_JOB_DONE 2    /* buffer holds the final image */
#define CT_JOB_FAILED 3  /* device error (see ct_last_error after ct_job_poll) */

#define CT_SERVE_FIFO 0 /* rounds filled in admission order */
#define CT_SERVE_SRPT 1 /* rounds filled shortest remaining work first */
#define CT_SERVE_SCENE 2 /* admission order, each scene's jobs together: a round
                            takes the oldest job's scene's jobs, then the next
                            scene's, so it touches fewer scenes' trees */

typedef struct ct_server_opts {
  uint64_t hbmBudget;   /* bytes of progressive buffers resident at once (0: 4 GiB) */
  uint64_t maxPaths;    /* paths per launch round (0: the tracer default, 2^26) */
  uint32_t maxActive;   /* admitted jobs at most (0: 1024) */
  uint32_t slicePasses; /* passes one job adds per round at most (0: as many as fit) */
  int32_t policy;       /* CT_SERVE_FIFO, CT_SERVE_SRPT or CT_SERVE_SCENE */
  int32_t graph;        /* replay rounds of a repeated shape as hipGraphs */
  int32_t precision;    /* CT_PRECISION_F64 or CT_PRECISION_F32 (compensation terms
                           double each job's buffer footprint) */
} ct_server_opts;

typedef struct ct_job_status {
  int32_t state;        /* CT_JOB_* */
  uint32_t passesDone;  /* passes added to the buffer */
  uint32_t spp;
  double submitSec;     /* server clock (seconds since ct_server_create) */
  double admitSec;      /* -1 while queued */
  double doneSec;       /* -1 until done */
} ct_job_status;

typedef struct ct_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t queued, active;       /* now */
  uint64_t rounds;               /* launch rounds traced */
  uint64_t paths;                /* camera paths traced */
  uint64_t extensionRays, shadowRays;
  uint64_t pixelSamples;         /* width*height*passes summed over rounds */
  uint64_t hbmInUse;             /* bytes of progressive buffers held */
  double busySec;                /* time spent inside launch rounds */
  double nowSec;                 /* server clock */
} ct_server_stats;

int ct_server_create(const ct_server_opts *opts, ct_server **out);
/* Stops the loop after the round in flight; unfinished jobs are dropped. */
int ct_server_destroy(ct_server *server);
int ct_job_submit(ct_server *server, const ct_job *job, uint64_t *id);
int ct_job_poll(ct_server *server, uint64_t id, ct_job_status *status);
/* Blocks until the job is done or failed, or timeoutSec passes (CT_ETIMEOUT). */
int ct_job_wait(ct_server *server, uint64_t id, double timeoutSec, ct_job_status *status);
/* Copies the job's buffer (width*height*3 float32): the final image once
 * done, the running sum of passesDone passes before.  dstOnDevice selects a
 * device or host destination. */
int ct_job_read(ct_server *server, uint64_t id, float *dst, int32_t dstOnDevice);
/* Frees the job's buffer and forgets it; a queued job is cancelled.  A
 * running job is released after its round. */
int ct_job_release(ct_server *server, uint64_t id);
int ct_server_stats_get(ct_server *server, ct_server_stats *out);

/* ---- display ----------------------------------------------------------------
 * tonemapBuffer (tonemap.ts:41-62) on the device: each channel of the float32
 * sum (width*height*3) times 2^exposureEv / samples, tone mapped (ACES fit or
 * Reinhard), sRGB encoded, rounded to 8 bits; alpha 255.  d_rgba holds
 * width*height*4 bytes.  Pair with a host P6 writer (ppm.ts:4-16). */
#define CT_TONEMAP_ACES 0
#define CT_TONEMAP_REINHARD 1
int ct_tonemap(const float *d_accum, uint32_t width, uint32_t height, double samples,
               double exposureEv, int32_t mode, uint8_t *d_rgba, void *stream);
/* A served job's frame as 8-bit RGBA (width*height*4 bytes): the final image
 * once done, the progressive average of the passes added so far before. */
int ct_job_read_rgba(ct_server *server, uint64_t id, double exposureEv, int32_t mode,
                     uint8_t *dst, int32_t dstOnDevice);

/* ---- BVH build --------------------------------------------------------------
 * buildBvh (bvh.ts:299-301) on the GPU from device-resident PrimBounds
 * (bvh.ts:22-28: 3 binary64 per primitive for min, max and centroid).
 * Writes nodes (room for max(1, 2*count-1)*8 floats), *nodeCount (host) and
 * order (count uint32) bit-identical to the reference builder.  bins is the
 * SAH bucket count (12 in the reference, bvh.ts:20). */
int ct_bvh_build(const double *d_min, const double *d_max, const double *d_centroid,
                 uint32_t count, int32_t bins, float *d_nodes, uint32_t *h_nodeCount,
                 uint32_t *d_order, void *stream);
/* ct_bvh_build plus the exact leaf word of every node in d_leafWords
 * (room for max(1, 2*count-1) uint32; see ct_scene_create_ex). */
int ct_bvh_build_ex(const double *d_min, const double *d_max, const double *d_centroid,
                    uint32_t count, int32_t bins, float *d_nodes, uint32_t *h_nodeCount,
                    uint32_t *d_order, uint32_t *d_leafWords, void *stream);

/* ct_bvh_build_ex without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops that must not call
 * hipMalloc, harnesses that forbid it).  No counterpart in the reference.
 * The build's working memory is a static device array of
 * CT_BVH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: at 12
 * bins that is room for about 83,000 primitives (the call returns CT_ELIMIT
 * with the bytes it needs when count is larger).  Every call on one GPU
 * shares that array: calls that overlap on one GPU race, so the caller
 * serialises them (the call synchronises its stream before returning). */
#define CT_BVH_NOALLOC_SCRATCH_BYTES (64u << 20)
int ct_bvh_build_noalloc(const double *d_min, const double *d_max, const double *d_centroid,
                         uint32_t count, int32_t bins, float *d_nodes, uint32_t *h_nodeCount,
                         uint32_t *d_order, uint32_t *d_leafWords, void *stream);

#ifdef __cplusplus
}
#endif
#endif
