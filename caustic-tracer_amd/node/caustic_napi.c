/*
 * N-API addon: the Node/TypeScript side of the C ABI (include/caustic_amd.h).
 *
 * Each export takes the same typed arrays the reference's packer and tracer
 * use (pack.ts:46-59, tracer.ts:337-395) and forwards them to libcaustic_amd.so.
 * Bulk arrays cross as host typed arrays; the addon stages them through
 * device buffers.  A non-zero status from the library becomes a thrown Error
 * carrying ct_last_error().
 *
 *   createScene(prims, primCount, nodes, nodeCount, mats, matCount, lights,
 *               primObject, sky?) -> scene handle            (PackedScene upload)
 *   buildBvh(min, max, centroid, bins) -> {nodes, nodeCount, order}   (bvh.ts:299)
 *   closestHit(scene, rays, tMin, tMax|null, walk)
 *       -> {prim: Int32Array, t, u, v: Float64Array}          (intersect.ts:202/290)
 *   occluded(scene, rays, tMin, tMax) -> Uint8Array            (intersect.ts:368)
 *   accumulatePass(scene, frame16, width, height, frameIndex, maxBounces,
 *                  russianRoulette, nee, accum)               (tracer.ts:361)
 *   renderImage(scene, frame16, width, height, spp, seed, maxBounces,
 *               russianRoulette, nee) -> Float32Array          (tracer.ts:348)
 *   radiance(scene, rays Float64Array(6n), states Uint32Array(n), maxBounces,
 *            russianRoulette, nee) -> Float64Array(3n); states advanced in
 *            place                                              (tracer.ts:191)
 *   setPrecision(0|1)            arithmetic of the two calls above
 *   setPathKernel(0|1|2)         auto / wavefront / per-path launches
 *   createServer({hbmBudget, maxPaths, maxActive, slicePasses, policy,
 *                 precision}) -> server handle                 (serving loop)
 *   submitJob(server, scene, frame16, width, height, spp, seed, maxBounces,
 *             russianRoulette, nee) -> job id
 *   pollJob(server, id) -> {state, passesDone, spp, submitSec, admitSec, doneSec}
 *   waitJob(server, id) -> Promise of the same status (N-API async work: the
 *             wait runs on a libuv worker thread, the event loop stays free)
 *   readJob(server, id, width, height) -> Float32Array (final image once done)
 *   releaseJob(server, id); serverStats(server) -> counters
 *   tonemap(accum, width, height, samples, exposureEv, mode)
 *       -> Uint8ClampedArray RGBA                              (tonemap.ts:41)
 */
#define NAPI_VERSION 4
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/caustic_amd.h"

#define CALL(env, expr)                                               \
  do {                                                                \
    if ((expr) != napi_ok) {                                          \
      napi_throw_error((env), NULL, "caustic_amd: N-API call failed: " #expr); \
      return NULL;                                                    \
    }                                                                 \
  } while (0)

static napi_value throw_ct(napi_env env, int rc, const char *what) {
  char msg[512];
  snprintf(msg, sizeof msg, "%s failed (%d): %s", what, rc, ct_last_error());
  napi_throw_error(env, NULL, msg);
  return NULL;
}

static int g_init = 0;
static int ensure_init(napi_env env) {
  if (g_init) return 1;
  int rc = ct_init(0);
  if (rc) {
    throw_ct(env, rc, "ct_init");
    return 0;
  }
  g_init = 1;
  return 1;
}

/* typed-array view: data pointer and element count (NULL/0 for null/undefined) */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, void **data,
                 size_t *len) {
  napi_valuetype t;
  *data = NULL;
  *len = 0;
  if (napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_null || t == napi_undefined) return 1;
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, NULL, "caustic_amd: expected a typed array");
    return 0;
  }
  napi_typedarray_type ty;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &ty, len, data, &ab, &off) != napi_ok) return 0;
  if (ty != want) {
    napi_throw_type_error(env, NULL, "caustic_amd: typed array of the wrong element type");
    return 0;
  }
  return 1;
}

static double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

static void *to_device(const void *h, size_t bytes, int *rc) {
  void *d = NULL;
  *rc = ct_device_alloc(&d, bytes);
  if (*rc) return NULL;
  if (bytes) *rc = ct_copy_h2d(d, h, bytes);
  if (*rc) {
    ct_device_free(d);
    return NULL;
  }
  return d;
}

static napi_value new_typed(napi_env env, napi_typedarray_type ty, size_t elems, size_t esize,
                            void **data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, elems * esize, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, ty, elems, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

/* ---- scene handle ----------------------------------------------------- */
static void scene_finalize(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  ct_scene_destroy((ct_scene *)data);
}

static napi_value create_scene(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value a[9];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (argc < 8) {
    napi_throw_error(env, NULL, "createScene: 8 arguments required");
    return NULL;
  }
  if (!ensure_init(env)) return NULL;
  void *prims, *nodes, *mats, *lights, *pobj;
  size_t np_, nn, nm, nl, no;
  if (!typed(env, a[0], napi_float32_array, &prims, &np_)) return NULL;
  uint32_t primCount = (uint32_t)num(env, a[1]);
  if (!typed(env, a[2], napi_float32_array, &nodes, &nn)) return NULL;
  uint32_t nodeCount = (uint32_t)num(env, a[3]);
  if (!typed(env, a[4], napi_float32_array, &mats, &nm)) return NULL;
  uint32_t matCount = (uint32_t)num(env, a[5]);
  if (!typed(env, a[6], napi_int32_array, &lights, &nl)) return NULL;
  if (!typed(env, a[7], napi_uint32_array, &pobj, &no)) return NULL;
  if (np_ < (size_t)primCount * 16 || nn < (size_t)nodeCount * 8 || nm < (size_t)matCount * 16 ||
      no < primCount) {
    napi_throw_range_error(env, NULL, "createScene: array shorter than its count");
    return NULL;
  }
  ct_sky sky;
  memset(&sky, 0, sizeof sky);
  if (argc > 8) {
    napi_valuetype t;
    napi_typeof(env, a[8], &t);
    if (t == napi_object) {
      /* {kind: 'gradient', zenith:{r,g,b}, horizon:{r,g,b}, strength} (scene.ts:57-59) */
      napi_value kind, z, h, s, c;
      char buf[16] = {0};
      size_t l;
      if (napi_get_named_property(env, a[8], "kind", &kind) == napi_ok &&
          napi_get_value_string_utf8(env, kind, buf, sizeof buf, &l) == napi_ok &&
          strcmp(buf, "gradient") == 0) {
        static const char *ch[3] = {"r", "g", "b"};
        sky.kind = 1;
        CALL(env, napi_get_named_property(env, a[8], "zenith", &z));
        CALL(env, napi_get_named_property(env, a[8], "horizon", &h));
        CALL(env, napi_get_named_property(env, a[8], "strength", &s));
        for (int k = 0; k < 3; k++) {
          CALL(env, napi_get_named_property(env, z, ch[k], &c));
          sky.zenith[k] = num(env, c);
          CALL(env, napi_get_named_property(env, h, ch[k], &c));
          sky.horizon[k] = num(env, c);
        }
        sky.strength = num(env, s);
      }
    }
  }
  ct_scene *scene = NULL;
  int rc = ct_scene_create((const float *)prims, primCount, (const float *)nodes, nodeCount,
                           (const float *)mats, matCount, (const int32_t *)lights, (uint32_t)nl,
                           (const uint32_t *)pobj, &sky, &scene);
  if (rc) return throw_ct(env, rc, "ct_scene_create");
  napi_value ext;
  CALL(env, napi_create_external(env, scene, scene_finalize, NULL, &ext));
  return ext;
}

static ct_scene *get_scene(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "caustic_amd: not a scene handle");
    return NULL;
  }
  return (ct_scene *)p;
}

/* ---- buildBvh ----------------------------------------------------------- */
static napi_value build_bvh(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (!ensure_init(env)) return NULL;
  void *mn, *mx, *cn;
  size_t l0, l1, l2;
  if (!typed(env, a[0], napi_float64_array, &mn, &l0) ||
      !typed(env, a[1], napi_float64_array, &mx, &l1) ||
      !typed(env, a[2], napi_float64_array, &cn, &l2))
    return NULL;
  if (l0 != l1 || l1 != l2 || l0 % 3) {
    napi_throw_range_error(env, NULL, "buildBvh: PrimBounds arrays must be 3*count long");
    return NULL;
  }
  int bins = argc > 3 ? (int)num(env, a[3]) : 12;
  uint32_t n = (uint32_t)(l0 / 3);
  size_t maxNodes = n ? 2 * (size_t)n - 1 : 1;
  int rc = 0;
  void *dmn = to_device(mn, l0 * 8, &rc), *dmx = NULL, *dcn = NULL, *dnodes = NULL, *dord = NULL;
  if (!rc) dmx = to_device(mx, l1 * 8, &rc);
  if (!rc) dcn = to_device(cn, l2 * 8, &rc);
  if (!rc) rc = ct_device_alloc(&dnodes, maxNodes * 32);
  if (!rc) rc = ct_device_alloc(&dord, (n ? n : 1) * 4);
  uint32_t nodeCount = 0;
  if (!rc) rc = ct_bvh_build((const double *)dmn, (const double *)dmx, (const double *)dcn, n,
                             bins, (float *)dnodes, &nodeCount, (uint32_t *)dord, NULL);
  napi_value res = NULL;
  if (!rc) {
    void *hn, *ho;
    napi_value nodes = new_typed(env, napi_float32_array, (size_t)nodeCount * 8, 4, &hn);
    napi_value order = new_typed(env, napi_uint32_array, n, 4, &ho);
    rc = ct_copy_d2h(hn, dnodes, (size_t)nodeCount * 32);
    if (!rc && n) rc = ct_copy_d2h(ho, dord, (size_t)n * 4);
    if (!rc) {
      napi_value nc;
      napi_create_object(env, &res);
      napi_create_uint32(env, nodeCount, &nc);
      napi_set_named_property(env, res, "nodes", nodes);
      napi_set_named_property(env, res, "nodeCount", nc);
      napi_set_named_property(env, res, "order", order);
    }
  }
  ct_device_free(dmn);
  ct_device_free(dmx);
  ct_device_free(dcn);
  ct_device_free(dnodes);
  ct_device_free(dord);
  if (rc) return throw_ct(env, rc, "ct_bvh_build");
  return res;
}

/* ---- closestHit / occluded --------------------------------------------- */
static napi_value closest_hit(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value a[5];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_scene *s = get_scene(env, a[0]);
  if (!s) return NULL;
  void *rays, *tmax;
  size_t nr, nt;
  if (!typed(env, a[1], napi_float64_array, &rays, &nr)) return NULL;
  double tMin = num(env, a[2]);
  if (!typed(env, a[3], napi_float64_array, &tmax, &nt)) return NULL;
  int walk = argc > 4 ? (int)num(env, a[4]) : CT_WALK_ORDERED;
  uint32_t n = (uint32_t)(nr / 6);
  if (tmax && nt < n) {
    napi_throw_range_error(env, NULL, "closestHit: tMax shorter than the ray count");
    return NULL;
  }
  int rc = 0;
  void *dr = to_device(rays, (size_t)n * 48, &rc), *dt = NULL, *dp = NULL, *dT = NULL,
       *dU = NULL, *dV = NULL;
  if (!rc && tmax) dt = to_device(tmax, (size_t)n * 8, &rc);
  if (!rc) rc = ct_device_alloc(&dp, (size_t)n * 4);
  if (!rc) rc = ct_device_alloc(&dT, (size_t)n * 8);
  if (!rc) rc = ct_device_alloc(&dU, (size_t)n * 8);
  if (!rc) rc = ct_device_alloc(&dV, (size_t)n * 8);
  if (!rc) rc = ct_closest_hit(s, (const double *)dr, n, tMin, (const double *)dt, walk,
                               (int32_t *)dp, (double *)dT, (double *)dU, (double *)dV, NULL);
  napi_value res = NULL;
  if (!rc) {
    void *hp, *hT, *hU, *hV;
    napi_value P = new_typed(env, napi_int32_array, n, 4, &hp);
    napi_value T = new_typed(env, napi_float64_array, n, 8, &hT);
    napi_value U = new_typed(env, napi_float64_array, n, 8, &hU);
    napi_value V = new_typed(env, napi_float64_array, n, 8, &hV);
    if (n) {
      rc = ct_copy_d2h(hp, dp, (size_t)n * 4);
      if (!rc) rc = ct_copy_d2h(hT, dT, (size_t)n * 8);
      if (!rc) rc = ct_copy_d2h(hU, dU, (size_t)n * 8);
      if (!rc) rc = ct_copy_d2h(hV, dV, (size_t)n * 8);
    }
    napi_create_object(env, &res);
    napi_set_named_property(env, res, "prim", P);
    napi_set_named_property(env, res, "t", T);
    napi_set_named_property(env, res, "u", U);
    napi_set_named_property(env, res, "v", V);
  }
  ct_device_free(dr);
  ct_device_free(dt);
  ct_device_free(dp);
  ct_device_free(dT);
  ct_device_free(dU);
  ct_device_free(dV);
  if (rc) return throw_ct(env, rc, "ct_closest_hit");
  return res;
}

static napi_value occluded(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_scene *s = get_scene(env, a[0]);
  if (!s) return NULL;
  void *rays, *tmax;
  size_t nr, nt;
  if (!typed(env, a[1], napi_float64_array, &rays, &nr)) return NULL;
  double tMin = num(env, a[2]);
  if (!typed(env, a[3], napi_float64_array, &tmax, &nt)) return NULL;
  uint32_t n = (uint32_t)(nr / 6);
  if (!tmax || nt < n) {
    napi_throw_range_error(env, NULL, "occluded: one tMax per ray required");
    return NULL;
  }
  int rc = 0;
  void *dr = to_device(rays, (size_t)n * 48, &rc), *dt = NULL, *dout = NULL;
  if (!rc) dt = to_device(tmax, (size_t)n * 8, &rc);
  if (!rc) rc = ct_device_alloc(&dout, n ? n : 1);
  if (!rc) rc = ct_occluded(s, (const double *)dr, n, tMin, (const double *)dt, (uint8_t *)dout, NULL);
  napi_value res = NULL;
  if (!rc) {
    void *h;
    res = new_typed(env, napi_uint8_array, n, 1, &h);
    if (n) rc = ct_copy_d2h(h, dout, n);
  }
  ct_device_free(dr);
  ct_device_free(dt);
  ct_device_free(dout);
  if (rc) return throw_ct(env, rc, "ct_occluded");
  return res;
}

/* ---- accumulatePass / renderImage ---------------------------------------- */
static ct_tracer *g_tracer = NULL;

static napi_value accumulate_pass(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value a[9];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_scene *s = get_scene(env, a[0]);
  if (!s) return NULL;
  void *frame, *accum;
  size_t nf, na;
  if (!typed(env, a[1], napi_float64_array, &frame, &nf)) return NULL;
  uint32_t w = (uint32_t)num(env, a[2]), h = (uint32_t)num(env, a[3]);
  uint32_t frameIndex = (uint32_t)(int64_t)num(env, a[4]);
  ct_trace_opts o = {(int32_t)num(env, a[5]), (int32_t)num(env, a[6]), (int32_t)num(env, a[7])};
  if (!typed(env, a[8], napi_float32_array, &accum, &na)) return NULL;
  if (nf < 16 || na < (size_t)w * h * 3) {
    napi_throw_range_error(env, NULL, "accumulatePass: frame needs 16 values, accum w*h*3");
    return NULL;
  }
  int rc = 0;
  if (!g_tracer) rc = ct_tracer_create(&g_tracer);
  void *d = NULL;
  if (!rc) d = to_device(accum, (size_t)w * h * 12, &rc);
  if (!rc) rc = ct_accumulate_pass(g_tracer, s, (const double *)frame, w, h, frameIndex, &o,
                                   (float *)d, NULL);
  if (!rc) rc = ct_copy_d2h(accum, d, (size_t)w * h * 12);
  ct_device_free(d);
  if (rc) return throw_ct(env, rc, "ct_accumulate_pass");
  return NULL;
}

static napi_value render_image(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value a[9];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_scene *s = get_scene(env, a[0]);
  if (!s) return NULL;
  void *frame;
  size_t nf;
  if (!typed(env, a[1], napi_float64_array, &frame, &nf)) return NULL;
  if (nf < 16) {
    napi_throw_range_error(env, NULL, "renderImage: frame needs 16 values");
    return NULL;
  }
  ct_render_opts o;
  o.width = (uint32_t)num(env, a[2]);
  o.height = (uint32_t)num(env, a[3]);
  o.spp = (uint32_t)num(env, a[4]);
  o.seed = (uint32_t)(int64_t)num(env, a[5]);
  o.maxBounces = (int32_t)num(env, a[6]);
  o.russianRoulette = (int32_t)num(env, a[7]);
  o.nee = (int32_t)num(env, a[8]);
  size_t n = (size_t)o.width * o.height * 3;
  int rc = 0;
  if (!g_tracer) rc = ct_tracer_create(&g_tracer);
  void *d = NULL;
  if (!rc) rc = ct_device_alloc(&d, n * 4);
  if (!rc) rc = ct_render_image(g_tracer, s, (const double *)frame, &o, (float *)d, NULL);
  napi_value res = NULL;
  if (!rc) {
    void *h;
    res = new_typed(env, napi_float32_array, n, 4, &h);
    rc = ct_copy_d2h(h, d, n * 4);
  }
  ct_device_free(d);
  if (rc) return throw_ct(env, rc, "ct_render_image");
  return res;
}

/* radiance(scene, rays Float64Array(6n), states Uint32Array(n), maxBounces, rr, nee)
 * -> Float64Array(3n); the states are updated in place (radianceInto's Rng). */
static napi_value radiance(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_scene *s = get_scene(env, a[0]);
  if (!s) return NULL;
  void *rays, *states;
  size_t nr, ns;
  if (!typed(env, a[1], napi_float64_array, &rays, &nr)) return NULL;
  if (!typed(env, a[2], napi_uint32_array, &states, &ns)) return NULL;
  if (nr % 6 != 0 || nr / 6 != ns || ns > 0xffffffffu) {
    napi_throw_range_error(env, NULL, "radiance: rays needs 6 values per state");
    return NULL;
  }
  const uint32_t n = (uint32_t)ns;
  ct_trace_opts o = {(int32_t)num(env, a[3]), (int32_t)num(env, a[4]), (int32_t)num(env, a[5])};
  void *h;
  napi_value res = new_typed(env, napi_float64_array, (size_t)n * 3, 8, &h);
  if (!res || n == 0) return res;
  int rc = 0;
  if (!g_tracer) rc = ct_tracer_create(&g_tracer);
  void *dr = NULL, *ds = NULL, *dout = NULL;
  if (!rc) dr = to_device(rays, (size_t)n * 48, &rc);
  if (!rc) ds = to_device(states, (size_t)n * 4, &rc);
  if (!rc) rc = ct_device_alloc(&dout, (size_t)n * 24);
  if (!rc) rc = ct_radiance(g_tracer, s, (const double *)dr, n, (uint32_t *)ds, &o,
                            (double *)dout, NULL);
  if (!rc) rc = ct_copy_d2h(h, dout, (size_t)n * 24);
  if (!rc) rc = ct_copy_d2h(states, ds, (size_t)n * 4);
  ct_device_free(dr);
  ct_device_free(ds);
  ct_device_free(dout);
  if (rc) return throw_ct(env, rc, "ct_radiance");
  return res;
}

static napi_value set_path_kernel(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (!ensure_init(env)) return NULL;
  int rc = 0;
  if (!g_tracer) rc = ct_tracer_create(&g_tracer);
  if (!rc) rc = ct_tracer_set_path_kernel(g_tracer, (int32_t)num(env, a[0]));
  if (rc) return throw_ct(env, rc, "ct_tracer_set_path_kernel");
  return NULL;
}

static napi_value set_precision(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (!ensure_init(env)) return NULL;
  int rc = 0;
  if (!g_tracer) rc = ct_tracer_create(&g_tracer);
  if (!rc) rc = ct_tracer_set_precision(g_tracer, (int32_t)num(env, a[0]));
  if (rc) return throw_ct(env, rc, "ct_tracer_set_precision");
  return NULL;
}

/* ---- serving ------------------------------------------------------------ */
static void server_finalize(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  ct_server_destroy((ct_server *)data);
}

static ct_server *get_server(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "caustic_amd: expected a server handle");
    return NULL;
  }
  return (ct_server *)p;
}

static double prop_num(napi_env env, napi_value obj, const char *key, double dflt) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_number ? num(env, v) : dflt;
}

static napi_value create_server(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (!ensure_init(env)) return NULL;
  ct_server_opts o;
  memset(&o, 0, sizeof o);
  napi_valuetype t = napi_undefined;
  if (argc >= 1) napi_typeof(env, a[0], &t);
  if (t == napi_object) {
    o.hbmBudget = (uint64_t)prop_num(env, a[0], "hbmBudget", 0);
    o.maxPaths = (uint64_t)prop_num(env, a[0], "maxPaths", 0);
    o.maxActive = (uint32_t)prop_num(env, a[0], "maxActive", 0);
    o.slicePasses = (uint32_t)prop_num(env, a[0], "slicePasses", 0);
    o.policy = (int32_t)prop_num(env, a[0], "policy", 0);
    o.precision = (int32_t)prop_num(env, a[0], "precision", 0);
  }
  ct_server *srv = NULL;
  int rc = ct_server_create(&o, &srv);
  if (rc) return throw_ct(env, rc, "ct_server_create");
  napi_value ext;
  CALL(env, napi_create_external(env, srv, server_finalize, NULL, &ext));
  return ext;
}

static napi_value submit_job(napi_env env, napi_callback_info info) {
  size_t argc = 10;
  napi_value a[10];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (argc < 10) {
    napi_throw_error(env, NULL, "submitJob: 10 arguments required");
    return NULL;
  }
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  ct_scene *s = get_scene(env, a[1]);
  if (!s) return NULL;
  void *frame;
  size_t nf;
  if (!typed(env, a[2], napi_float64_array, &frame, &nf)) return NULL;
  if (nf < 16) {
    napi_throw_range_error(env, NULL, "submitJob: frame needs 16 values");
    return NULL;
  }
  ct_job j;
  memset(&j, 0, sizeof j);
  j.scene = s;
  memcpy(j.frame, frame, sizeof j.frame);
  j.width = (uint32_t)num(env, a[3]);
  j.height = (uint32_t)num(env, a[4]);
  j.spp = (uint32_t)num(env, a[5]);
  j.seed = (uint32_t)(int64_t)num(env, a[6]);
  j.maxBounces = (int32_t)num(env, a[7]);
  j.russianRoulette = (int32_t)num(env, a[8]);
  j.nee = (int32_t)num(env, a[9]);
  uint64_t id = 0;
  int rc = ct_job_submit(srv, &j, &id);
  if (rc) return throw_ct(env, rc, "ct_job_submit");
  napi_value v;
  CALL(env, napi_create_double(env, (double)id, &v));
  return v;
}

static napi_value status_object(napi_env env, const ct_job_status *st) {
  napi_value o, v;
  if (napi_create_object(env, &o) != napi_ok) return NULL;
#define SET(name, val)                                        \
  if (napi_create_double(env, (double)(val), &v) != napi_ok || \
      napi_set_named_property(env, o, name, v) != napi_ok)    \
    return NULL;
  SET("state", st->state)
  SET("passesDone", st->passesDone)
  SET("spp", st->spp)
  SET("submitSec", st->submitSec)
  SET("admitSec", st->admitSec)
  SET("doneSec", st->doneSec)
#undef SET
  return o;
}

static napi_value poll_job(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  ct_job_status st;
  int rc = ct_job_poll(srv, (uint64_t)num(env, a[1]), &st);
  if (rc) return throw_ct(env, rc, "ct_job_poll");
  return status_object(env, &st);
}

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  ct_server *srv;
  uint64_t id;
  ct_job_status st;
  int rc;
  char err[512];
} WaitWork;

static void wait_execute(napi_env env, void *data) {
  (void)env;
  WaitWork *w = (WaitWork *)data;
  w->rc = ct_job_wait(w->srv, w->id, -1.0, &w->st);
  if (w->rc) snprintf(w->err, sizeof w->err, "ct_job_wait failed (%d): %s", w->rc, ct_last_error());
}

static void wait_complete(napi_env env, napi_status status, void *data) {
  WaitWork *w = (WaitWork *)data;
  if (status == napi_ok && w->rc == 0) {
    napi_value o = status_object(env, &w->st);
    napi_resolve_deferred(env, w->deferred, o);
  } else {
    napi_value msg, err;
    const char *m = w->rc ? w->err : "waitJob: async work failed";
    napi_create_string_utf8(env, m, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_async_work(env, w->work);
  free(w);
}

static napi_value wait_job(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  WaitWork *w = (WaitWork *)calloc(1, sizeof *w);
  if (!w) {
    napi_throw_error(env, NULL, "waitJob: out of memory");
    return NULL;
  }
  w->srv = srv;
  w->id = (uint64_t)num(env, a[1]);
  napi_value promise, name;
  if (napi_create_promise(env, &w->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "caustic_amd.waitJob", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, wait_execute, wait_complete, w, &w->work) != napi_ok ||
      napi_queue_async_work(env, w->work) != napi_ok) {
    free(w);
    napi_throw_error(env, NULL, "waitJob: could not queue the wait");
    return NULL;
  }
  return promise;
}

static napi_value read_job(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  size_t n = (size_t)num(env, a[2]) * (size_t)num(env, a[3]) * 3;
  void *h;
  napi_value res = new_typed(env, napi_float32_array, n, 4, &h);
  if (!res) return NULL;
  int rc = ct_job_read(srv, (uint64_t)num(env, a[1]), (float *)h, 0);
  if (rc) return throw_ct(env, rc, "ct_job_read");
  return res;
}

static napi_value release_job(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  int rc = ct_job_release(srv, (uint64_t)num(env, a[1]));
  if (rc) return throw_ct(env, rc, "ct_job_release");
  return NULL;
}

static napi_value server_stats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  ct_server *srv = get_server(env, a[0]);
  if (!srv) return NULL;
  ct_server_stats st;
  int rc = ct_server_stats_get(srv, &st);
  if (rc) return throw_ct(env, rc, "ct_server_stats_get");
  napi_value o, v;
  CALL(env, napi_create_object(env, &o));
#define SET(name, val)                                                   \
  CALL(env, napi_create_double(env, (double)(val), &v));                 \
  CALL(env, napi_set_named_property(env, o, name, v));
  SET("submitted", st.submitted)
  SET("completed", st.completed)
  SET("failed", st.failed)
  SET("queued", st.queued)
  SET("active", st.active)
  SET("rounds", st.rounds)
  SET("paths", st.paths)
  SET("extensionRays", st.extensionRays)
  SET("shadowRays", st.shadowRays)
  SET("pixelSamples", st.pixelSamples)
  SET("hbmInUse", st.hbmInUse)
  SET("busySec", st.busySec)
  SET("nowSec", st.nowSec)
#undef SET
  return o;
}

static napi_value tonemap(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (!ensure_init(env)) return NULL;
  void *acc;
  size_t na;
  if (!typed(env, a[0], napi_float32_array, &acc, &na)) return NULL;
  uint32_t w = (uint32_t)num(env, a[1]), h = (uint32_t)num(env, a[2]);
  size_t n = (size_t)w * h;
  if (na < n * 3) {
    napi_throw_range_error(env, NULL, "tonemap: accum needs width*height*3 values");
    return NULL;
  }
  int rc = 0;
  void *d_acc = to_device(acc, n * 12, &rc);
  void *d_out = NULL;
  if (!rc) rc = ct_device_alloc(&d_out, n * 4);
  if (!rc) rc = ct_tonemap((const float *)d_acc, w, h, num(env, a[3]), num(env, a[4]),
                           (int32_t)num(env, a[5]), (uint8_t *)d_out, NULL);
  napi_value res = NULL;
  if (!rc) {
    void *hbuf;
    napi_value ab;
    if (napi_create_arraybuffer(env, n * 4, &hbuf, &ab) == napi_ok &&
        napi_create_typedarray(env, napi_uint8_clamped_array, n * 4, ab, 0, &res) == napi_ok)
      rc = ct_copy_d2h(hbuf, d_out, n * 4);
  }
  ct_device_free(d_acc);
  ct_device_free(d_out);
  if (rc) return throw_ct(env, rc, "ct_tonemap");
  return res;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"createScene", NULL, create_scene, NULL, NULL, NULL, napi_default, NULL},
      {"buildBvh", NULL, build_bvh, NULL, NULL, NULL, napi_default, NULL},
      {"closestHit", NULL, closest_hit, NULL, NULL, NULL, napi_default, NULL},
      {"occluded", NULL, occluded, NULL, NULL, NULL, napi_default, NULL},
      {"accumulatePass", NULL, accumulate_pass, NULL, NULL, NULL, napi_default, NULL},
      {"renderImage", NULL, render_image, NULL, NULL, NULL, napi_default, NULL},
      {"radiance", NULL, radiance, NULL, NULL, NULL, napi_default, NULL},
      {"setPrecision", NULL, set_precision, NULL, NULL, NULL, napi_default, NULL},
      {"setPathKernel", NULL, set_path_kernel, NULL, NULL, NULL, napi_default, NULL},
      {"createServer", NULL, create_server, NULL, NULL, NULL, napi_default, NULL},
      {"submitJob", NULL, submit_job, NULL, NULL, NULL, napi_default, NULL},
      {"pollJob", NULL, poll_job, NULL, NULL, NULL, napi_default, NULL},
      {"waitJob", NULL, wait_job, NULL, NULL, NULL, napi_default, NULL},
      {"readJob", NULL, read_job, NULL, NULL, NULL, napi_default, NULL},
      {"releaseJob", NULL, release_job, NULL, NULL, NULL, napi_default, NULL},
      {"serverStats", NULL, server_stats, NULL, NULL, NULL, napi_default, NULL},
      {"tonemap", NULL, tonemap, NULL, NULL, NULL, napi_default, NULL},
  };
  CALL(env, napi_define_properties(env, exports, sizeof props / sizeof props[0], props));
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
