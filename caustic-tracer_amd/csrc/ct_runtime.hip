// Runtime part of the C ABI: errors, device memory, device-resident scenes.
#include <string.h>

#include <algorithm>
#include <cmath>

#include <string>
#include <thread>
#include <vector>

#include "ct_internal.h"

namespace ct {

static thread_local std::string g_error;

int set_error(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return CT_OK;
  int code = (e == hipErrorOutOfMemory) ? CT_ENOMEM : CT_EDEVICE;
  return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// f(i) for i in [0, n), in contiguous chunks on up to 16 host threads (the
// upload's per-primitive and per-node passes, each i independent of the rest).
template <typename F>
static void parallel_for(size_t n, F f) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({hw, 16, (n + 4095) / 4096});
  if (nt <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (size_t i = n * t / nt, e = n * (t + 1) / nt; i < e; i++) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace ct

using namespace ct;

extern "C" {

const char* ct_last_error(void) { return g_error.c_str(); }

int ct_init(int32_t device) {
  int n = 0;
  CT_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(CT_EINVAL, "ct_init: device " + std::to_string(device) + " of " +
                                    std::to_string(n));
  CT_HIP(hipSetDevice(device));
  return CT_OK;
}

int ct_device_alloc(void** p, size_t bytes) {
  if (!p) return set_error(CT_EINVAL, "ct_device_alloc: null out pointer");
  CT_HIP(hipMalloc(p, bytes ? bytes : 1));
  return CT_OK;
}

int ct_device_free(void* p) {
  if (p) CT_HIP(hipFree(p));
  return CT_OK;
}

int ct_copy_h2d(void* dst, const void* src, size_t bytes) {
  CT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CT_OK;
}

int ct_copy_d2h(void* dst, const void* src, size_t bytes) {
  CT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CT_OK;
}

int ct_stream_sync(void* stream) {
  CT_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CT_OK;
}

// Host-side check that the kernels' loops end and stay in bounds on these
// arrays: every skip link moves forward and stays <= nodeCount, interior nodes
// have both children inside the array, leaves reference existing primitives,
// light and object indices are in range.
static int validate_scene(const float* nodes, uint32_t nodeCount, uint32_t primCount,
                          uint32_t matCount, const int32_t* lights, uint32_t lightCount,
                          const uint32_t* primObject, const uint32_t* leafWords) {
  for (uint32_t i = 0; i < nodeCount; i++) {
    const float* nd = nodes + (size_t)i * NODE_STRIDE;
    double skip = nd[3], leaf = nd[7];
    if (!(skip > i && skip <= nodeCount))
      return set_error(CT_EINVAL, "ct_scene_create: node " + std::to_string(i) +
                                      " has a skip link that does not move forward");
    if ((leaf >= 0) != (leafWords[i] != 0xffffffffu))
      return set_error(CT_EINVAL, "ct_scene_create: leaf words disagree with the nodes");
    if (leaf >= 0) {
      uint32_t w = leafWords[i];
      uint32_t count = w & 7, start = w >> 3;
      if (count && (uint64_t)start + count > primCount)
        return set_error(CT_EINVAL, "ct_scene_create: leaf " + std::to_string(i) +
                                        " references primitives beyond primCount");
    } else {
      if (i + 1 >= nodeCount)
        return set_error(CT_EINVAL, "ct_scene_create: interior node without children");
      double right = nodes[(size_t)(i + 1) * NODE_STRIDE + 3];
      if (!(right < nodeCount))
        return set_error(CT_EINVAL, "ct_scene_create: right child beyond nodeCount");
    }
  }
  for (uint32_t k = 0; k < lightCount; k++)
    if (lights[k] < 0 || (uint32_t)lights[k] >= primCount)
      return set_error(CT_EINVAL, "ct_scene_create: light index out of range");
  uint32_t mc = matCount ? matCount : 1;
  for (uint32_t p = 0; p < primCount; p++)
    if (primObject[p] >= mc)
      return set_error(CT_EINVAL, "ct_scene_create: primObject beyond matCount");
  return CT_OK;
}

// Exact leaf words: the caller's (layout extension for scenes of 2^21 or more
// primitives), or read from the float32 words of the reference layout.
// Returns whether every float32 word equals the exact one.
static bool leaf_words(const float* nodes, uint32_t nodeCount, const uint32_t* given,
                       std::vector<uint32_t>& out) {
  out.resize(nodeCount);
  bool exact = true;
  for (uint32_t i = 0; i < nodeCount; i++) {
    float f = nodes[(size_t)i * NODE_STRIDE + 7];
    if (given) {
      out[i] = given[i];
      if (f >= 0 && (double)f != (double)given[i]) exact = false;
    } else {
      out[i] = f >= 0 ? (uint32_t)f : 0xffffffffu;
    }
  }
  return exact;
}

// Box of primitive p as the reference computes the box of a leaf that holds
// p alone: the padded primitive bounds of pack.ts:152-154 (spheres) and
// pack.ts:177-187 (triangles, from the float32 vertices the tests use), then
// the outward nudge of bvh.ts:139-145, rounded outward to float32.
static void prim_box(const float* prims, uint32_t p, float lo[3], float hi[3]) {
  const float* r = prims + (size_t)p * PRIM_STRIDE;
  double mn[3], mx[3];
  if (r[3] == 0.0f) {
    const double rad = r[4], rp = rad + 1e-6 + std::fabs(rad) * 1e-6;
    for (int a = 0; a < 3; a++) {
      mn[a] = (double)r[a] - rp;
      mx[a] = (double)r[a] + rp;
    }
  } else {
    for (int a = 0; a < 3; a++) {
      mn[a] = std::min({(double)r[a], (double)r[4 + a], (double)r[8 + a]}) - 1e-6;
      mx[a] = std::max({(double)r[a], (double)r[4 + a], (double)r[8 + a]}) + 1e-6;
    }
  }
  for (int a = 0; a < 3; a++) {
    const double l = mn[a] - std::fabs(mn[a]) * 1.2e-7 - 1e-9;
    const double h = mx[a] + std::fabs(mx[a]) * 1.2e-7 + 1e-9;
    float fl = (float)l, fh = (float)h;
    if ((double)fl > l) fl = std::nextafter(fl, -INFINITY);
    if ((double)fh < h) fh = std::nextafter(fh, INFINITY);
    lo[a] = fl;
    hi[a] = fh;
  }
}

// Union of the boxes of primitives p .. p + len - 1.
static void run_box(const float* pbox, uint32_t p, uint32_t len, float lo[3], float hi[3]) {
  for (int a = 0; a < 3; a++) {
    lo[a] = pbox[(size_t)p * 6 + a];
    hi[a] = pbox[(size_t)p * 6 + 3 + a];
  }
  for (uint32_t j = 1; j < len; j++)
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], pbox[(size_t)(p + j) * 6 + a]);
      hi[a] = std::max(hi[a], pbox[(size_t)(p + j) * 6 + 3 + a]);
    }
}

static double half_area(const float lo[3], const float hi[3]) {
  double x = (double)hi[0] - lo[0], y = (double)hi[1] - lo[1], z = (double)hi[2] - lo[2];
  return x * y + y * z + z * x;
}

// 4-wide nodes from the binary tree by the surface-area cost model, collapsed
// optimally by dynamic programming over the binary tree (the collapse of
// Ylitie, Karras and Laine, "Efficient incoherent ray traversal on GPUs through
// compressed wide BVHs", HPG 2017, for four children), with each reference
// leaf's primitives also offered in runs of consecutive ones, a box each:
//   D(n, k)  cheapest cover of the subtree of n by at most k roots, a root
//            being a wide node, a reference leaf or a run of its primitives
//   wide(n)  = A(n) + min over k of D(left, k) + D(right, 4 - k)
//   R(n, r)  cheapest split of leaf n into r runs: sum over the runs of
//            A(run) * C_PRIM * length, A(run) the union of its primitives' boxes
//   leaf(n)  = min(A(n) * C_PRIM * count(n),          the leaf tested whole
//                  A(n) + min over r <= 4 of R(n, r))  a wide node of its runs
//   D(leaf, k) may also be R(n, r) for 2 <= r <= k:    its runs as roots
// (the runs' terms weighted by SPLIT_MARGIN) with A the surface area of a
// box, in units of one 4-wide node step.  A primitive's own box is the box
// the reference gives a leaf of that primitive alone (prim_box), so it holds
// every hit on it, as the reference leaf box does, and a run's box holds its
// primitives' boxes.  Every box is such a box or a reference node's box, so
// the walks over these nodes visit a superset of the boxes on the path to any
// primitive and return the reference's closest hit.  A wide node holds at most
// four runs, whatever the leaf's count (leaf words allow seven).  Wide nodes are numbered breadth first
// until the first TOP_NODES slots are taken (the block traversal launches copy
// into LDS), then depth first.
//
// Measured on the 1M-triangle soup at 1920x1080x4 (binary64 / FP32 Mrays/s):
// reference leaves only 1249 / 1351 (the collapse itself takes 1.7% fewer
// node visits per ray than opening the largest child first; merging subtrees
// of up to 7 primitives into one leaf cost 4%); every leaf of two or more
// primitives a wide node of its primitives 1355 / 1439 (27.5 -> 4.4 triangle
// tests and 167 -> 223 node slots per ray); the cost model above 1383 / 1488
// with C_PRIM 0.5 (0.75, 1.0, 1.5: the same within 0.5%).  The caustic
// scene's tessellated spheres, whose neighbouring triangles' boxes overlap,
// lost 1-2% with it (FP32 3051 vs 3080, binary64 2682 vs 2748); counting the
// primitive boxes at SPLIT_MARGIN = 1.5 times their cost keeps the caustic
// scene where it was (3089 / 2743) and the soup at 1370 / 1479 (2 or 3:
// 1360 / 1474, 1344 / 1457).  Runs of several primitives, beside single ones:
// soup 1381 / 1499 against 1368 / 1480, caustic scene 2742 / 3091 against
// 2738 / 3086 (314k instead of 340k wide nodes on the soup).
#ifndef CT_C_PRIM
#define CT_C_PRIM 0.5
#endif
constexpr double C_PRIM = CT_C_PRIM;  // one binary64 primitive test, in node steps
#ifndef CT_SPLIT_MARGIN
#define CT_SPLIT_MARGIN 1.5
#endif
constexpr double SPLIT_MARGIN = CT_SPLIT_MARGIN;  // weight of the primitive boxes' cost
#ifndef CT_LEAF_RUNS
#define CT_LEAF_RUNS 1
#endif
constexpr bool LEAF_RUNS = CT_LEAF_RUNS;  // runs of several primitives, not only single ones

static void build_wide(const float* nodes, uint32_t nodeCount, const uint32_t* words,
                       const float* prims, uint32_t primCount, std::vector<float>& out) {
  out.clear();
  const uint32_t N = nodeCount;
  std::vector<float> pbox((size_t)primCount * 6);  // prim_box of each primitive: lo, hi
  parallel_for(primCount, [&](size_t p) {
    prim_box(prims, (uint32_t)p, &pbox[p * 6], &pbox[p * 6 + 3]);
  });
  auto isLeaf = [&](uint32_t i) { return words[i] != 0xffffffffu; };
  auto rightOf = [&](uint32_t i) { return (uint32_t)nodes[(size_t)(i + 1) * NODE_STRIDE + 3]; };
  auto area = [&](uint32_t i) {
    const float* b = nodes + (size_t)i * NODE_STRIDE;
    double x = (double)b[4] - b[0], y = (double)b[5] - b[1], z = (double)b[6] - b[2];
    return x * y + y * z + z * x;
  };
  // Children have larger indices than their parent (skip links move forward),
  // so one pass from the last node to the first sees children before parents.
  std::vector<float> D((size_t)N * 4);    // D[4n + k-1] = D(n, k)
  std::vector<uint8_t> how((size_t)N * 4);  // interior, k = 1: 1 (wide); k > 1: roots given
                                            // to the left child (0: as for k - 1); leaf,
                                            // k = 1: 0 tested whole, 2 wide node of runs;
                                            // k > 1: 3 runs as roots (0: as for k - 1)
  std::vector<uint8_t> cuts((size_t)N * 4);  // leaf runs: bit j = a run ends at primitive j
  std::vector<uint8_t> split(N);          // wide(n): slots given to the left child
  const float INF = __builtin_inff();
  for (int64_t ii = (int64_t)N - 1; ii >= 0; ii--) {
    const uint32_t i = (uint32_t)ii;
    float* d = &D[(size_t)i * 4];
    uint8_t* h = &how[(size_t)i * 4];
    const double a = area(i);
    if (isLeaf(i)) {
      const uint32_t c = words[i] & 7, start = words[i] >> 3;
      d[0] = d[1] = d[2] = d[3] = (float)(a * C_PRIM * c);
      h[0] = h[1] = h[2] = h[3] = 0;
      if (c >= 2) {
        // the leaf's primitives in 2-4 runs of consecutive ones (a box each,
        // the union of its primitives' boxes; leaf word start*8 + length):
        // the cheapest partition for each number of runs
        double best[5] = {INF, INF, INF, INF, INF};
        uint8_t bestCuts[5] = {0, 0, 0, 0, 0};
        const uint32_t last = 1u << (c - 1);  // bit c-1: the last run ends at the last primitive
        for (uint32_t m = last; m < 2 * last; m++) {
          const int runs = __builtin_popcount(m);
          if (runs < 2 || runs > 4 || (!LEAF_RUNS && runs != (int)c)) continue;
          double cost = 0;
          for (uint32_t j = 0, rs = 0; j < c; j++) {
            if (!((m >> j) & 1)) continue;
            float lo[3], hi[3];
            run_box(pbox.data(), start + rs, j - rs + 1, lo, hi);
            cost += half_area(lo, hi) * C_PRIM * SPLIT_MARGIN * (j - rs + 1);
            rs = j + 1;
          }
          if (cost < best[runs]) {
            best[runs] = cost;
            bestCuts[runs] = (uint8_t)m;
          }
        }
        for (int r = 2; r <= 4; r++)  // a wide node of the runs
          if ((float)(a + best[r]) < d[0]) {
            d[0] = (float)(a + best[r]);
            h[0] = 2;
            cuts[(size_t)i * 4] = bestCuts[r];
          }
        for (int k = 2; k <= 4; k++) {  // the runs as roots
          d[k - 1] = d[k - 2];
          if ((float)best[k] < d[k - 1]) {
            d[k - 1] = (float)best[k];
            h[k - 1] = 3;
            cuts[(size_t)i * 4 + k - 1] = bestCuts[k];
          }
        }
      }
      continue;
    }
    const uint32_t l = i + 1, r = rightOf(i);
    const float* dl = &D[(size_t)l * 4];
    const float* dr = &D[(size_t)r * 4];
    float bestF = INF;
    uint8_t bestK = 1;
    for (int k = 1; k <= 3; k++) {
      float c = dl[k - 1] + dr[4 - k - 1];
      if (c < bestF) {
        bestF = c;
        bestK = (uint8_t)k;
      }
    }
    split[i] = bestK;
    d[0] = (float)a + bestF;  // n as a wide node
    h[0] = 1;
    for (int k = 2; k <= 4; k++) {
      d[k - 1] = d[k - 2];
      h[k - 1] = 0;
      for (int kl = 1; kl < k; kl++) {
        float c = dl[kl - 1] + dr[k - kl - 1];
        if (c < d[k - 1]) {
          d[k - 1] = c;
          h[k - 1] = (uint8_t)kl;
        }
      }
    }
  }
  // the roots covering subtree m with at most k of them: a reference node
  // (count 0), or a run of `count` primitives from `start` of a leaf whose runs
  // are roots
  struct Root {
    uint32_t n, start, count;
  };
  std::vector<Root> roots;
  std::vector<std::pair<uint32_t, int>> pend;
  auto expand = [&](uint32_t m, int k) {
    pend.clear();
    pend.push_back({m, k});
    while (!pend.empty()) {
      auto [n, kk] = pend.back();
      pend.pop_back();
      while (kk > 1 && how[(size_t)n * 4 + kk - 1] == 0) kk--;
      if (kk == 1) {
        roots.push_back({n, 0, 0});
      } else if (isLeaf(n)) {  // its runs as roots
        const uint32_t c = words[n] & 7, start = words[n] >> 3, m = cuts[(size_t)n * 4 + kk - 1];
        for (uint32_t j = 0, rs = 0; j < c; j++)
          if ((m >> j) & 1) {
            roots.push_back({n, start + rs, j - rs + 1});
            rs = j + 1;
          }
      } else {
        int kl = how[(size_t)n * 4 + kk - 1];
        pend.push_back({rightOf(n), kk - kl});
        pend.push_back({n + 1, kl});  // left first
      }
    }
  };
  auto alloc = [&]() {
    uint32_t slot = (uint32_t)(out.size() / 32);
    out.resize(out.size() + 32);
    float* rr = &out[(size_t)slot * 32];
    for (int k = 0; k < 4; k++) {
      // an empty slot: a box at +inf on every axis, which no ray enters
      rr[k] = rr[4 + k] = rr[8 + k] = __builtin_inff();
      rr[12 + k] = rr[16 + k] = rr[20 + k] = __builtin_inff();
      int32_t empty = -1;
      memcpy(&rr[24 + k], &empty, 4);
    }
    return slot;
  };
  auto put = [&](uint32_t slot, int k, uint32_t n2, int32_t c) {
    float* rr = &out[(size_t)slot * 32];
    const float* b = nodes + (size_t)n2 * NODE_STRIDE;
    rr[k] = b[0]; rr[4 + k] = b[1]; rr[8 + k] = b[2];
    rr[12 + k] = b[4]; rr[16 + k] = b[5]; rr[20 + k] = b[6];
    memcpy(&rr[24 + k], &c, 4);
  };
  auto putRun = [&](uint32_t slot, int k, uint32_t p, uint32_t len) {  // leaf word p*8 + len
    float lo[3], hi[3];
    run_box(pbox.data(), p, len, lo, hi);
    float* rr = &out[(size_t)slot * 32];
    for (int a = 0; a < 3; a++) {
      rr[4 * a + k] = lo[a];
      rr[12 + 4 * a + k] = hi[a];
    }
    int32_t c = -2 - (int32_t)((p << 3) | len);
    memcpy(&rr[24 + k], &c, 4);
  };
  uint32_t root = alloc();
  if (isLeaf(0)) {
    put(root, 0, 0, -2 - (int32_t)words[0]);
    return;
  }
  struct Item { uint32_t n2; uint32_t slot; };  // binary node expanded into wide node `slot`
  std::vector<Item> work;
  work.push_back({0, root});
  size_t head = 0;  // work[head..] pending; FIFO while the front slot is below TOP_NODES
  while (head < work.size()) {
    Item it;
    if (work[head].slot < (uint32_t)TOP_NODES) {
      it = work[head++];
    } else {
      it = work.back();
      work.pop_back();
    }
    roots.clear();
    expand(it.n2 + 1, split[it.n2]);
    expand(rightOf(it.n2), 4 - split[it.n2]);
    for (size_t k = 0; k < roots.size(); k++) {
      const uint32_t g = roots[k].n;
      if (roots[k].count) {
        putRun(it.slot, (int)k, roots[k].start, roots[k].count);
      } else if (isLeaf(g) && how[(size_t)g * 4] == 2) {  // a wide node of its runs
        const uint32_t ws = alloc();
        put(it.slot, (int)k, g, (int32_t)ws);
        const uint32_t c = words[g] & 7, start = words[g] >> 3, m = cuts[(size_t)g * 4];
        int slot = 0;
        for (uint32_t j = 0, rs = 0; j < c; j++)
          if ((m >> j) & 1) {
            putRun(ws, slot++, start + rs, j - rs + 1);
            rs = j + 1;
          }
      } else if (isLeaf(g)) {
        put(it.slot, (int)k, g, -2 - (int32_t)words[g]);
      } else {
        uint32_t ws = alloc();
        put(it.slot, (int)k, g, (int32_t)ws);
        work.push_back({g, ws});
      }
    }
  }
}

// Exact difference of two floats in binary64, when it is representable.
static bool exact_diff(float a, float b, double& d) {
  double x = a, y = -(double)b;
  d = x + y;
  double bb = d - x;
  double err = (x - (d - bb)) + (y - bb);
  return err == 0.0;
}

// Smallest e in [-126, 127] with 255 * 2^e >= ext (127 when there is none, -126
// for ext <= 0 or NaN), as a search upwards from -126 finds it: frexp gives the
// candidate and the exact comparisons (255 * 2^e is exact in binary64) settle
// it in a step or two instead of up to 253.
static int scale_exponent(double ext) {
  if (!(ext > 0)) return -126;
  if (ext == INFINITY) return 127;
  int k;
  std::frexp(ext / 255.0, &k);
  int e = std::max(-126, std::min(127, k));
  while (e > -126 && 255.0 * std::ldexp(1.0, e - 1) >= ext) e--;
  while (e < 127 && 255.0 * std::ldexp(1.0, e) < ext) e++;
  return e;
}

// 8-bit parent-relative quantization of the 4-wide nodes (precision form).
// Per node and axis the origin is the children's minimum and the scale the
// smallest power of two with 255*scale covering the children's extent; child
// planes are rounded outward (floor below, ceil above) and then checked in
// exact arithmetic, so every decoded box origin + q*scale contains the float32
// box it replaces and a walk over them enters every subtree the float32 boxes
// would (the kernels bound their own rounding separately).
static void quantize_wide(const std::vector<float>& wide, std::vector<float>& out) {
  size_t nw = wide.size() / 32;
  out.assign(nw * 16, 0.0f);
  parallel_for(nw, [&](size_t n) {
    const float* w = &wide[n * 32];
    float* q = &out[n * 16];
    int32_t codes[4];
    memcpy(codes, w + 24, 16);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool any = false;
    for (int k = 0; k < 4; k++) {
      if (codes[k] == -1) continue;
      any = true;
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], w[4 * a + k]);
        hi[a] = std::max(hi[a], w[12 + 4 * a + k]);
      }
    }
    float origin[3] = {0, 0, 0}, scale[3] = {1, 1, 1};
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    if (any) {
      for (int a = 0; a < 3; a++) {
        origin[a] = lo[a];
        double ext;
        bool ok = exact_diff(hi[a], lo[a], ext);
        int e = scale_exponent(ext);
        if (!ok && e < 127) e++;  // inexact extent: one more doubling of margin
        scale[a] = std::ldexp(1.0f, e);
      }
      for (int k = 0; k < 4; k++) {
        if (codes[k] == -1) continue;
        for (int a = 0; a < 3; a++) {
          const double s = scale[a];
          double dl, dh;
          uint32_t ql = 0, qh = 255;
          if (exact_diff(w[4 * a + k], origin[a], dl)) {
            double f = std::floor(dl / s);
            ql = (uint32_t)std::min(255.0, std::max(0.0, f));
            while (ql > 0 && ql * s > dl) ql--;  // decoded plane <= box plane
          }
          if (exact_diff(w[12 + 4 * a + k], origin[a], dh)) {
            double c = std::ceil(dh / s);
            qh = (uint32_t)std::min(255.0, std::max(0.0, c));
            while (qh < 255 && qh * s < dh) qh++;  // decoded plane >= box plane
          }
          qlo[a] |= ql << (8 * k);
          qhi[a] |= qh << (8 * k);
        }
      }
    }
    q[0] = origin[0]; q[1] = origin[1]; q[2] = origin[2]; q[3] = scale[0];
    q[4] = scale[1]; q[5] = scale[2];
    memcpy(&q[6], &qlo[0], 4);
    memcpy(&q[7], &qlo[1], 4);
    memcpy(&q[8], &qlo[2], 4);
    memcpy(&q[9], &qhi[0], 4);
    memcpy(&q[10], &qhi[1], 4);
    memcpy(&q[11], &qhi[2], 4);
    memcpy(&q[12], codes, 16);
  });
}

static int scene_create_impl(const float* prims, uint32_t primCount, const float* nodes,
                             uint32_t nodeCount, const float* mats, uint32_t matCount,
                             const int32_t* lights, uint32_t lightCount,
                             const uint32_t* primObject, const ct_sky* sky,
                             const uint32_t* leafWordsIn, hipMemcpyKind kind,
                             hipStream_t stream, ct_scene** out) {
  if (!out) return set_error(CT_EINVAL, "ct_scene_create: null out");
  *out = nullptr;
  if (nodeCount == 0 || !nodes)
    return set_error(CT_EINVAL, "ct_scene_create: a packed scene has at least one node");
  if ((primCount && !prims) || (matCount && !mats) || (lightCount && !lights) ||
      (primCount && !primObject))
    return set_error(CT_EINVAL, "ct_scene_create: null array with non-zero count");
  if (nodeCount > (1u << 24))
    return set_error(CT_ELIMIT,
                     "ct_scene_create: skip links are float32 and exact only below 2^24 nodes");
  if (primCount >= (1u << 28))
    return set_error(CT_ELIMIT, "ct_scene_create: more than 2^28 primitives");
  std::vector<float> wide, qwide;
  std::vector<uint32_t> words;
  bool binaryExact = true;
  if (kind == hipMemcpyHostToDevice) {
    binaryExact = leaf_words(nodes, nodeCount, leafWordsIn, words);
    int v = validate_scene(nodes, nodeCount, primCount, matCount, lights, lightCount, primObject,
                           words.data());
    if (v) return v;
    build_wide(nodes, nodeCount, words.data(), prims, primCount, wide);
  } else {
    std::vector<float> hn((size_t)nodeCount * NODE_STRIDE);
    std::vector<int32_t> hl(lightCount);
    std::vector<uint32_t> ho(primCount);
    CT_HIP(hipMemcpyAsync(hn.data(), nodes, hn.size() * 4, hipMemcpyDeviceToHost, stream));
    if (lightCount)
      CT_HIP(hipMemcpyAsync(hl.data(), lights, hl.size() * 4, hipMemcpyDeviceToHost, stream));
    if (primCount)
      CT_HIP(hipMemcpyAsync(ho.data(), primObject, ho.size() * 4, hipMemcpyDeviceToHost, stream));
    std::vector<float> hp((size_t)primCount * PRIM_STRIDE);
    if (primCount)
      CT_HIP(hipMemcpyAsync(hp.data(), prims, hp.size() * 4, hipMemcpyDeviceToHost, stream));
    CT_HIP(hipStreamSynchronize(stream));
    binaryExact = leaf_words(hn.data(), nodeCount, leafWordsIn, words);
    int v = validate_scene(hn.data(), nodeCount, primCount, matCount, hl.data(), lightCount,
                           ho.data(), words.data());
    if (v) return v;
    build_wide(hn.data(), nodeCount, words.data(), hp.data(), primCount, wide);
  }
  quantize_wide(wide, qwide);
  // The array lengths PackedScene carries: max(1, n) records each (pack.ts:136-138, 111).
  size_t pc = primCount ? primCount : 1, mc = matCount ? matCount : 1;
  size_t szPrims = align_up(pc * PRIM_STRIDE * sizeof(float));
  size_t szNodes = align_up((size_t)nodeCount * NODE_STRIDE * sizeof(float));
  size_t szMats = align_up(mc * MAT_STRIDE * sizeof(float));
  size_t szLights = align_up((lightCount ? lightCount : 1) * sizeof(int32_t));
  size_t szObj = align_up(pc * sizeof(uint32_t));
  size_t szQwide = align_up(qwide.size() * sizeof(float));
  size_t szWords = align_up(words.size() * sizeof(uint32_t));
  size_t total = szPrims + szNodes + szMats + szLights + szObj + szQwide + szWords;
  ct_scene* s = new ct_scene();
  hipError_t e = hipMalloc(&s->buf, total);
  if (e != hipSuccess) {
    delete s;
    return check_hip(e, "ct_scene_create: hipMalloc");
  }
  s->bytes = total;
  if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
  char* base = (char*)s->buf;
  char* pPrims = base;
  char* pNodes = pPrims + szPrims;
  char* pMats = pNodes + szNodes;
  char* pLights = pMats + szMats;
  char* pObj = pLights + szLights;
  char* pQwide = pObj + szObj;
  char* pWords = pQwide + szQwide;
  e = hipMemsetAsync(base, 0, total, stream);
  if (e == hipSuccess && primCount)
    e = hipMemcpyAsync(pPrims, prims, (size_t)primCount * PRIM_STRIDE * sizeof(float), kind, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pNodes, nodes, (size_t)nodeCount * NODE_STRIDE * sizeof(float), kind, stream);
  if (e == hipSuccess && matCount)
    e = hipMemcpyAsync(pMats, mats, (size_t)matCount * MAT_STRIDE * sizeof(float), kind, stream);
  if (e == hipSuccess && lightCount)
    e = hipMemcpyAsync(pLights, lights, (size_t)lightCount * sizeof(int32_t), kind, stream);
  if (e == hipSuccess && primCount)
    e = hipMemcpyAsync(pObj, primObject, (size_t)primCount * sizeof(uint32_t), kind, stream);
  if (e == hipSuccess && !qwide.empty())
    e = hipMemcpyAsync(pQwide, qwide.data(), qwide.size() * sizeof(float), hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pWords, words.data(), words.size() * sizeof(uint32_t),
                       hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    (void)hipFree(s->buf);
    delete s;
    return check_hip(e, "ct_scene_create: upload");
  }
  DevScene& d = s->dev;
  d.prims = (const float4*)pPrims;
  d.nodes = (const float4*)pNodes;
  d.mats = (const float4*)pMats;
  d.lights = (const int32_t*)pLights;
  d.primObject = (const uint32_t*)pObj;
  d.wideCount = (int)(wide.size() / 32);
  d.qwide = (const float4*)pQwide;
  d.leafWords = (const uint32_t*)pWords;
  d.binaryExact = binaryExact ? 1 : 0;
  d.primCount = (int)primCount;
  d.nodeCount = (int)nodeCount;
  d.lightCount = (int)lightCount;
  d.skyKind = (sky && sky->kind == 1) ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    d.zenith[k] = d.skyKind ? sky->zenith[k] : 0;
    d.horizon[k] = d.skyKind ? sky->horizon[k] : 0;
  }
  d.skyStrength = d.skyKind ? sky->strength : 0;
  *out = s;
  return CT_OK;
}

int ct_scene_create(const float* prims, uint32_t primCount, const float* nodes,
                    uint32_t nodeCount, const float* mats, uint32_t matCount,
                    const int32_t* lights, uint32_t lightCount, const uint32_t* primObject,
                    const ct_sky* sky, ct_scene** out) {
  return scene_create_impl(prims, primCount, nodes, nodeCount, mats, matCount, lights,
                           lightCount, primObject, sky, nullptr, hipMemcpyHostToDevice, nullptr,
                           out);
}

int ct_scene_create_ex(const float* prims, uint32_t primCount, const float* nodes,
                       uint32_t nodeCount, const float* mats, uint32_t matCount,
                       const int32_t* lights, uint32_t lightCount, const uint32_t* primObject,
                       const ct_sky* sky, const uint32_t* leafWords, ct_scene** out) {
  if (!leafWords) return set_error(CT_EINVAL, "ct_scene_create_ex: null leafWords");
  return scene_create_impl(prims, primCount, nodes, nodeCount, mats, matCount, lights,
                           lightCount, primObject, sky, leafWords, hipMemcpyHostToDevice,
                           nullptr, out);
}

int ct_scene_create_device(const float* prims, uint32_t primCount, const float* nodes,
                           uint32_t nodeCount, const float* mats, uint32_t matCount,
                           const int32_t* lights, uint32_t lightCount,
                           const uint32_t* primObject, const ct_sky* sky, void* stream,
                           ct_scene** out) {
  return scene_create_impl(prims, primCount, nodes, nodeCount, mats, matCount, lights,
                           lightCount, primObject, sky, nullptr, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream, out);
}

int ct_scene_destroy(ct_scene* s) {
  if (!s) return CT_OK;
  hipError_t e = hipFree(s->buf);
  delete s;
  return check_hip(e, "ct_scene_destroy");
}

}  // extern "C"
