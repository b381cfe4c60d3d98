// GPU binned-SAH BVH builder, bit-identical to the reference's recursive
// builder (bvh.ts:72-301).
//
// The reference builds depth first: buildNode(start,end) reduces the node's
// FP64 bounds, makes a leaf at <= 4 prims, else bins centroids into 12 buckets
// per axis, sweeps the SAH cost, partitions order[start,end) in place with a
// two-pointer loop (or stable-sorts it when every centroid axis is
// degenerate) and recurses left then right, numbering nodes in DFS pre-order.
//
// Here the same tree is built breadth first, one level per round of kernels,
// over node records that are kept in position order:
//   bounds     segmented min/max of prim and centroid bounds (order-independent,
//              exact through sortable-integer atomics), wave/LDS aggregated
//   setup      leaf / SAH / median decision per node, bucket scale per axis
//   bin        LDS histograms per (node, axis, bucket) for the first nodes a
//              block touches, global atomics for the rest
//   sweep      the reference's suffix/prefix sweeps per node (one lane per node)
//   partition  the permutation of the two-pointer loop from prefix counts:
//              L-goods stay, the k-th R-good fills the k-th L-bad's slot, the
//              k-th L-bad lands at end-2-b_{k-1}, R-bads move one slot inward,
//              up to the meeting point K = max{k : a_k + b_k + 2 <= n}
//   median     stable sort by (centroid, position): bitonic in LDS up to 2048
//              primitives, a tiled merge sort over several workgroups above
//   children   child records appended in parent order (exclusive scan)
// After the last level, subtree sizes (bottom-up) give DFS indices and skip
// links (top-down), and each record is written as one 8-float node.
#include <string.h>

#include <stdio.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "ct_internal.h"

namespace ct {

constexpr int BB = 256;          // block size of position kernels
constexpr int CHUNK = BB * 8;    // positions per block in the aggregated kernels
constexpr int LSLOTS = 4;        // nodes per block aggregated in LDS
constexpr int MAX_BINS = 32;
constexpr int MEDIAN_MAX = 2048;  // largest all-degenerate node the LDS sort takes;
                                  // larger ones go through k_msort_tiles/k_msort_merge
// Nodes of up to this many positions are reduced by one wave each
// (k_bounds_small, k_bin_small); the chunked kernels take the larger ones,
// of which a chunk then touches at most three, all within its LDS slots.
constexpr uint32_t SMALL_NODE = 2048;
// Bounds of nodes of up to this many positions are reduced by one lane each.
constexpr uint32_t TINY_NODE = 16;
constexpr int WAVES = BB / 64;

enum : int32_t { KIND_LEAF = 0, KIND_SAH = 1, KIND_MEDIAN = 2 };

// Sortable 64-bit image of a double: unsigned order == numeric order, -0 < +0.
__device__ __forceinline__ unsigned long long dkey(double x) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(unsigned long long k) {
  unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}
#define KEY_POS_INF 0xfff0000000000000ull  // dkey(+inf)
#define KEY_NEG_INF 0x000fffffffffffffull  // dkey(-inf)

struct Build {
  // inputs
  const double* bmin;
  const double* bmax;
  const double* bcen;
  // the inputs gathered per primitive, 9 doubles: bmin.xyz bmax.xyz bcen.xyz
  // (one primitive's bounds in one or two cache lines for the per-level gathers)
  double* box;
  uint32_t n;
  int bins;
  // position arrays
  uint32_t* order;      // current
  uint32_t* orderNext;  // next level
  int32_t* nodeOf;      // record of each position, -1 once in a finished leaf
  uint32_t* good;       // partition flag per position (0/1)
  uint32_t* goodScan;   // exclusive scan of good, n+1 entries
  uint32_t* A;          // per node: positions of bads from the left (relative)
  uint32_t* B;          // per node: R-indices of goods from the right
  // records (capacity 2n)
  uint32_t* recStart;
  uint32_t* recEnd;
  int32_t* recLeft;
  int32_t* recRight;
  double* recBox;       // 6 per record: min.xyz, max.xyz (un-nudged)
  uint32_t* recSize;
  uint32_t* recIdx;
  // per active node of the current level (local index j)
  unsigned long long* keys;  // 12 per node: bmin3 bmax3 cmin3 cmax3
  int32_t* kind;
  int32_t* axis;             // SAH best axis / median axis
  int32_t* split;            // SAH best split bucket
  double* origin;            // SAH bestOrigin (cmin of axis)
  double* scale;             // SAH bestScale (bins / extent)
  double* cmin3;             // 3 per node
  double* tob3;              // 3 per node: bins/extent or 0 when the axis is skipped
  uint32_t* flag;            // scan input (several uses)
  uint32_t* flagScan;        // scan output (m+1)
  uint32_t* binSlot;         // scan of SAH flags
  uint32_t* mid;             // split position per node
  uint32_t* meetX;           // L-read count of the partition per node
  // bins per SAH node: 3 * bins * (count + 6 keys)
  uint32_t* binCount;
  unsigned long long* binKey;
  uint32_t* levelInfo;       // [0] splits [1] median nodes [2] inconsistent partitions
                             // [3] median nodes above the LDS sort
  uint32_t* scanTmp;         // block sums for scans
};

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (out has n+1 entries; out[n] = total)

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_BLOCK)
k_scan_tile(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
            uint32_t* __restrict__ tileSums) {
  __shared__ uint32_t s[SCAN_BLOCK];
  uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint32_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint32_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (threadIdx.x == SCAN_BLOCK - 1) {
    tileSums[blockIdx.x] = s[SCAN_BLOCK - 1];
    if (gridDim.x == 1) out[n] = s[SCAN_BLOCK - 1];  // one tile: its sum is the total
  }
}

// tileOffsets: the exclusive scan of the tile sums, gridDim.x + 1 entries (the
// last is the total, written to out[n])
__global__ void __launch_bounds__(SCAN_BLOCK)
k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ tileOffsets) {
  uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t add = tileOffsets[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tileOffsets[gridDim.x];
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint32_t i = base + k;
    if (i < n) out[i] += add;
  }
}

__global__ void k_store_u32(uint32_t* p, uint32_t v) { *p = v; }

// out[0..n] = exclusive scan of in[0..n-1]; tmp needs room for the tile sums of
// every recursion level (about n / 1000 entries).
static int scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                    hipStream_t st) {
  if (n == 0) {
    hipLaunchKernelGGL(k_store_u32, dim3(1), dim3(1), 0, st, out, 0u);
    CT_LAUNCH_CHECK("k_store_u32");
    return CT_OK;
  }
  uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(SCAN_BLOCK), 0, st, in, out, n, tmp);
  CT_LAUNCH_CHECK("k_scan_tile");
  if (tiles > 1) {
    uint32_t* offs = tmp + tiles;  // tiles+1 entries
    int rc = scan_u32(tmp, offs, tiles, offs + tiles + 1, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_BLOCK), 0, st, out, n, offs);
    CT_LAUNCH_CHECK("k_scan_add");
  }
  return CT_OK;
}

static size_t scan_tmp_entries(size_t n) {
  size_t total = 0;
  while (n > 1) {
    size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    total += 2 * tiles + 2;
    n = tiles;
  }
  return total + 4;
}

// ---------------------------------------------------------------------------
// level kernels

__global__ void k_init_root(Build b) {
  b.recStart[0] = 0;
  b.recEnd[0] = b.n;
  b.recLeft[0] = -1;
  b.recRight[0] = -1;
}

__global__ void k_init_positions(Build b) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    b.order[i] = i;
    b.nodeOf[i] = 0;
    double* x = b.box + 9 * (size_t)i;
    for (int k = 0; k < 3; k++) {
      x[k] = b.bmin[3 * (size_t)i + k];
      x[3 + k] = b.bmax[3 * (size_t)i + k];
      x[6 + k] = b.bcen[3 * (size_t)i + k];
    }
  }
}

__global__ void k_level_init(Build b, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    for (int k = 0; k < 12; k++)
      b.keys[12 * (size_t)j + k] = (k < 3 || (k >= 6 && k < 9)) ? KEY_POS_INF : KEY_NEG_INF;
  }
  // [0] splits, [1] median nodes, [3] median nodes above the LDS sort; [2]
  // (inconsistent partitions) accumulates over the levels
  if (blockIdx.x == 0 && (threadIdx.x < 2 || threadIdx.x == 3)) b.levelInfo[threadIdx.x] = 0;
}

__device__ __forceinline__ bool small_node(const Build& b, uint32_t rec) {
  return b.recEnd[rec] - b.recStart[rec] <= SMALL_NODE;
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

// First active local node index in this block's chunk that the chunked
// kernels reduce (not a small node), or m if none.
__device__ uint32_t block_first_node(Build& b, uint32_t c0, uint32_t lvStart, uint32_t m,
                                     uint32_t* sFirst) {
  if (threadIdx.x == 0) *sFirst = m;
  __syncthreads();
  uint32_t mine = m;
  for (uint32_t i = c0 + threadIdx.x; i < c0 + CHUNK && i < b.n; i += BB) {
    int32_t r = b.nodeOf[i];
    if (r >= 0 && !small_node(b, (uint32_t)r)) {
      mine = (uint32_t)r - lvStart;
      break;
    }
  }
  if (mine < m) atomicMin(sFirst, mine);
  __syncthreads();
  return *sFirst;
}

// Bounds of every active node: box of prim bounds and box of centroids.
__global__ void __launch_bounds__(BB) k_bounds(Build b, uint32_t lvStart, uint32_t m) {
  __shared__ unsigned long long sKeys[LSLOTS * 12];
  __shared__ uint32_t sFirst;
  uint32_t c0 = blockIdx.x * CHUNK;
  for (int k = threadIdx.x; k < LSLOTS * 12; k += BB) {
    int f = k % 12;
    sKeys[k] = (f < 3 || (f >= 6 && f < 9)) ? KEY_POS_INF : KEY_NEG_INF;
  }
  uint32_t j0 = block_first_node(b, c0, lvStart, m, &sFirst);
  for (uint32_t i = c0 + threadIdx.x; i < c0 + CHUNK && i < b.n; i += BB) {
    int32_t r = b.nodeOf[i];
    if (r < 0 || small_node(b, (uint32_t)r)) continue;
    uint32_t j = (uint32_t)r - lvStart;
    const double* x = b.box + 9 * (size_t)b.order[i];
    unsigned long long v[12];
    for (int k = 0; k < 3; k++) {
      v[k] = dkey(x[k]);
      v[3 + k] = dkey(x[3 + k]);
      v[6 + k] = dkey(x[6 + k]);
      v[9 + k] = v[6 + k];
    }
    if (j - j0 < (uint32_t)LSLOTS) {
      unsigned long long* s = sKeys + 12 * (j - j0);
      for (int k = 0; k < 3; k++) {
        atomicMin(s + k, v[k]);
        atomicMax(s + 3 + k, v[3 + k]);
        atomicMin(s + 6 + k, v[6 + k]);
        atomicMax(s + 9 + k, v[9 + k]);
      }
    } else {
      unsigned long long* g = b.keys + 12 * (size_t)j;
      for (int k = 0; k < 3; k++) {
        atomicMin(g + k, v[k]);
        atomicMax(g + 3 + k, v[3 + k]);
        atomicMin(g + 6 + k, v[6 + k]);
        atomicMax(g + 9 + k, v[9 + k]);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < LSLOTS * 12; k += BB) {
    uint32_t j = j0 + k / 12;
    if (j >= m) continue;
    int f = k % 12;
    unsigned long long x = sKeys[k];
    bool isMin = f < 3 || (f >= 6 && f < 9);
    if (x == (isMin ? KEY_POS_INF : KEY_NEG_INF)) continue;
    unsigned long long* g = b.keys + 12 * (size_t)j + f;
    if (isMin) atomicMin(g, x);
    else atomicMax(g, x);
  }
}

// Bounds of the tiny nodes: one lane per node, stored without atomics (the
// chunked k_bounds skips every small node).
__global__ void k_bounds_tiny(Build b, uint32_t lvStart, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    const uint32_t s = b.recStart[lvStart + j], e = b.recEnd[lvStart + j];
    if (e - s > TINY_NODE) continue;
    unsigned long long v[12];
    for (int k = 0; k < 12; k++)
      v[k] = (k < 3 || (k >= 6 && k < 9)) ? KEY_POS_INF : KEY_NEG_INF;
    for (uint32_t i = s; i < e; i++) {
      const double* x = b.box + 9 * (size_t)b.order[i];
      for (int k = 0; k < 3; k++) {
        unsigned long long lo = dkey(x[k]), hi = dkey(x[3 + k]);
        unsigned long long c = dkey(x[6 + k]);
        v[k] = lo < v[k] ? lo : v[k];
        v[3 + k] = hi > v[3 + k] ? hi : v[3 + k];
        v[6 + k] = c < v[6 + k] ? c : v[6 + k];
        v[9 + k] = c > v[9 + k] ? c : v[9 + k];
      }
    }
    for (int k = 0; k < 12; k++) b.keys[12 * (size_t)j + k] = v[k];
  }
}

// Bounds of the small nodes above TINY_NODE: one wave per node (wave w takes
// nodes w, w + waves, ...), keys reduced in registers and across the wave,
// stored without atomics.
__global__ void __launch_bounds__(BB) k_bounds_small(Build b, uint32_t lvStart, uint32_t m) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t j = wave; j < m; j += nwaves) {  // wave-uniform
    const uint32_t s = b.recStart[lvStart + j], e = b.recEnd[lvStart + j];
    if (e - s <= TINY_NODE || e - s > SMALL_NODE) continue;
    unsigned long long v[12];
    for (int k = 0; k < 12; k++)
      v[k] = (k < 3 || (k >= 6 && k < 9)) ? KEY_POS_INF : KEY_NEG_INF;
    for (uint32_t i = s + lane; i < e; i += 64) {
      const double* x = b.box + 9 * (size_t)b.order[i];
      for (int k = 0; k < 3; k++) {
        unsigned long long lo = dkey(x[k]), hi = dkey(x[3 + k]);
        unsigned long long c = dkey(x[6 + k]);
        v[k] = lo < v[k] ? lo : v[k];
        v[3 + k] = hi > v[3 + k] ? hi : v[3 + k];
        v[6 + k] = c < v[6 + k] ? c : v[6 + k];
        v[9 + k] = c > v[9 + k] ? c : v[9 + k];
      }
    }
    for (int k = 0; k < 3; k++) {
      v[k] = wave_min(v[k]);
      v[3 + k] = wave_max(v[3 + k]);
      v[6 + k] = wave_min(v[6 + k]);
      v[9 + k] = wave_max(v[9 + k]);
    }
    if (lane == 0)
      for (int k = 0; k < 12; k++) b.keys[12 * (size_t)j + k] = v[k];
  }
}

// Leaf / SAH / median decision (bvh.ts:149-160, 187-192, 284-289).
__global__ void k_setup(Build b, uint32_t lvStart, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    uint32_t rec = lvStart + j;
    const unsigned long long* kk = b.keys + 12 * (size_t)j;
    for (int k = 0; k < 6; k++) b.recBox[6 * (size_t)rec + k] = dval(kk[k]);
    uint32_t count = b.recEnd[rec] - b.recStart[rec];
    int32_t kd;
    if (count <= 4) {
      kd = KIND_LEAF;
    } else {
      bool any = false;
      for (int a = 0; a < 3; a++) {
        double cmin = dval(kk[6 + a]), cmax = dval(kk[9 + a]);
        double extent = cmax - cmin;
        b.cmin3[3 * (size_t)j + a] = cmin;
        if (extent > 1e-12) {
          b.tob3[3 * (size_t)j + a] = b.bins / extent;
          any = true;
        } else {
          b.tob3[3 * (size_t)j + a] = 0;
        }
      }
      if (any) {
        kd = KIND_SAH;
      } else {
        kd = KIND_MEDIAN;
        double ex = dval(kk[9]) - dval(kk[6]);
        double ey = dval(kk[10]) - dval(kk[7]);
        double ez = dval(kk[11]) - dval(kk[8]);
        b.axis[j] = ex >= ey && ex >= ez ? 0 : ey >= ez ? 1 : 2;
        atomicAdd(b.levelInfo + 1, 1u);
      }
    }
    b.kind[j] = kd;
    b.flag[j] = kd == KIND_SAH ? 1u : 0u;
  }
}

__global__ void k_bin_init(Build b, uint32_t lvStart, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  const int per = 3 * b.bins;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    if (b.kind[j] != KIND_SAH || small_node(b, lvStart + j)) continue;
    size_t s = (size_t)b.binSlot[j] * per;
    for (int q = 0; q < per; q++) {
      b.binCount[s + q] = 0;
      for (int k = 0; k < 3; k++) {
        b.binKey[6 * (s + q) + k] = KEY_POS_INF;
        b.binKey[6 * (s + q) + 3 + k] = KEY_NEG_INF;
      }
    }
  }
}

__device__ __forceinline__ int bucket_of(double c, double cmin, double tob, int bins) {
  double fb = floor((c - cmin) * tob);  // bvh.ts:200-202
  if (fb >= bins) return bins - 1;
  if (fb < 0) return 0;
  return (int)fb;
}

// SAH bucket histograms (bvh.ts:198-211).
__global__ void __launch_bounds__(BB) k_bin(Build b, uint32_t lvStart, uint32_t m) {
  __shared__ uint32_t sCount[LSLOTS * 3 * MAX_BINS];
  __shared__ unsigned long long sKey[LSLOTS * 3 * MAX_BINS * 6];
  __shared__ uint32_t sFirst;
  const int bins = b.bins;
  const int per = 3 * bins;
  uint32_t c0 = blockIdx.x * CHUNK;
  for (int k = threadIdx.x; k < LSLOTS * per; k += BB) {
    sCount[k] = 0;
    for (int f = 0; f < 3; f++) {
      sKey[6 * k + f] = KEY_POS_INF;
      sKey[6 * k + 3 + f] = KEY_NEG_INF;
    }
  }
  uint32_t j0 = block_first_node(b, c0, lvStart, m, &sFirst);
  for (uint32_t i = c0 + threadIdx.x; i < c0 + CHUNK && i < b.n; i += BB) {
    int32_t r = b.nodeOf[i];
    if (r < 0) continue;
    uint32_t j = (uint32_t)r - lvStart;
    if (b.kind[j] != KIND_SAH || small_node(b, (uint32_t)r)) continue;
    const double* x = b.box + 9 * (size_t)b.order[i];
    unsigned long long mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
      mn[k] = dkey(x[k]);
      mx[k] = dkey(x[3 + k]);
    }
    bool local = j - j0 < (uint32_t)LSLOTS;
    size_t gs = (size_t)b.binSlot[j] * per;
    for (int a = 0; a < 3; a++) {
      double tob = b.tob3[3 * (size_t)j + a];
      if (tob == 0) continue;
      int bk = bucket_of(x[6 + a], b.cmin3[3 * (size_t)j + a], tob, bins);
      if (local) {
        int q = (int)(j - j0) * per + a * bins + bk;
        atomicAdd(sCount + q, 1u);
        for (int k = 0; k < 3; k++) {
          atomicMin(sKey + 6 * q + k, mn[k]);
          atomicMax(sKey + 6 * q + 3 + k, mx[k]);
        }
      } else {
        size_t q = gs + a * bins + bk;
        atomicAdd(b.binCount + q, 1u);
        for (int k = 0; k < 3; k++) {
          atomicMin(b.binKey + 6 * q + k, mn[k]);
          atomicMax(b.binKey + 6 * q + 3 + k, mx[k]);
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < LSLOTS * per; k += BB) {
    uint32_t c = sCount[k];
    if (c == 0) continue;
    uint32_t j = j0 + k / per;
    if (j >= m || b.kind[j] != KIND_SAH) continue;
    size_t q = (size_t)b.binSlot[j] * per + (k % per);
    atomicAdd(b.binCount + q, c);
    for (int f = 0; f < 3; f++) {
      atomicMin(b.binKey + 6 * q + f, sKey[6 * k + f]);
      atomicMax(b.binKey + 6 * q + 3 + f, sKey[6 * k + 3 + f]);
    }
  }
}

__device__ __forceinline__ double surface_area(double dx, double dy, double dz) {
  return 2 * (dx * dy + dy * dz + dz * dx);  // bvh.ts:70
}

// Suffix/prefix sweeps of one axis of SAH node j (bvh.ts:213-257): the first
// split of strictly least cost, or split -1 when no split has both sides
// non-empty.  cnt/key: the node's 3*bins bucket counts and 6 keys per bucket.
__device__ void sweep_axis(const Build& b, uint32_t j, int a, const uint32_t* cnt0,
                           const unsigned long long* key0, double& bestCost, int& bestSplit) {
  const int bins = b.bins;
  bestCost = __builtin_inf();
  bestSplit = -1;
  if (b.tob3[3 * (size_t)j + a] == 0) return;
  double rightArea[MAX_BINS];
  uint32_t rightCount[MAX_BINS];
  const uint32_t* cnt = cnt0 + a * bins;
  const unsigned long long* key = key0 + 6 * (a * bins);
  double rx0 = __builtin_inf(), ry0 = rx0, rz0 = rx0;
  double rx1 = -__builtin_inf(), ry1 = rx1, rz1 = rx1;
  uint32_t rc = 0;
  for (int s = bins - 1; s >= 1; s--) {
    if (cnt[s] > 0) {
      double v;
      v = dval(key[6 * s + 0]); if (v < rx0) rx0 = v;
      v = dval(key[6 * s + 1]); if (v < ry0) ry0 = v;
      v = dval(key[6 * s + 2]); if (v < rz0) rz0 = v;
      v = dval(key[6 * s + 3]); if (v > rx1) rx1 = v;
      v = dval(key[6 * s + 4]); if (v > ry1) ry1 = v;
      v = dval(key[6 * s + 5]); if (v > rz1) rz1 = v;
      rc += cnt[s];
    }
    rightCount[s - 1] = rc;
    rightArea[s - 1] = rc > 0 ? surface_area(rx1 - rx0, ry1 - ry0, rz1 - rz0) : 0;
  }
  double lx0 = __builtin_inf(), ly0 = lx0, lz0 = lx0;
  double lx1 = -__builtin_inf(), ly1 = lx1, lz1 = lx1;
  uint32_t lc = 0;
  for (int s = 0; s < bins - 1; s++) {
    if (cnt[s] > 0) {
      double v;
      v = dval(key[6 * s + 0]); if (v < lx0) lx0 = v;
      v = dval(key[6 * s + 1]); if (v < ly0) ly0 = v;
      v = dval(key[6 * s + 2]); if (v < lz0) lz0 = v;
      v = dval(key[6 * s + 3]); if (v > lx1) lx1 = v;
      v = dval(key[6 * s + 4]); if (v > ly1) ly1 = v;
      v = dval(key[6 * s + 5]); if (v > lz1) lz1 = v;
      lc += cnt[s];
    }
    uint32_t rcs = rightCount[s];
    if (lc == 0 || rcs == 0) continue;
    double cost = (double)lc * surface_area(lx1 - lx0, ly1 - ly0, lz1 - lz0) +
                  (double)rcs * rightArea[s];
    if (cost < bestCost) {
      bestCost = cost;
      bestSplit = s;
    }
  }
}

// The sweeps of every SAH node, four lanes per node (lane a < 3 sweeps axis a),
// and the choice over the axes in the reference's order: a later axis wins
// only with a strictly lower cost (bvh.ts:249-256); no split on any axis
// falls back to the median split.
__global__ void k_sweep(Build b, uint32_t m) {
  const uint32_t stride = gridDim.x * blockDim.x;  // a multiple of 4
  for (uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x; tid < 4 * m; tid += stride) {
    // the four lanes of a group share j and run the same iterations
    const uint32_t j = tid >> 2;
    const int a = (int)(tid & 3);
    const bool sah = b.kind[j] == KIND_SAH;
    double cost = __builtin_inf();
    int split = -1;
    if (sah && a < 3) {
      size_t base = (size_t)b.binSlot[j] * 3 * b.bins;
      sweep_axis(b, j, a, b.binCount + base, b.binKey + 6 * base, cost, split);
    }
    double c1 = __shfl_down(cost, 1, 4), c2 = __shfl_down(cost, 2, 4);
    int s1 = __shfl_down(split, 1, 4), s2 = __shfl_down(split, 2, 4);
    if (!sah || a != 0) continue;
    int bestAxis = split >= 0 ? 0 : -1, bestSplit = split;
    double bestCost = cost;
    if (s1 >= 0 && c1 < bestCost) {
      bestCost = c1;
      bestAxis = 1;
      bestSplit = s1;
    }
    if (s2 >= 0 && c2 < bestCost) {
      bestCost = c2;
      bestAxis = 2;
      bestSplit = s2;
    }
    if (bestAxis < 0) {
      const unsigned long long* kk = b.keys + 12 * (size_t)j;
      double ex = dval(kk[9]) - dval(kk[6]);
      double ey = dval(kk[10]) - dval(kk[7]);
      double ez = dval(kk[11]) - dval(kk[8]);
      b.axis[j] = ex >= ey && ex >= ez ? 0 : ey >= ez ? 1 : 2;
      b.kind[j] = KIND_MEDIAN;
      atomicAdd(b.levelInfo + 1, 1u);
    } else {
      b.axis[j] = bestAxis;
      b.split[j] = bestSplit;
      b.origin[j] = b.cmin3[3 * (size_t)j + bestAxis];
      b.scale[j] = b.tob3[3 * (size_t)j + bestAxis];
    }
  }
}

// Completes this wave's LDS traffic before the wave reads what it wrote (the
// slice is the wave's own: no other wave touches it).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Buckets of the small SAH nodes: one wave per node (wave w takes nodes w,
// w + waves, ...), the histogram in the wave's own LDS slice, then stored
// without atomics (the chunked k_bin skips these nodes, and k_bin_init leaves
// their buckets to this kernel).
__global__ void __launch_bounds__(BB) k_bin_small(Build b, uint32_t lvStart, uint32_t m) {
  __shared__ uint32_t sCount[WAVES * 3 * MAX_BINS];
  __shared__ unsigned long long sKey[WAVES * 3 * MAX_BINS * 6];
  const int bins = b.bins;
  const int per = 3 * bins;
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  uint32_t* cnt = sCount + w * 3 * MAX_BINS;
  unsigned long long* key = sKey + (size_t)w * 3 * MAX_BINS * 6;
  for (uint32_t j = wave; j < m; j += nwaves) {  // wave-uniform
    const uint32_t s = b.recStart[lvStart + j], e = b.recEnd[lvStart + j];
    if (e - s > SMALL_NODE || b.kind[j] != KIND_SAH) continue;
    for (int q = (int)lane; q < per; q += 64) {
      cnt[q] = 0;
      for (int k = 0; k < 3; k++) {
        key[6 * q + k] = KEY_POS_INF;
        key[6 * q + 3 + k] = KEY_NEG_INF;
      }
    }
    wave_lds_sync();
    double cmin[3], tob[3];
    for (int a = 0; a < 3; a++) {
      cmin[a] = b.cmin3[3 * (size_t)j + a];
      tob[a] = b.tob3[3 * (size_t)j + a];
    }
    for (uint32_t i = s + lane; i < e; i += 64) {
      const double* x = b.box + 9 * (size_t)b.order[i];
      unsigned long long mn[3], mx[3];
      for (int k = 0; k < 3; k++) {
        mn[k] = dkey(x[k]);
        mx[k] = dkey(x[3 + k]);
      }
      for (int a = 0; a < 3; a++) {
        if (tob[a] == 0) continue;
        int q = a * bins + bucket_of(x[6 + a], cmin[a], tob[a], bins);
        atomicAdd(cnt + q, 1u);
        for (int k = 0; k < 3; k++) {
          atomicMin(key + 6 * q + k, mn[k]);
          atomicMax(key + 6 * q + 3 + k, mx[k]);
        }
      }
    }
    wave_lds_sync();
    const size_t gs = (size_t)b.binSlot[j] * per;
    for (int q = (int)lane; q < per; q += 64) {
      b.binCount[gs + q] = cnt[q];
      for (int k = 0; k < 6; k++) b.binKey[6 * (gs + q) + k] = key[6 * q + k];
    }
    wave_lds_sync();
  }
}

// Partition flag of each position of a SAH node (bvh.ts:264-269).
__global__ void k_flags(Build b, uint32_t lvStart) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    uint32_t g = 0;
    int32_t r = b.nodeOf[i];
    if (r >= 0) {
      uint32_t j = (uint32_t)r - lvStart;
      if (b.kind[j] == KIND_SAH) {
        double c = b.box[9 * (size_t)b.order[i] + 6 + b.axis[j]];
        int bk = bucket_of(c, b.origin[j], b.scale[j], b.bins);
        g = bk <= b.split[j] ? 1u : 0u;
      }
    }
    b.good[i] = g;
  }
}

// a_k (bads from the left) and b_k (goods from the right) per node.
__global__ void k_ab(Build b, uint32_t lvStart) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    int32_t r = b.nodeOf[i];
    if (r < 0) continue;
    uint32_t j = (uint32_t)r - lvStart;
    if (b.kind[j] != KIND_SAH) continue;
    uint32_t start = b.recStart[r], end = b.recEnd[r];
    uint32_t rel = i - start;
    uint32_t gBefore = b.goodScan[i] - b.goodScan[start];
    uint32_t gTotal = b.goodScan[end] - b.goodScan[start];
    if (b.good[i]) {
      uint32_t gAfter = gTotal - gBefore - 1;
      if (gAfter < end - start) b.B[start + gAfter] = (end - 1) - i;
    } else {
      uint32_t badBefore = rel - gBefore;
      if (badBefore < end - start) b.A[start + badBefore] = rel;
    }
  }
}

// Meeting point of the two-pointer loop per SAH node.
__global__ void k_meet(Build b, uint32_t lvStart, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    if (b.kind[j] != KIND_SAH) continue;
    uint32_t rec = lvStart + j;
    uint32_t start = b.recStart[rec], end = b.recEnd[rec];
    uint32_t n = end - start;
    uint32_t nGood = b.goodScan[end] - b.goodScan[start];
    uint32_t nBad = n - nGood;
    const uint32_t* A = b.A + start;
    const uint32_t* B = b.B + start;
    // K = number of complete (L-bad, R-good) pairs: max k with a_k + b_k + 2 <= n
    uint32_t lo = 0, hi = nBad < nGood ? nBad : nGood;
    while (lo < hi) {
      uint32_t mk = (lo + hi + 1) >> 1;  // test k = mk (1-based)
      if ((uint64_t)A[mk - 1] + B[mk - 1] + 2 <= n) lo = mk;
      else hi = mk - 1;
    }
    uint32_t K = lo;
    int64_t bK = K > 0 ? (int64_t)B[K - 1] : -1;
    uint32_t x;
    if (K < nBad && (int64_t)A[K] + bK + 2 <= (int64_t)n) x = A[K] + 1;
    else x = (uint32_t)((int64_t)n - (bK + 1));
    b.meetX[j] = x;
    b.mid[j] = start + nGood;
  }
}

// Scatter every position to its place after the partition (SAH nodes) or keep it.
__global__ void k_scatter(Build b, uint32_t lvStart) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    int32_t r = b.nodeOf[i];
    uint32_t dst = i;
    if (r >= 0) {
      uint32_t j = (uint32_t)r - lvStart;
      int32_t kd = b.kind[j];
      if (kd == KIND_MEDIAN) continue;  // written by k_median
      if (kd == KIND_SAH) {
        uint32_t start = b.recStart[r], end = b.recEnd[r];
        uint32_t n = end - start, rel = i - start;
        uint32_t gBefore = b.goodScan[i] - b.goodScan[start];
        uint32_t gTotal = b.goodScan[end] - b.goodScan[start];
        bool g = b.good[i] != 0;
        uint32_t x = b.meetX[j];
        uint32_t d;
        if (rel < x) {
          if (g) {
            d = rel;
          } else {
            uint32_t k = rel - gBefore;  // rank among L-bads
            int64_t bprev = k > 0 ? (int64_t)b.B[start + k - 1] : -1;
            d = (uint32_t)((int64_t)n - 2 - bprev);
          }
        } else {
          uint32_t rr = n - 1 - rel;  // R-index
          if (g) {
            uint32_t k = gTotal - gBefore - 1;  // rank among R-goods
            d = b.A[start + k];
          } else {
            d = n - 2 - rr;
          }
        }
        if (d >= n) {  // cannot happen for a consistent partition; never write out of range
          atomicOr(b.levelInfo + 2, 1u);
          continue;
        }
        dst = start + d;
      }
    }
    b.orderNext[dst] = b.order[i];
  }
}

// Stable sort of an all-degenerate node by centroid (bvh.ts:284-296): bitonic
// sort of (centroid, position) pairs, ties keep position order.
__global__ void __launch_bounds__(BB) k_median(Build b, uint32_t lvStart, uint32_t m) {
  __shared__ double sKey[MEDIAN_MAX];
  __shared__ uint32_t sPos[MEDIAN_MAX];
  for (uint32_t j = blockIdx.x; j < m; j += gridDim.x) {
    if (b.kind[j] != KIND_MEDIAN) continue;
    uint32_t rec = lvStart + j;
    uint32_t start = b.recStart[rec], end = b.recEnd[rec];
    uint32_t n = end - start;
    if (n > (uint32_t)MEDIAN_MAX) {  // sorted by the merge kernels, after the level
      if (threadIdx.x == 0) {
        b.mid[j] = (start + end) >> 1;
        atomicAdd(b.levelInfo + 3, 1u);
      }
      continue;
    }
    int axis = b.axis[j];
    uint32_t P2 = 1;
    while (P2 < n) P2 <<= 1;
    for (uint32_t t = threadIdx.x; t < P2; t += BB) {
      if (t < n) {
        sKey[t] = b.box[9 * (size_t)b.order[start + t] + 6 + axis];
        sPos[t] = t;
      } else {
        sKey[t] = __builtin_inf();
        sPos[t] = 0xffffffffu;
      }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= P2; k <<= 1) {
      for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
        for (uint32_t t = threadIdx.x; t < P2; t += BB) {
          uint32_t u = t ^ jj;
          if (u > t) {
            double ka = sKey[t], kb = sKey[u];
            uint32_t pa = sPos[t], pb = sPos[u];
            bool aFirst = (ka < kb) || (!(kb < ka) && pa < pb);
            bool up = (t & k) == 0;
            if (up != aFirst) {
              sKey[t] = kb; sKey[u] = ka;
              sPos[t] = pb; sPos[u] = pa;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t t = threadIdx.x; t < n; t += BB)
      b.orderNext[start + t] = b.order[start + sPos[t]];
    if (threadIdx.x == 0) b.mid[j] = (start + end) >> 1;
    __syncthreads();
  }
}

// Stable sort of a large all-degenerate node (more than MEDIAN_MAX primitives):
// pairs (centroid, position) with position the index in the node's range, so
// every pair is distinct and "a before b" is (ka < kb) || (!(kb < ka) && pa < pb),
// the order of the LDS sort and of a stable sort by centroid (-0 and +0 compare
// equal, as in the reference's comparator).  First every tile of MEDIAN_MAX
// positions is sorted in LDS, then runs of width w, 2w, ... are merged: an
// element's place is its index in its own run plus its rank in the partner
// run, found by binary search.  Runs cover contiguous position ranges, so a
// left-run element precedes every right-run element of equal centroid.
__device__ __forceinline__ bool pair_before(double ka, uint32_t pa, double kb, uint32_t pb) {
  return (ka < kb) || (!(kb < ka) && pa < pb);
}

__global__ void __launch_bounds__(BB) k_msort_tiles(Build b, uint32_t start, uint32_t n,
                                                    int axis, double* __restrict__ key,
                                                    uint32_t* __restrict__ pos) {
  __shared__ double sKey[MEDIAN_MAX];
  __shared__ uint32_t sPos[MEDIAN_MAX];
  for (uint32_t t0 = blockIdx.x * (uint32_t)MEDIAN_MAX; t0 < n;
       t0 += gridDim.x * (uint32_t)MEDIAN_MAX) {
    uint32_t len = min((uint32_t)MEDIAN_MAX, n - t0);
    for (uint32_t t = threadIdx.x; t < (uint32_t)MEDIAN_MAX; t += BB) {
      if (t < len) {
        sKey[t] = b.box[9 * (size_t)b.order[start + t0 + t] + 6 + axis];
        sPos[t] = t0 + t;
      } else {
        sKey[t] = __builtin_inf();
        sPos[t] = 0xffffffffu;
      }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= (uint32_t)MEDIAN_MAX; k <<= 1) {
      for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
        for (uint32_t t = threadIdx.x; t < (uint32_t)MEDIAN_MAX; t += BB) {
          uint32_t u = t ^ jj;
          if (u > t) {
            double ka = sKey[t], kb = sKey[u];
            uint32_t pa = sPos[t], pb = sPos[u];
            bool aFirst = pair_before(ka, pa, kb, pb);
            bool up = (t & k) == 0;
            if (up != aFirst) {
              sKey[t] = kb; sKey[u] = ka;
              sPos[t] = pb; sPos[u] = pa;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t t = threadIdx.x; t < len; t += BB) {
      key[t0 + t] = sKey[t];
      pos[t0 + t] = sPos[t];
    }
    __syncthreads();
  }
}

__global__ void k_msort_merge(uint32_t n, uint32_t w, const double* __restrict__ keyIn,
                              const uint32_t* __restrict__ posIn, double* __restrict__ keyOut,
                              uint32_t* __restrict__ posOut) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t run = i / w;
    uint32_t own0 = run * w;
    uint32_t part0 = (run ^ 1u) * w;
    double k = keyIn[i];
    uint32_t p = posIn[i];
    uint32_t rank = 0;
    if (part0 < n) {
      uint32_t lo = 0, hi = min(w, n - part0);  // partners in [part0, part0 + hi)
      while (lo < hi) {                         // count partners that come before (k, p)
        uint32_t mid = (lo + hi) >> 1;
        if (pair_before(keyIn[part0 + mid], posIn[part0 + mid], k, p)) lo = mid + 1;
        else hi = mid;
      }
      rank = lo;
    }
    uint32_t dst = min(own0, part0) + (i - own0) + rank;
    keyOut[dst] = k;
    posOut[dst] = p;
  }
}

__global__ void k_msort_emit(Build b, uint32_t j, uint32_t start, uint32_t end,
                             const uint32_t* __restrict__ pos) {
  uint32_t n = end - start;
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
    b.orderNext[start + t] = b.order[start + pos[t]];
  if (blockIdx.x == 0 && threadIdx.x == 0) b.mid[j] = (start + end) >> 1;
}

__global__ void k_split_flags(Build b, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
    b.flag[j] = b.kind[j] != KIND_LEAF ? 1u : 0u;
}

// Children of split nodes appended at lvEnd in parent order.
__global__ void k_children(Build b, uint32_t lvStart, uint32_t lvEnd, uint32_t m) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    uint32_t rec = lvStart + j;
    if (b.kind[j] == KIND_LEAF) {
      b.recLeft[rec] = -1;
      b.recRight[rec] = -1;
      continue;
    }
    uint32_t L = lvEnd + 2 * b.flagScan[j];
    uint32_t R = L + 1;
    uint32_t mid = b.mid[j];
    b.recLeft[rec] = (int32_t)L;
    b.recRight[rec] = (int32_t)R;
    b.recStart[L] = b.recStart[rec];
    b.recEnd[L] = mid;
    b.recStart[R] = mid;
    b.recEnd[R] = b.recEnd[rec];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) b.levelInfo[0] = b.flagScan[m];
}

__global__ void k_node_of(Build b, uint32_t lvStart) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    int32_t r = b.nodeOf[i];
    if (r < 0) continue;
    uint32_t j = (uint32_t)r - lvStart;
    if (b.kind[j] == KIND_LEAF) {
      b.nodeOf[i] = -1;
    } else {
      b.nodeOf[i] = i < b.mid[j] ? b.recLeft[r] : b.recRight[r];
    }
  }
}

// subtree node counts, bottom-up per level
__global__ void k_sizes(Build b, uint32_t r0, uint32_t r1) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x; r < r1; r += stride) {
    int32_t L = b.recLeft[r];
    b.recSize[r] = L < 0 ? 1u : 1u + b.recSize[L] + b.recSize[b.recRight[r]];
  }
}

// DFS pre-order indices, top-down per level (bvh.ts:111-112, 161-164)
__global__ void k_dfs(Build b, uint32_t r0, uint32_t r1) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x; r < r1; r += stride) {
    int32_t L = b.recLeft[r];
    if (L < 0) continue;
    uint32_t idx = b.recIdx[r];
    b.recIdx[L] = idx + 1;
    b.recIdx[b.recRight[r]] = idx + 1 + b.recSize[L];
  }
}

// One 8-float node per record (bvh.ts:140-164).
__global__ void k_emit(Build b, uint32_t nrec, float* __restrict__ nodes,
                       uint32_t* __restrict__ words) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += stride) {
    const double* bx = b.recBox + 6 * (size_t)r;
    uint32_t idx = b.recIdx[r];
    float* nd = nodes + 8 * (size_t)idx;
    nd[0] = (float)(bx[0] - fabs(bx[0]) * 1.2e-7 - 1e-9);
    nd[1] = (float)(bx[1] - fabs(bx[1]) * 1.2e-7 - 1e-9);
    nd[2] = (float)(bx[2] - fabs(bx[2]) * 1.2e-7 - 1e-9);
    nd[4] = (float)(bx[3] + fabs(bx[3]) * 1.2e-7 + 1e-9);
    nd[5] = (float)(bx[4] + fabs(bx[4]) * 1.2e-7 + 1e-9);
    nd[6] = (float)(bx[5] + fabs(bx[5]) * 1.2e-7 + 1e-9);
    nd[3] = (float)(idx + b.recSize[r]);
    if (b.recLeft[r] < 0) {
      uint32_t start = b.recStart[r], count = b.recEnd[r] - start;
      nd[7] = (float)((double)start * 8 + count);  // encodeLeaf (bvh.ts:60-62)
      if (words) words[idx] = start * 8 + count;
    } else {
      nd[7] = -1.0f;
      if (words) words[idx] = 0xffffffffu;
    }
  }
}

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src, uint32_t n) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace ct

using namespace ct;

extern "C" int ct_bvh_build(const double* d_min, const double* d_max, const double* d_cen,
                            uint32_t n, int32_t bins, float* d_nodes, uint32_t* h_nodeCount,
                            uint32_t* d_order, void* stream) {
  return ct_bvh_build_ex(d_min, d_max, d_cen, n, bins, d_nodes, h_nodeCount, d_order, nullptr,
                         stream);
}

namespace {

constexpr int SCRATCH_SECTIONS = 31;

// Device bytes of the build's working memory for n primitives and `bins` buckets.
size_t build_scratch_bytes(size_t N, int bins, size_t* offsets) {
  const size_t NR = 2 * N, M = N, NS = N / 5 + 1;
  const size_t sizes[] = {N * 4, N * 4, N * 4, (N + 1) * 4, N * 4, N * 4,
                          NR * 4, NR * 4, NR * 4, NR * 4, NR * 48, NR * 4, NR * 4,
                          M * 96, M * 4, M * 4, M * 4, M * 8, M * 8, M * 24, M * 24,
                          M * 4, (M + 1) * 4, (M + 1) * 4, M * 4, M * 4,
                          NS * 3 * (size_t)bins * 4, NS * 3 * (size_t)bins * 48, 16,
                          scan_tmp_entries(N + 1) * 4, N * 72};
  static_assert(sizeof(sizes) / sizeof(sizes[0]) == SCRATCH_SECTIONS, "one offset per section");
  size_t sz = 0, k = 0;
  for (size_t b : sizes) {
    if (offsets) offsets[k] = sz;
    k++;
    sz += al(b);
  }
  return sz;
}

// The build proper.  `mem` is working memory of build_scratch_bytes(n, bins)
// bytes (uninitialised: every kernel clears or writes what it reads), or null
// to allocate it on the stream and free it at the end.
int bvh_build(const double* d_min, const double* d_max, const double* d_cen, uint32_t n,
              int32_t bins, float* d_nodes, uint32_t* h_nodeCount, uint32_t* d_order,
              uint32_t* d_leafWords, hipStream_t st, char* mem);

}  // namespace

extern "C" int ct_bvh_build_ex(const double* d_min, const double* d_max, const double* d_cen,
                               uint32_t n, int32_t bins, float* d_nodes, uint32_t* h_nodeCount,
                               uint32_t* d_order, uint32_t* d_leafWords, void* stream) {
  return bvh_build(d_min, d_max, d_cen, n, bins, d_nodes, h_nodeCount, d_order, d_leafWords,
                   (hipStream_t)stream, nullptr);
}

// Working memory of ct_bvh_build_noalloc: one array per device, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a stand-in, not the device address): the address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) char scratch[CT_BVH_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratchMu;
char* g_scratchPtr[MAX_DEVICES] = {};

int scratch_for_current_device(char** out) {
  int dev = 0;
  CT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES)
    return set_error(CT_ELIMIT, "ct_bvh_build_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return check_hip(e, "ct_bvh_build_noalloc: hipGetSymbolAddress");
    g_scratchPtr[dev] = (char*)p;
  }
  *out = g_scratchPtr[dev];
  return CT_OK;
}
}  // namespace

extern "C" int ct_bvh_build_noalloc(const double* d_min, const double* d_max,
                                    const double* d_cen, uint32_t n, int32_t bins,
                                    float* d_nodes, uint32_t* h_nodeCount, uint32_t* d_order,
                                    uint32_t* d_leafWords, void* stream) {
  if (bins < 2 || bins > MAX_BINS)
    return set_error(CT_EINVAL, "ct_bvh_build_noalloc: bins must be in [2, 32]");
  size_t need = n ? build_scratch_bytes(n, bins, nullptr) : 0;
  if (need > (size_t)CT_BVH_NOALLOC_SCRATCH_BYTES) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "ct_bvh_build_noalloc: %u primitives at %d bins need %zu bytes of working "
             "memory; the static buffer holds %zu",
             n, bins, need, (size_t)CT_BVH_NOALLOC_SCRATCH_BYTES);
    return set_error(CT_ELIMIT, msg);
  }
  char* mem = nullptr;
  if (n) {
    int rc = scratch_for_current_device(&mem);
    if (rc) return rc;
  }
  return bvh_build(d_min, d_max, d_cen, n, bins, d_nodes, h_nodeCount, d_order, d_leafWords,
                   (hipStream_t)stream, mem);
}

namespace {

int bvh_build(const double* d_min, const double* d_max, const double* d_cen, uint32_t n,
              int32_t bins, float* d_nodes, uint32_t* h_nodeCount, uint32_t* d_order,
              uint32_t* d_leafWords, hipStream_t st, char* scratchMem) {
  if (!d_nodes || !h_nodeCount || (n && (!d_min || !d_max || !d_cen || !d_order)))
    return set_error(CT_EINVAL, "ct_bvh_build: null argument");
  if (bins < 2 || bins > MAX_BINS)
    return set_error(CT_EINVAL, "ct_bvh_build: bins must be in [2, 32]");
  if (n > (1u << 30)) return set_error(CT_ELIMIT, "ct_bvh_build: more than 2^30 primitives");
  if (n == 0) {
    // a single empty leaf (bvh.ts:93-100)
    float node[8] = {0, 0, 0, 1, 0, 0, 0, 0};
    uint32_t word = 0;
    CT_HIP(hipMemcpyAsync(d_nodes, node, sizeof node, hipMemcpyHostToDevice, st));
    if (d_leafWords)
      CT_HIP(hipMemcpyAsync(d_leafWords, &word, sizeof word, hipMemcpyHostToDevice, st));
    CT_HIP(hipStreamSynchronize(st));
    *h_nodeCount = 1;
    return CT_OK;
  }
  const size_t N = n, NR = 2 * N;  // records: 2n-1 at most
  size_t off[SCRATCH_SECTIONS];
  const size_t sz = build_scratch_bytes(N, bins, off);
  const bool owned = scratchMem == nullptr;
  char* mem = scratchMem;
  if (owned) CT_HIP(hipMallocAsync((void**)&mem, sz, st));
  const size_t oOrderB = off[0], oNodeOf = off[1], oGood = off[2], oGoodScan = off[3],
               oA = off[4], oB = off[5], oRecStart = off[6], oRecEnd = off[7],
               oRecLeft = off[8], oRecRight = off[9], oRecBox = off[10], oRecSize = off[11],
               oRecIdx = off[12], oKeys = off[13], oKind = off[14], oAxis = off[15],
               oSplit = off[16], oOrigin = off[17], oScale = off[18], oCmin3 = off[19],
               oTob3 = off[20], oFlag = off[21], oFlagScan = off[22], oBinSlot = off[23],
               oMid = off[24], oMeetX = off[25], oBinCount = off[26], oBinKey = off[27],
               oLevelInfo = off[28], oScanTmp = off[29], oBox = off[30];
  Build b;
  b.bmin = d_min;
  b.bmax = d_max;
  b.bcen = d_cen;
  b.n = n;
  b.bins = bins;
  b.order = d_order;
  b.orderNext = (uint32_t*)(mem + oOrderB);
  b.nodeOf = (int32_t*)(mem + oNodeOf);
  b.good = (uint32_t*)(mem + oGood);
  b.goodScan = (uint32_t*)(mem + oGoodScan);
  b.A = (uint32_t*)(mem + oA);
  b.B = (uint32_t*)(mem + oB);
  b.recStart = (uint32_t*)(mem + oRecStart);
  b.recEnd = (uint32_t*)(mem + oRecEnd);
  b.recLeft = (int32_t*)(mem + oRecLeft);
  b.recRight = (int32_t*)(mem + oRecRight);
  b.recBox = (double*)(mem + oRecBox);
  b.recSize = (uint32_t*)(mem + oRecSize);
  b.recIdx = (uint32_t*)(mem + oRecIdx);
  b.keys = (unsigned long long*)(mem + oKeys);
  b.kind = (int32_t*)(mem + oKind);
  b.axis = (int32_t*)(mem + oAxis);
  b.split = (int32_t*)(mem + oSplit);
  b.origin = (double*)(mem + oOrigin);
  b.scale = (double*)(mem + oScale);
  b.cmin3 = (double*)(mem + oCmin3);
  b.tob3 = (double*)(mem + oTob3);
  b.flag = (uint32_t*)(mem + oFlag);
  b.flagScan = (uint32_t*)(mem + oFlagScan);
  b.binSlot = (uint32_t*)(mem + oBinSlot);
  b.mid = (uint32_t*)(mem + oMid);
  b.meetX = (uint32_t*)(mem + oMeetX);
  b.binCount = (uint32_t*)(mem + oBinCount);
  b.binKey = (unsigned long long*)(mem + oBinKey);
  b.levelInfo = (uint32_t*)(mem + oLevelInfo);
  b.scanTmp = (uint32_t*)(mem + oScanTmp);
  b.box = (double*)(mem + oBox);

  int rc = CT_OK;
  std::vector<std::pair<uint32_t, uint32_t>> levels;
  auto fail = [&](int code) {
    if (owned) (void)hipFreeAsync(mem, st);
    (void)hipStreamSynchronize(st);
    return code;
  };
#define BVH_LAUNCH(kern, grid, block, ...)                              \
  do {                                                                  \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, __VA_ARGS__); \
    hipError_t e_ = hipGetLastError();                                  \
    if (e_ != hipSuccess) return fail(check_hip(e_, #kern));            \
  } while (0)

  const unsigned gN = grid_for(N, BB);
  const unsigned gChunks = (unsigned)((N + CHUNK - 1) / CHUNK);
  BVH_LAUNCH(k_init_root, 1, 1, b);
  BVH_LAUNCH(k_set_u32, 1, 1, b.levelInfo + 2, 0u);
  BVH_LAUNCH(k_init_positions, gN, BB, b);
  uint32_t lvStart = 0, lvEnd = 1;
  while (lvEnd > lvStart) {
    uint32_t m = lvEnd - lvStart;
    if (levels.size() >= 8192)
      return fail(set_error(CT_ELIMIT, "ct_bvh_build: tree deeper than 8192 levels"));
    levels.push_back({lvStart, lvEnd});
    unsigned gM = grid_for(m, BB);
    BVH_LAUNCH(k_level_init, gM, BB, b, m);
    // waves of the per-node kernels: one per node, at most 32768
    const unsigned gW = (unsigned)((std::min<uint32_t>(m, 32768u) + WAVES - 1) / WAVES);
    BVH_LAUNCH(k_bounds, gChunks, BB, b, lvStart, m);
    BVH_LAUNCH(k_bounds_tiny, gM, BB, b, lvStart, m);
    BVH_LAUNCH(k_bounds_small, gW, BB, b, lvStart, m);
    BVH_LAUNCH(k_setup, gM, BB, b, lvStart, m);
    rc = scan_u32(b.flag, b.binSlot, m, b.scanTmp, st);
    if (rc) return fail(rc);
    BVH_LAUNCH(k_bin_init, gM, BB, b, lvStart, m);
    BVH_LAUNCH(k_bin, gChunks, BB, b, lvStart, m);
    BVH_LAUNCH(k_bin_small, gW, BB, b, lvStart, m);
    BVH_LAUNCH(k_sweep, grid_for(4 * (size_t)m, BB), BB, b, m);
    BVH_LAUNCH(k_flags, gN, BB, b, lvStart);
    rc = scan_u32(b.good, b.goodScan, n, b.scanTmp, st);
    if (rc) return fail(rc);
    BVH_LAUNCH(k_ab, gN, BB, b, lvStart);
    BVH_LAUNCH(k_meet, gM, BB, b, lvStart, m);
    BVH_LAUNCH(k_scatter, gN, BB, b, lvStart);
    // median nodes up to MEDIAN_MAX positions are sorted here (the kernel
    // passes over a level without any quickly); it sets every median node's
    // split position, so the children can be laid out before larger ones are
    // sorted
    BVH_LAUNCH(k_median, std::min<uint32_t>(m, 4096u), BB, b, lvStart, m);
    BVH_LAUNCH(k_split_flags, gM, BB, b, m);
    rc = scan_u32(b.flag, b.flagScan, m, b.scanTmp, st);
    if (rc) return fail(rc);
    BVH_LAUNCH(k_children, gM, BB, b, lvStart, lvEnd, m);
    BVH_LAUNCH(k_node_of, gN, BB, b, lvStart);
    // one host round trip per level: the split count (the next level's size)
    // and whether any median node is too large for the LDS sort
    uint32_t info[4] = {0, 0, 0, 0};
    {
      hipError_t e = hipMemcpyAsync(info, b.levelInfo, sizeof info, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return fail(check_hip(e, "ct_bvh_build: level info"));
    }
    if (info[3]) {
      // nodes above the LDS sort's size: tiled merge sort, one node at a time,
      // in the level's key buffer (free once k_setup has read it; 96 B per
      // position, the sort takes 24), into the next level's order
      std::vector<int32_t> kinds(m);
      std::vector<uint32_t> st0(m), en0(m);
      hipError_t e = hipMemcpyAsync(kinds.data(), b.kind, m * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipMemcpyAsync(st0.data(), b.recStart + lvStart, m * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipMemcpyAsync(en0.data(), b.recEnd + lvStart, m * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return fail(check_hip(e, "ct_bvh_build: median check"));
      double* keyA = (double*)b.keys;
      double* keyB = keyA + N;
      uint32_t* posA = (uint32_t*)(keyB + N);
      uint32_t* posB = posA + N;
      for (uint32_t j = 0; j < m; j++) {
        uint32_t len = en0[j] - st0[j];
        if (kinds[j] != KIND_MEDIAN || len <= (uint32_t)MEDIAN_MAX) continue;
        int32_t ax = 0;
        hipError_t ea = hipMemcpyAsync(&ax, b.axis + j, 4, hipMemcpyDeviceToHost, st);
        if (ea == hipSuccess) ea = hipStreamSynchronize(st);
        if (ea != hipSuccess) return fail(check_hip(ea, "ct_bvh_build: median axis"));
        unsigned tiles = (len + MEDIAN_MAX - 1) / MEDIAN_MAX;
        BVH_LAUNCH(k_msort_tiles, std::min(tiles, 4096u), BB, b, st0[j], len, (int)ax, keyA,
                   posA);
        double *ki = keyA, *ko = keyB;
        uint32_t *pi = posA, *po = posB;
        for (uint32_t w = MEDIAN_MAX; w < len; w *= 2) {
          BVH_LAUNCH(k_msort_merge, grid_for(len, BB), BB, len, w, ki, pi, ko, po);
          std::swap(ki, ko);
          std::swap(pi, po);
        }
        BVH_LAUNCH(k_msort_emit, grid_for(len, BB), BB, b, j, st0[j], en0[j], pi);
      }
    }
    std::swap(b.order, b.orderNext);
    if (info[2]) return fail(set_error(CT_EDEVICE, "ct_bvh_build: inconsistent partition"));
    uint32_t nSplit = info[0];
    lvStart = lvEnd;
    lvEnd = lvEnd + 2 * nSplit;
    if (lvEnd > NR) return fail(set_error(CT_EDEVICE, "ct_bvh_build: record overflow"));
  }
  uint32_t nrec = lvStart;
  // the final order must end up in the caller's buffer
  if (b.order != d_order) BVH_LAUNCH(k_copy_u32, gN, BB, d_order, b.order, n);
  for (size_t L = levels.size(); L-- > 0;)
    BVH_LAUNCH(k_sizes, grid_for(levels[L].second - levels[L].first, BB), BB, b,
               levels[L].first, levels[L].second);
  BVH_LAUNCH(k_set_u32, 1, 1, b.recIdx, 0u);
  for (size_t L = 0; L < levels.size(); L++)
    BVH_LAUNCH(k_dfs, grid_for(levels[L].second - levels[L].first, BB), BB, b,
               levels[L].first, levels[L].second);
  BVH_LAUNCH(k_emit, grid_for(nrec, BB), BB, b, nrec, d_nodes, d_leafWords);
#undef BVH_LAUNCH
  hipError_t e = owned ? hipFreeAsync(mem, st) : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return check_hip(e, "ct_bvh_build: finish");
  *h_nodeCount = nrec;
  return CT_OK;
}

}  // namespace
