// Device-side structures of the wavefront engine shared by the FP64 kernels
// (reference arithmetic) and the FP32 precision kernels: the job table, the
// SoA path state and the wave64 queue helpers.
#pragma once

#include "ct_engine.h"

namespace ct {

constexpr int RBLOCK = 256;
constexpr int MAX_BOUNCE_SLOTS = 256;  // counters per kind (bounces + 1 <= this)

// One job of a launch round (device copy).
struct JobDev {
  DevScene S;
  Frame F;
  TraceOpts opts;
  uint32_t width, height, npix;
  uint32_t passes;     // passes of this job in the round
  uint32_t frameBase;  // frame index of its first pass (uint32 arithmetic)
  uint32_t pathBase;   // first path id
  uint32_t pixelBase;  // first finalize index
  int doScale;         // multiply the sum by `scale` after adding (renderImage)
  double scale;
  float* accum;        // width*height*3 float32 sum
  float* comp;         // width*height*3 compensation terms (FP32 precision form)
  int fold;            // add the compensation into the sum after this round (FP32)
  // radianceInto jobs (binary64): the paths start from these rays (6 doubles
  // each, one pass, width = ray count) with these RNG states, and end with
  // their radiance (3 doubles each) and RNG states written back; accum unused
  const double* rays;
  uint32_t* rngIO;
  double* radiance;
};

struct PathSoA {
  double *ox, *oy, *oz, *dx, *dy, *dz;
  double *br, *bg, *bb;
  double *Lr, *Lg, *Lb;
  double *prevPdf, *hitT;
  uint32_t* rng;
  int32_t* hitPrim;
  uint32_t* jobOf;
  uint8_t* specular;
  uint32_t *q0, *q1;
  // shadow queue (one slot per live path at most)
  uint32_t* sPath;
  double *sdx, *sdy, *sdz, *sTmax, *sCr, *sCg, *sCb;
  // counters: live[b] at [b], shadow[b] at [MAX_BOUNCE_SLOTS + b], ray-fetch
  // cursors of the extend/shadow launches at [2*MAX + b] and [3*MAX + b], one
  // such block per path group; 64-bit work counters at ctr64
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

// FP32 path state of the precision form (same queues and counters).
struct PathSoA32 {
  float *ox, *oy, *oz, *dx, *dy, *dz;
  float *br, *bg, *bb;
  float *Lr, *Lg, *Lb;
  float *prevPdf, *hitT;
  uint32_t* rng;
  int32_t* hitPrim;
  uint32_t* jobOf;
  uint8_t* specular;
  uint32_t *q0, *q1;
  uint32_t* sPath;
  float *sdx, *sdy, *sdz, *sTmax, *sCr, *sCg, *sCb;
  uint32_t* ctr;
  unsigned long long* ctr64;
};

// Job of a path / finalize index: the last job whose base is <= i.
template <bool MULTI, bool PIXEL>
__device__ __forceinline__ int job_of(const JobDev* __restrict__ jobs, int njobs, size_t i) {
  if (!MULTI) return 0;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    size_t base = PIXEL ? jobs[mid].pixelBase : jobs[mid].pathBase;
    if (base <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Lane refill shared by the persistent traversal kernels: lanes without a ray
// take the next queue slots (one atomic per wave).  Returns the slot for this
// lane or UINT32_MAX.
__device__ __forceinline__ uint32_t refill(bool need, uint32_t n, uint32_t* fetch,
                                           bool& exhausted) {
  unsigned long long m = __ballot(need);
  if (!m) return 0xffffffffu;
  const int lane = __lane_id();
  int leader = __ffsll((long long)__ballot(1)) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(fetch, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (base + (uint32_t)__popcll(m) >= n) exhausted = true;
  if (!need) return 0xffffffffu;
  uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  return slot < n ? slot : 0xffffffffu;
}

// Leaf parking of the persistent walks.  A lane that reaches a leaf parks it
// and keeps walking; the wave leaves its interior phase once every lane holds
// at least PARK_LEAVES parked leaves (or has nothing left), and lanes that
// already hold them go on walking and parking, up to PARK_SLOTS, while the
// wave waits for the others: steps that would idle find leaves ahead of time.
// Then the wave tests all parked leaves together.  Parking only delays when
// the closest distance shrinks: the walk visits a superset of the unparked
// walk's boxes and returns the same minimum.  Measured on the 1080p soup frame
// (binary64 / FP32 Mrays/s): exit 3 with 3 slots 1036 / 1115, exit 3 with 5
// slots 1065 / 1160, exit 2 with 5 slots 1083 / 1214, exit 1 with 4 slots
// 1131 / 1236 (3 slots: 1128 / 1230; 2 slots: 1051 / 1140).  Since the
// traversal grids were retuned, the binary64 extension kernel does best with
// 3 slots (1166 against 1152 with 4), the closest-hit queries with 4 (3
// slots: 872 against 885; 5 slots: 862; exit 2: 842).  Once a triangle's
// loads were issued together the FP32 extension kernel moved to 5 slots
// (1350 against 1339 with 4; the binary64 one stays at 3: 1240 with 4; the
// queries stay at 4: 3 / 4 / 5 slots 928 / 937 / 913).
#ifndef CT_PARK_LEAVES
#define CT_PARK_LEAVES 1
#endif
constexpr int PARK_LEAVES = CT_PARK_LEAVES;
#ifndef CT_PARK_SLOTS
#define CT_PARK_SLOTS 4
#endif
constexpr int PARK_SLOTS = CT_PARK_SLOTS;  // closest-hit queries
#ifndef CT_PARK_SLOTS64
#define CT_PARK_SLOTS64 3
#endif
constexpr int PARK_SLOTS64 = CT_PARK_SLOTS64;  // binary64 extension rays
#ifndef CT_PARK_SLOTS32
#define CT_PARK_SLOTS32 5
#endif
constexpr int PARK_SLOTS32 = CT_PARK_SLOTS32;  // FP32 extension rays
static_assert(PARK_SLOTS >= PARK_LEAVES && PARK_SLOTS64 >= PARK_LEAVES &&
                  PARK_SLOTS32 >= PARK_LEAVES,
              "parking slots");
// The same for any-hit (NEE) rays, whose first blocking primitive ends the
// ray; their slots differ between the binary64 and the FP32 kernels (with
// exit 1, binary64 with 2 / 3 / 4 slots: 1158 / 1160 / 1144; FP32 with
// 3 / 4 / 5 slots: 1244 / 1240 / 1237).
#ifndef CT_SHADOW_PARK_LEAVES
#define CT_SHADOW_PARK_LEAVES 1
#endif
constexpr int SHADOW_PARK_LEAVES = CT_SHADOW_PARK_LEAVES;
#ifndef CT_SHADOW_PARK_SLOTS
#define CT_SHADOW_PARK_SLOTS 3
#endif
#ifndef CT_SHADOW_PARK_SLOTS32
#define CT_SHADOW_PARK_SLOTS32 4
#endif
constexpr int SHADOW_PARK_SLOTS = CT_SHADOW_PARK_SLOTS;
constexpr int SHADOW_PARK_SLOTS32 = CT_SHADOW_PARK_SLOTS32;
static_assert(SHADOW_PARK_LEAVES >= 1 && SHADOW_PARK_LEAVES <= SHADOW_PARK_SLOTS &&
                  SHADOW_PARK_LEAVES <= SHADOW_PARK_SLOTS32,
              "parking slots");

// A lane's parked leaves; N slots, held in registers.
template <int N>
struct Parked {
  int w[N];  // leaf words
  int n;
};

template <int N>
__device__ __forceinline__ void park(Parked<N>& p, int word) {
#pragma unroll
  for (int k = 0; k < N; k++)
    if (p.n == k) p.w[k] = word;
  p.n++;
}


// Second stream for the NEE launches.  The shadow rays of bounce b and the
// extension rays of bounce b+1 are independent (one adds to the path's
// radiance, the other writes its hit), so the two persistent launches run side
// by side and each fills the other's tail; the shading of bounce b+1 waits for
// both.  `side` null: one stream, launches in order.
struct SideStream {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// Work queued on `st` so far happens before what is queued on the side stream next.
inline hipError_t side_fork(const SideStream& s, hipStream_t st) {
  hipError_t e = hipEventRecord(s.fork, st);
  return e == hipSuccess ? hipStreamWaitEvent(s.side, s.fork, 0) : e;
}
// Work queued on the side stream so far happens before what is queued on `st` next.
inline hipError_t side_join(const SideStream& s, hipStream_t st) {
  hipError_t e = hipEventRecord(s.join, s.side);
  return e == hipSuccess ? hipStreamWaitEvent(st, s.join, 0) : e;
}

// Streams of one round.  The round's paths are traced as `groups` (1 or 2)
// independent groups, each on its own stream (group 0 on the caller's stream,
// group 1 on `split.side`, forked from it at the start and joined before the
// sums are finalized), so that one group's shading overlaps the other's
// traversal and each launch's tail is filled by the other group's work.
// Within a group the NEE launches may run on a side stream of their own.
struct RoundStreams {
  int groups = 1;
  SideStream split;
  SideStream shadow[2];
};

// Counters per path group: live[b], shadow[b], extend and shadow fetch cursors.
constexpr int GROUP_COUNTERS = 4 * MAX_BOUNCE_SLOTS;

// Parts of a round's launches: clearing the counters, each path group's
// launches (generate, then per bounce extend, shade and NEE), and the
// per-pixel sums once every group is done.  A captured round holds one graph
// per path group, launched on that group's stream between the other parts.
enum RoundPart { ROUND_CLEAR = 0, ROUND_GROUP0 = 1, ROUND_GROUP1 = 2, ROUND_FINAL = 3 };

// Paths [base, base + count) of path group g of a round of `paths` paths.
inline void group_range(const RoundStreams& rs, size_t paths, int g, size_t& base,
                        size_t& count) {
  size_t half = rs.groups == 2 ? paths / 2 : paths;
  base = g == 0 ? 0 : half;
  count = g == 0 ? half : paths - half;
}

// The queue and counter view of path group g whose paths start at `base`:
// queues and the shadow queue hold the group's entries from `base` on; the
// per-path arrays stay indexed by path id.
template <typename SoA>
inline SoA group_view(const SoA& P, size_t base, int g) {
  SoA Q = P;
  Q.q0 += base; Q.q1 += base; Q.sPath += base;
  Q.sdx += base; Q.sdy += base; Q.sdz += base; Q.sTmax += base;
  Q.sCr += base; Q.sCg += base; Q.sCb += base;
  Q.ctr += (size_t)g * GROUP_COUNTERS;
  return Q;
}

// One launch round of the FP32 precision kernels (ct_render32.hip).
struct RoundLaunch {
  PathSoA32 P32;
  const JobDev* jobs;
  int njobs;
  size_t paths, pix;
  int nb;  // bounce slots (maxBounces + 1)
  int stack, block;
  unsigned persistBlocks;
  bool counting;
  // profiling hook: kind 0 extend, 1 shade, 2 shadow
  void (*mark)(void* ctx, hipStream_t st, int kind, bool begin);
  void* markCtx;
  RoundStreams rs;
};
int launch_part32(const RoundLaunch& R, bool multi, int part, hipStream_t st);

}  // namespace ct
