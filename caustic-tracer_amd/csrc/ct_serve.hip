// Serving loop: continuous batching of render jobs over resident scenes.
//
// One worker thread per server owns a tracer and a stream.  Each iteration
// admits queued jobs while a slot and HBM for their progressive buffer are
// free, then traces one launch round that takes passes from every admitted
// job (admission order, or shortest remaining work first), so jobs that
// arrive while a round is on the GPU join the next one.  A job's passes are
// added to its buffer in pass order whatever rounds they fall in, and the
// final 1/spp scale is applied with its last pass, so the buffer ends equal
// to renderImage of the job alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ct_engine.h"

namespace ct {
namespace {

using Clock = std::chrono::steady_clock;

struct SJob {
  uint64_t id = 0;
  ct_job desc{};
  size_t bytes = 0;
  float* accum = nullptr;
  float* comp = nullptr;  // compensation terms (FP32 precision form)
  uint32_t done = 0;      // passes added
  uint32_t inFlight = 0;  // passes in the round being traced
  int state = CT_JOB_QUEUED;
  bool pending = false;   // admitted; the worker is allocating and clearing its buffer
  bool release = false;   // free once the round in flight ends
  bool charged = false;   // counted in the server's HBM in use
  double submit = 0, admit = -1, finish = -1;
  std::string error;
  uint64_t remaining() const {
    return (uint64_t)desc.width * desc.height * (desc.spp - done);
  }
};

}  // namespace
}  // namespace ct

using namespace ct;

struct ct_server {
  ct_server_opts opts{};
  int device = 0;
  ct_tracer* T = nullptr;
  hipStream_t stream = nullptr;      // launch rounds
  hipStream_t readStream = nullptr;  // client copies (never the legacy stream: the
                                     // worker may be capturing a graph meanwhile)
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cvWork, cvDone;
  std::deque<SJob*> queue;
  std::vector<SJob*> active;  // admission order
  std::unordered_map<uint64_t, SJob*> jobs;
  std::vector<SJob*> released;  // buffers the worker frees between rounds
  uint64_t nextId = 1;
  bool stop = false;
  std::thread worker;
  ct_server_stats st{};
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Worker thread only (or after it has stopped): device frees never overlap
// a graph capture of the worker's rounds.
void free_job(ct_server* S, SJob* j) {
  if (j->accum) (void)hipFree(j->accum);  // one allocation holds the sum and the terms
  if (j->charged) S->st.hbmInUse -= j->bytes;
  delete j;
}

void free_released(ct_server* S) {
  for (SJob* j : S->released) free_job(S, j);
  S->released.clear();
}

void drop_active(ct_server* S, SJob* j) {
  S->active.erase(std::remove(S->active.begin(), S->active.end(), j), S->active.end());
}

void serve_loop(ct_server* S) {
  (void)hipSetDevice(S->device);
  const size_t capacity = tracer_capacity(S->T);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cvWork.wait(lk, [&] {
      return S->stop || !S->queue.empty() || !S->active.empty() || !S->released.empty();
    });
    if (S->stop) break;
    free_released(S);
    // admission: FIFO while a slot and HBM for the buffer are free
    std::vector<SJob*> fresh;
    while (!S->queue.empty()) {
      SJob* j = S->queue.front();
      if (S->active.size() >= S->opts.maxActive ||
          S->st.hbmInUse + j->bytes > S->opts.hbmBudget)
        break;
      S->queue.pop_front();
      j->state = CT_JOB_RUNNING;
      j->pending = true;  // no client touches the buffer until it is allocated and zeroed
      j->admit = S->now();
      S->st.hbmInUse += j->bytes;
      j->charged = true;
      S->active.push_back(j);
      fresh.push_back(j);
    }
    if (S->active.empty()) {
      // nothing admitted and nothing running: wait for a release or a stop
      S->cvWork.wait(lk, [&] {
        return S->stop || !S->released.empty() ||
               (!S->queue.empty() &&
                S->st.hbmInUse + S->queue.front()->bytes <= S->opts.hbmBudget);
      });
      continue;
    }
    // one round: passes from every admitted job, up to the path capacity
    std::vector<SJob*> order = S->active;
    if (S->opts.policy == CT_SERVE_SRPT) {
      std::stable_sort(order.begin(), order.end(),
                       [](const SJob* a, const SJob* b) { return a->remaining() < b->remaining(); });
    } else if (S->opts.policy == CT_SERVE_SCENE) {
      // scenes ranked by their oldest admitted job; each scene's jobs together
      std::unordered_map<const ct_scene*, size_t> rank;
      for (SJob* j : order) rank.emplace(j->desc.scene, rank.size());
      std::stable_sort(order.begin(), order.end(), [&](const SJob* a, const SJob* b) {
        return rank[a->desc.scene] < rank[b->desc.scene];
      });
    }
    size_t budget = capacity;
    std::vector<SJob*> picked;
    std::vector<PlanJob> plan;
    uint64_t samples = 0;
    for (SJob* j : order) {
      const ct_job& d = j->desc;
      size_t npix = (size_t)d.width * d.height;
      uint32_t left = d.spp - j->done;
      size_t fit = budget / npix;
      if (left == 0 || fit == 0) continue;
      uint32_t p = (uint32_t)std::min<size_t>(left, fit);
      if (S->opts.slicePasses) p = std::min(p, S->opts.slicePasses);
      PlanJob pj{d.scene, make_frame(d.frame), {d.maxBounces, d.russianRoulette != 0, d.nee != 0},
                 d.width, d.height, d.seed * 7919u, j->done, p, nullptr,
                 j->done + p == d.spp, 1.0 / d.spp, nullptr};
      plan.push_back(pj);
      picked.push_back(j);
      j->inFlight = p;
      budget -= p * npix;
      samples += (uint64_t)p * npix;
    }
    lk.unlock();
    // buffers of newly admitted jobs (allocation outside the lock)
    for (SJob* j : fresh) {
      hipError_t e = hipMalloc((void**)&j->accum, j->bytes);
      if (e != hipSuccess) {
        j->accum = nullptr;
        j->error = std::string("hipMalloc of a progressive buffer: ") + hipGetErrorString(e);
        continue;
      }
      if (zero_floats(j->accum, j->bytes / sizeof(float), S->stream) != CT_OK)
        j->error = ct_last_error();
      if (S->opts.precision == CT_PRECISION_F32)
        j->comp = j->accum + (size_t)j->desc.width * j->desc.height * 3;
    }
    // the clears complete before any client may read a fresh buffer (reads go
    // through another stream, which is not ordered after this one)
    if (!fresh.empty()) {
      hipError_t e = hipStreamSynchronize(S->stream);
      if (e != hipSuccess)
        for (SJob* j : fresh)
          if (j->error.empty()) j->error = std::string("clearing a progressive buffer: ") +
                                           hipGetErrorString(e);
    }
    std::vector<PlanJob> run;
    for (size_t k = 0; k < picked.size(); k++) {
      if (!picked[k]->accum || !picked[k]->error.empty()) continue;
      plan[k].accum = picked[k]->accum;
      plan[k].comp = picked[k]->comp;
      run.push_back(plan[k]);
    }
    Clock::time_point t = Clock::now();
    int rc = run.empty() ? CT_OK : run_plan(S->T, run, S->stream);
    std::string err = rc ? ct_last_error() : "";
    ct_render_stats rs;
    memset(&rs, 0, sizeof rs);
    if (!rc) rc = ct_tracer_stats(S->T, &rs);
    double busy = std::chrono::duration<double>(Clock::now() - t).count();
    lk.lock();
    S->st.rounds++;
    S->st.busySec += busy;
    if (!rc) {
      S->st.paths += rs.paths;
      S->st.extensionRays += rs.extensionRays;
      S->st.shadowRays += rs.shadowRays;
      S->st.pixelSamples += samples;
    }
    double now = S->now();
    // fresh jobs this round did not reach: their buffers are ready now (before the
    // picked jobs, which may be freed below)
    for (SJob* j : fresh) {
      if (std::find(picked.begin(), picked.end(), j) != picked.end()) continue;
      j->pending = false;
      if (!j->error.empty() || !j->accum) {
        if (j->error.empty()) j->error = "no progressive buffer";
        j->state = CT_JOB_FAILED;
        j->finish = now;
        S->st.failed++;
        drop_active(S, j);
      }
      if (j->release) {
        if (j->state == CT_JOB_RUNNING) drop_active(S, j);
        S->jobs.erase(j->id);
        free_job(S, j);
      }
    }
    for (SJob* j : picked) {
      uint32_t p = j->inFlight;
      j->inFlight = 0;
      j->pending = false;
      bool failed = !j->error.empty() || !j->accum || rc;
      if (failed) {
        if (j->error.empty()) j->error = err.empty() ? "render round failed" : err;
        j->state = CT_JOB_FAILED;
        j->finish = now;
        S->st.failed++;
        drop_active(S, j);
      } else {
        j->done += p;
        if (j->done == j->desc.spp) {
          j->state = CT_JOB_DONE;
          j->finish = now;
          S->st.completed++;
          drop_active(S, j);
        }
      }
      if (j->release) {
        if (j->state == CT_JOB_RUNNING) drop_active(S, j);
        S->jobs.erase(j->id);
        free_job(S, j);
      }
    }
    S->cvDone.notify_all();
  }
}

SJob* find(ct_server* S, uint64_t id) {
  auto it = S->jobs.find(id);
  return it == S->jobs.end() ? nullptr : it->second;
}

void fill_status(const ct_server* S, const SJob* j, ct_job_status* out) {
  (void)S;
  out->state = j->state;
  out->passesDone = j->done;
  out->spp = j->desc.spp;
  out->submitSec = j->submit;
  out->admitSec = j->admit;
  out->doneSec = j->finish;
}

}  // namespace

extern "C" {

int ct_server_create(const ct_server_opts* opts, ct_server** out) {
  if (!out) return set_error(CT_EINVAL, "ct_server_create: null out");
  *out = nullptr;
  ct_server_opts o{};
  if (opts) o = *opts;
  if (!o.hbmBudget) o.hbmBudget = uint64_t(4) << 30;
  if (!o.maxActive) o.maxActive = 1024;
  if (o.policy != CT_SERVE_FIFO && o.policy != CT_SERVE_SRPT && o.policy != CT_SERVE_SCENE)
    return set_error(CT_EINVAL, "ct_server_create: unknown policy");
  if (o.precision != CT_PRECISION_F64 && o.precision != CT_PRECISION_F32)
    return set_error(CT_EINVAL, "ct_server_create: unknown precision");
  ct_server* S = new ct_server();
  S->opts = o;
  hipError_t e = hipGetDevice(&S->device);
  if (e != hipSuccess) {
    delete S;
    return check_hip(e, "hipGetDevice");
  }
  int rc = ct_tracer_create(&S->T);
  if (!rc && o.maxPaths) rc = ct_tracer_set_capacity(S->T, o.maxPaths);
  if (!rc && o.graph) rc = ct_tracer_set_graph(S->T, 1);
  if (!rc) rc = ct_tracer_set_precision(S->T, o.precision);
  // the round stream, then the tracer's three side streams, so that a round's
  // four streams take four hardware queues; the client read stream comes last
  // and shares one (measured on a 431-request stream over 128 scenes: 1018 ->
  // 1245 spp*Mpix/s once no side stream shares the round stream's queue)
  if (!rc) {
    e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = check_hip(e, "hipStreamCreateWithFlags");
  }
  if (!rc) rc = tracer_streams(S->T);
  if (!rc) {
    e = hipStreamCreateWithFlags(&S->readStream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = check_hip(e, "hipStreamCreateWithFlags");
  }
  if (rc) {
    if (S->T) ct_tracer_destroy(S->T);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return rc;
  }
  S->t0 = Clock::now();
  S->worker = std::thread(serve_loop, S);
  *out = S;
  return CT_OK;
}

int ct_server_destroy(ct_server* S) {
  if (!S) return CT_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cvWork.notify_all();
  if (S->worker.joinable()) S->worker.join();
  free_released(S);
  for (auto& kv : S->jobs) free_job(S, kv.second);
  S->jobs.clear();
  int rc = ct_tracer_destroy(S->T);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  if (S->readStream) (void)hipStreamDestroy(S->readStream);
  delete S;
  return rc;
}

int ct_job_submit(ct_server* S, const ct_job* job, uint64_t* id) {
  if (!S || !job || !id) return set_error(CT_EINVAL, "ct_job_submit: null argument");
  const ct_job& d = *job;
  if (!d.scene) return set_error(CT_EINVAL, "ct_job_submit: null scene");
  size_t npix = (size_t)d.width * d.height;
  if (npix == 0 || d.spp == 0)
    return set_error(CT_EINVAL, "ct_job_submit: width, height and spp must be positive");
  if (d.maxBounces < 0 || d.maxBounces > 254)
    return set_error(CT_ELIMIT, "ct_job_submit: maxBounces must be in [0, 254]");
  if (npix > tracer_capacity(S->T))
    return set_error(CT_ELIMIT, "ct_job_submit: one pass exceeds the round's path capacity");
  size_t bytes = npix * 3 * sizeof(float) * (S->opts.precision == CT_PRECISION_F32 ? 2 : 1);
  if (bytes > S->opts.hbmBudget)
    return set_error(CT_ELIMIT, "ct_job_submit: the progressive buffer exceeds the HBM budget");
  SJob* j = new SJob();
  j->desc = d;
  j->desc.d_out = nullptr;
  j->bytes = bytes;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    j->id = S->nextId++;
    j->submit = S->now();
    S->jobs[j->id] = j;
    S->queue.push_back(j);
    S->st.submitted++;
    *id = j->id;
  }
  S->cvWork.notify_all();
  return CT_OK;
}

int ct_job_poll(ct_server* S, uint64_t id, ct_job_status* status) {
  if (!S || !status) return set_error(CT_EINVAL, "ct_job_poll: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  SJob* j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_poll: unknown job");
  fill_status(S, j, status);
  if (j->state == CT_JOB_FAILED) set_error(CT_EDEVICE, j->error);
  return CT_OK;
}

int ct_job_wait(ct_server* S, uint64_t id, double timeoutSec, ct_job_status* status) {
  if (!S) return set_error(CT_EINVAL, "ct_job_wait: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  SJob* j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_wait: unknown job");
  auto ended = [&] {
    SJob* x = find(S, id);
    return !x || x->state == CT_JOB_DONE || x->state == CT_JOB_FAILED;
  };
  bool ok = true;
  if (timeoutSec < 0)
    S->cvDone.wait(lk, ended);
  else
    ok = S->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSec), ended);
  j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_wait: job released while waiting");
  if (status) fill_status(S, j, status);
  if (!ok) return set_error(CT_ETIMEOUT, "ct_job_wait: timed out");
  if (j->state == CT_JOB_FAILED) return set_error(CT_EDEVICE, j->error);
  return CT_OK;
}

int ct_job_read(ct_server* S, uint64_t id, float* dst, int32_t dstOnDevice) {
  if (!S || !dst) return set_error(CT_EINVAL, "ct_job_read: null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  SJob* j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_read: unknown job");
  // never read a buffer while a round writes it or before it is cleared
  S->cvDone.wait(lk, [&] {
    SJob* x = find(S, id);
    return !x || (x->inFlight == 0 && !x->pending);
  });
  j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_read: job released while reading");
  if (!j->accum) return set_error(CT_EINVAL, "ct_job_read: the job has no buffer yet");
  CT_HIP(hipMemcpyAsync(dst, j->accum, (size_t)j->desc.width * j->desc.height * 3 * sizeof(float),
                        dstOnDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        S->readStream));
  CT_HIP(hipStreamSynchronize(S->readStream));
  return CT_OK;
}

int ct_job_read_rgba(ct_server* S, uint64_t id, double exposureEv, int32_t mode, uint8_t* dst,
                     int32_t dstOnDevice) {
  if (!S || !dst) return set_error(CT_EINVAL, "ct_job_read_rgba: null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  SJob* j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_read_rgba: unknown job");
  S->cvDone.wait(lk, [&] {
    SJob* x = find(S, id);
    return !x || (x->inFlight == 0 && !x->pending);
  });
  j = find(S, id);
  if (!j) return set_error(CT_EINVAL, "ct_job_read_rgba: job released while reading");
  if (!j->accum) return set_error(CT_EINVAL, "ct_job_read_rgba: the job has no buffer yet");
  size_t npix = (size_t)j->desc.width * j->desc.height;
  // a finished buffer holds the average, a running one the sum of its passes
  double samples = j->state == CT_JOB_DONE ? 1.0 : (double)j->done;
  uint8_t* d = dst;
  if (!dstOnDevice) CT_HIP(hipMallocAsync((void**)&d, npix * 4, S->readStream));
  int rc = launch_tonemap(j->accum, npix, samples, exposureEv, mode, d, S->readStream);
  hipError_t e = hipSuccess;
  if (!rc && !dstOnDevice)
    e = hipMemcpyAsync(dst, d, npix * 4, hipMemcpyDeviceToHost, S->readStream);
  if (!dstOnDevice) {
    hipError_t e2 = hipFreeAsync(d, S->readStream);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(S->readStream);
  if (rc) return rc;
  return check_hip(e, "ct_job_read_rgba");
}

int ct_job_release(ct_server* S, uint64_t id) {
  if (!S) return set_error(CT_EINVAL, "ct_job_release: null server");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    SJob* j = find(S, id);
    if (!j) return set_error(CT_EINVAL, "ct_job_release: unknown job");
    if (j->state == CT_JOB_QUEUED) {
      S->queue.erase(std::remove(S->queue.begin(), S->queue.end(), j), S->queue.end());
      S->jobs.erase(id);
      delete j;  // no buffer yet
    } else if (j->inFlight || j->pending) {
      j->release = true;  // the worker frees it after the round
    } else {
      if (j->state == CT_JOB_RUNNING) drop_active(S, j);
      S->jobs.erase(id);
      S->released.push_back(j);
    }
  }
  S->cvWork.notify_all();
  S->cvDone.notify_all();
  return CT_OK;
}

int ct_server_stats_get(ct_server* S, ct_server_stats* out) {
  if (!S || !out) return set_error(CT_EINVAL, "ct_server_stats_get: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  out->queued = S->queue.size();
  out->active = S->active.size();
  out->nowSec = S->now();
  return CT_OK;
}

}  // extern "C"
