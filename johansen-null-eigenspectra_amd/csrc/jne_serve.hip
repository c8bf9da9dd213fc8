// jne_serve.hip -- the critical-value query server.
//
// A query (model, dim, steps, runs, seed, precision) asks for the 0.90, 0.95
// and 0.99 quantiles of the trace (sum of eigenvalues) and max-eigenvalue
// statistics over `runs` walks, with the statistic and quantile definitions
// of src/simulation_analyzers.rs:4-40.  Walk w of a query is Philox walk
// seed w+1 on key word `seed`, so a query's walks depend only on the query:
// its result is the same alone and in any mix.
//
// Continuous batching: a scheduler thread keeps the admitted queries and
// hands their walks out in chunks to `lanes` launch lanes (each a HIP stream
// with its own engine context, default priority), the next chunk of a free
// lane going to the query with the least work left (walks x steps x dim:
// shortest remaining work first, which minimises the mean latency; queries
// with equal work go in arrival order).
// A query that arrives while others run is admitted as soon as its HBM
// footprint (2 doubles per walk for its per-walk statistics) fits the
// budget, and its chunks interleave with the running queries' chunks on the
// device.  When a query's last chunk is done a finisher thread selects the
// three quantiles of each of its two statistics arrays on the device (bucket
// selection, jne_quantile.hip: no full sort) and copies back six values.
#include <hip/hip_runtime.h>
#include <math.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jne.h"
#include "jne_host.h"
#include "jne_internal.h"

namespace {

using jne_host::Clock;

// consecutive walk seeds first, first+1, ... of one chunk
__global__ void iota_seeds(uint32_t* __restrict__ out, uint32_t first, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = first + i;
}

// per-walk trace and max eigenvalue (simulation_analyzers.rs:26-39)
__global__ void walk_statistics(const double* __restrict__ eigs, int nf, uint32_t n, double* __restrict__ trace,
                                double* __restrict__ maxeig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* e = eigs + (size_t)i * nf;
  double s = 0.0, m = -1.7976931348623157e308;
  for (int j = 0; j < nf; ++j) {
    s += e[j];
    m = fmax(m, e[j]);
  }
  trace[i] = s;
  maxeig[i] = m;
}

// A chunk is at most 65,536 walks and at most 2^31 normals (walks x steps x
// dim; about 3 ms of the whole GPU at dim 12): long high-dim queries come in
// smaller chunks, so a query that arrives behind them waits less for a lane.
constexpr uint32_t kChunkWalks = 1u << 16;
constexpr uint64_t kChunkNormals = 1ull << 31;
constexpr uint32_t kMinChunkWalks = 4096;

uint32_t chunk_walks(uint32_t steps, int dim) {
  const uint64_t per_walk = (uint64_t)steps * (uint64_t)dim;
  const uint64_t n = kChunkNormals / (per_walk ? per_walk : 1);
  if (n >= kChunkWalks) return kChunkWalks;
  return n <= kMinChunkWalks ? kMinChunkWalks : (uint32_t)(n & ~uint64_t(255));
}

struct Request {
  uint64_t id;
  int model, dim, precision;
  uint32_t steps, seed;
  uint64_t runs;
  int nf;
  Clock::time_point submitted, finished;
  // device statistics: trace[runs], maxeig[runs] at arena offset
  size_t arena_off = 0;
  size_t bytes = 0;
  bool has_arena = false;         // arena_off holds an allocation
  int inflight = 0;               // chunks issued and not yet retired
  uint64_t issued = 0, done = 0;  // walks
  int status = 0;                 // JNE_OK or an error code
  std::string error;
  bool complete = false;
  double q[6] = {0, 0, 0, 0, 0, 0};
};

struct Lane {
  jne_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  uint32_t* d_seeds = nullptr;
  double* d_eigs = nullptr;
  hipEvent_t ev = nullptr;
  std::shared_ptr<Request> req;  // chunk in flight
  uint32_t walks = 0;
};

// first-fit allocator over one device buffer
class Arena {
 public:
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~size_t(255);
    size_t at = 0;
    for (auto it = used_.begin(); it != used_.end(); ++it) {
      if (it->first - at >= bytes) break;
      at = it->first + it->second;
    }
    if (at + bytes > cap_) return false;
    used_[at] = bytes;
    *off = at;
    return true;
  }
  void free(size_t off) { used_.erase(off); }
  size_t cap_ = 0;

 private:
  std::map<size_t, size_t> used_;
};

}  // namespace

struct jne_server {
  int device = 0;
  char* arena_base = nullptr;
  Arena arena;
  std::vector<Lane> lanes;
  std::mutex mu;
  std::condition_variable cv;       // scheduler wake-up
  std::condition_variable done_cv;  // completions
  std::deque<std::shared_ptr<Request>> pending;
  std::vector<std::shared_ptr<Request>> active;  // admitted, in arrival order
  std::map<uint64_t, std::shared_ptr<Request>> all;
  std::deque<std::shared_ptr<Request>> to_finish;
  uint64_t next_id = 1;
  uint64_t epoch = 0;  // bumped by submissions and freed memory: wakes an idle scheduler
  bool stop = false;
  std::thread scheduler, finisher;
  hipStream_t fin_st = nullptr;  // the finisher's stream, created before the lanes' streams
  // counters
  uint64_t walks_done = 0;
  double busy_s = 0.0;
};

namespace {

// A failed query keeps its statistics memory until none of its chunks is in
// flight (they still write there); then the memory goes back to the arena.
void release_if_drained(jne_server* s, const std::shared_ptr<Request>& r) {
  if (r->status != 0 && r->has_arena && r->inflight == 0) {
    s->arena.free(r->arena_off);
    r->has_arena = false;
    ++s->epoch;
  }
}

void fail_request(jne_server* s, const std::shared_ptr<Request>& r, int code, const std::string& msg) {
  if (r->status == 0) {
    r->status = code;
    r->error = msg;
    r->complete = true;
    r->finished = Clock::now();
    s->done_cv.notify_all();
  }
  release_if_drained(s, r);
}

// Issues the next chunk of request r on lane L (called with s->mu held).
int issue(jne_server* s, Lane& L, const std::shared_ptr<Request>& r) {
  const uint32_t first = (uint32_t)(r->issued + 1);
  const uint32_t n = (uint32_t)std::min<uint64_t>(chunk_walks(r->steps, r->dim), r->runs - r->issued);
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(iota_seeds, dim3((n + 255) / 256), dim3(256), 0, L.st, L.d_seeds, first, n);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return JNE_ERR_HIP;
  int rc = jne_eigenvalues_by_seed_device(L.ctx, r->model, r->dim, r->steps, r->seed, L.d_seeds, n, r->precision,
                                          L.d_eigs, L.st);
  if (rc) return rc;
  double* trace = reinterpret_cast<double*>(s->arena_base + r->arena_off) + r->issued;
  double* maxeig = reinterpret_cast<double*>(s->arena_base + r->arena_off) + r->runs + r->issued;
  hipLaunchKernelGGL(walk_statistics, dim3((n + 255) / 256), dim3(256), 0, L.st, L.d_eigs, r->nf, n, trace, maxeig);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipEventRecord(L.ev, L.st)) != hipSuccess) {
    // walk_statistics may be enqueued and write into the query's statistics:
    // drain the lane before the caller fails the query and frees that memory
    (void)hipStreamSynchronize(L.st);
    return JNE_ERR_HIP;
  }
  r->issued += n;
  ++r->inflight;
  L.req = r;
  L.walks = n;
  return 0;
}

void scheduler_loop(jne_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto t_busy = Clock::now();
  bool was_busy = false;
  while (!s->stop) {
    const uint64_t seen = s->epoch;
    // admit pending queries whose statistics fit the budget (arrival order)
    while (!s->pending.empty()) {
      auto r = s->pending.front();
      size_t off;
      if (r->bytes > s->arena.cap_) {  // can never fit
        s->pending.pop_front();
        fail_request(s, r, JNE_ERR_NOMEM, "query footprint exceeds the server's HBM budget");
        continue;
      }
      if (!s->arena.alloc(r->bytes, &off)) break;  // wait for memory, keep arrival order
      r->arena_off = off;
      r->has_arena = true;
      s->pending.pop_front();
      s->active.push_back(r);
    }
    // retire finished chunks
    bool any_busy = false;
    for (Lane& L : s->lanes) {
      if (!L.req) continue;
      const hipError_t q = hipEventQuery(L.ev);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      auto r = L.req;
      L.req.reset();
      --r->inflight;
      if (q != hipSuccess) {
        fail_request(s, r, JNE_ERR_HIP, std::string("chunk failed: ") + hipGetErrorString(q));
        continue;
      }
      r->done += L.walks;
      s->walks_done += L.walks;
      if (r->done == r->runs && r->status == 0) {
        s->to_finish.push_back(r);
        s->done_cv.notify_all();
      }
      release_if_drained(s, r);  // a query failed by another lane's chunk
    }
    // hand out chunks: each free lane takes the next chunk of the admitted
    // query with the least work left
    for (Lane& L : s->lanes) {
      if (L.req) continue;
      std::shared_ptr<Request> pick;
      uint64_t least = UINT64_MAX;
      for (const auto& r : s->active) {
        if (r->status != 0 || r->issued >= r->runs) continue;
        const uint64_t left = (r->runs - r->issued) * (uint64_t)r->steps * (uint64_t)r->dim;
        if (left < least) {
          least = left;
          pick = r;
        }
      }
      if (!pick) break;
      const int rc = issue(s, L, pick);
      if (rc) {
        fail_request(s, pick, rc, L.ctx ? jne_last_error(L.ctx) : "launch failed");
        continue;
      }
      any_busy = true;
    }
    // drop queries that are fully issued and done or failed from the ring
    s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                   [](const std::shared_ptr<Request>& r) {
                                     return r->status != 0 || (r->issued == r->runs && r->done == r->runs);
                                   }),
                    s->active.end());
    const auto now = Clock::now();
    if (was_busy) s->busy_s += std::chrono::duration<double>(now - t_busy).count();
    t_busy = now;
    was_busy = any_busy;
    // both waits release the lock; an idle scheduler sleeps until a query
    // arrives or memory is freed (a pending query that does not fit yet must
    // not keep the lock from the finisher that frees memory)
    if (any_busy)
      s->cv.wait_for(lk, std::chrono::microseconds(20));
    else
      s->cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return s->stop || s->epoch != seen; });
  }
}

void finisher_loop(jne_server* s) {
  hipStream_t st = s->fin_st;
  (void)hipSetDevice(s->device);
  {
    jne::QuantileScratch sc;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
      s->done_cv.wait(lk, [&] { return s->stop || !s->to_finish.empty(); });
      if (s->to_finish.empty() && s->stop) break;
      auto r = s->to_finish.front();
      s->to_finish.pop_front();
      double* const dev = reinterpret_cast<double*>(s->arena_base + r->arena_off);
      lk.unlock();
      // the 0.90/0.95/0.99 quantiles of trace and max-eig, selected on the device
      static const double ps[3] = {0.90, 0.95, 0.99};
      double q6[6];
      const double* const arrays[2] = {dev, dev + r->runs};
      const int code = jne::quantile_select(sc, arrays, 2, r->runs, ps, 3, q6, st, nullptr);
      std::string msg;
      if (code == JNE_ERR_NUMERIC)
        msg = "a walk's S11 is not positive definite (NaN eigenvalues)";
      else if (code)
        msg = "quantile selection failed (error " + std::to_string(code) + ")";
      lk.lock();
      s->arena.free(r->arena_off);
      r->has_arena = false;
      ++s->epoch;
      if (code) {
        fail_request(s, r, code, msg);
      } else {
        memcpy(r->q, q6, sizeof q6);
        r->complete = true;
        r->finished = Clock::now();
        s->done_cv.notify_all();
      }
      s->cv.notify_all();  // memory freed: the scheduler may admit more
    }
  }
}

}  // namespace

extern "C" {

int jne_server_create(int device, uint64_t hbm_budget_bytes, int lanes, jne_server** out) {
  if (!out || lanes < 1 || lanes > 8 || hbm_budget_bytes < (1u << 20)) {
    jne_host::set_error("jne_server_create: invalid argument (lanes 1..8, budget >= 1 MiB)");
    return JNE_ERR_INVALID_ARG;
  }
  *out = nullptr;
  jne_server* s = new jne_server();
  s->device = device;
  hipError_t e = hipSetDevice(device);
  // The finisher's stream: HIP maps ordinary streams onto a few hardware
  // queues (4 by default), each running its work in order, so a finisher
  // stream that shares a queue with a lane waits behind that lane's chunks
  // at every selection.  A stream with a CU mask (here: every CU) gets a
  // queue of its own; an ordinary stream, created before the lanes', is the
  // fallback.
  if (e == hipSuccess) {
    std::vector<uint32_t> all_cus(16, 0xFFFFFFFFu);  // up to 512 CUs; extra bits are ignored
    if (hipExtStreamCreateWithCUMask(&s->fin_st, (uint32_t)all_cus.size(), all_cus.data()) != hipSuccess) {
      (void)hipGetLastError();
      s->fin_st = nullptr;
      e = hipStreamCreateWithFlags(&s->fin_st, hipStreamNonBlocking);
    }
  }
  if (e == hipSuccess) e = hipMalloc((void**)&s->arena_base, hbm_budget_bytes);
  if (e != hipSuccess) {
    jne_host::set_error(std::string("jne_server_create: ") + hipGetErrorString(e));
    if (s->fin_st) (void)hipStreamDestroy(s->fin_st);
    delete s;
    return JNE_ERR_NOMEM;
  }
  s->arena.cap_ = hbm_budget_bytes;
  s->lanes.resize(lanes);
  for (Lane& L : s->lanes) {
    int rc = jne_create(device, &L.ctx);
    if (rc == JNE_OK) {
      e = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipMalloc((void**)&L.d_seeds, kChunkWalks * sizeof(uint32_t));
      if (e == hipSuccess) e = hipMalloc((void**)&L.d_eigs, (size_t)kChunkWalks * (JNE_MAX_DIM + 1) * sizeof(double));
      if (e == hipSuccess) e = hipEventCreateWithFlags(&L.ev, hipEventDisableTiming);
      if (e != hipSuccess) rc = JNE_ERR_HIP;
    }
    if (rc != JNE_OK) {
      jne_host::set_error(std::string("jne_server_create: lane setup failed: ") +
                          (L.ctx ? hipGetErrorString(e) : jne_last_error(nullptr)));
      jne_server_destroy(s);
      return rc;
    }
  }
  s->scheduler = std::thread(scheduler_loop, s);
  s->finisher = std::thread(finisher_loop, s);
  *out = s;
  return JNE_OK;
}

void jne_server_destroy(jne_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
    s->done_cv.notify_all();
  }
  if (s->scheduler.joinable()) s->scheduler.join();
  if (s->finisher.joinable()) s->finisher.join();
  (void)hipSetDevice(s->device);
  for (Lane& L : s->lanes) {
    if (L.st) (void)hipStreamSynchronize(L.st);
    if (L.ev) (void)hipEventDestroy(L.ev);
    if (L.d_eigs) (void)hipFree(L.d_eigs);
    if (L.d_seeds) (void)hipFree(L.d_seeds);
    if (L.st) (void)hipStreamDestroy(L.st);
    if (L.ctx) jne_destroy(L.ctx);
  }
  if (s->arena_base) (void)hipFree(s->arena_base);
  if (s->fin_st) (void)hipStreamDestroy(s->fin_st);
  delete s;
}

int jne_server_submit(jne_server* s, int model, int dim, uint32_t steps, uint64_t runs, uint32_t seed, int precision,
                      uint64_t* id) {
  if (!s || !id) {
    jne_host::set_error("jne_server_submit: NULL argument");
    return JNE_ERR_INVALID_ARG;
  }
  const int nf = jne_num_eigenvalues(model, dim);
  if (nf < 0 || dim > JNE_MAX_DIM || steps < 2 || steps > (1u << 30) || runs < 1 || runs > 0x7FFFFFFFull) {
    jne_host::set_error("jne_server_submit: invalid model, dim, steps or runs (runs 1..2^31-1)");
    return JNE_ERR_INVALID_ARG;
  }
  if (precision != JNE_PREC_FP64 && precision != JNE_PREC_MFMA_FP16X2 && precision != JNE_PREC_MFMA_BF16X2 &&
      precision != JNE_PREC_MFMA_BF16X3 && precision != JNE_PREC_MFMA_FP32 && precision != JNE_PREC_AUTO) {
    jne_host::set_error("jne_server_submit: unsupported precision");
    return JNE_ERR_INVALID_ARG;
  }
  if (dim > JNE_MAX_MIXED_DIM && precision != JNE_PREC_FP64 && precision != JNE_PREC_AUTO) {
    jne_host::set_error("jne_server_submit: the reduced precisions cover dims 1..12; above, use fp64 or auto");
    return JNE_ERR_UNSUPPORTED;
  }
  auto r = std::make_shared<Request>();
  r->model = model;
  r->dim = dim;
  r->steps = steps;
  r->runs = runs;
  r->seed = seed;
  r->precision = precision;
  r->nf = nf;
  r->bytes = (size_t)runs * 2 * sizeof(double);
  r->submitted = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  s->all[r->id] = r;
  s->pending.push_back(r);
  ++s->epoch;
  s->cv.notify_all();
  *id = r->id;
  return JNE_OK;
}

int jne_server_wait(jne_server* s, uint64_t id, int timeout_ms, double* quantiles, double* latency_ms) {
  if (!s) {
    jne_host::set_error("jne_server_wait: NULL server");
    return JNE_ERR_INVALID_ARG;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) {
    jne_host::set_error("jne_server_wait: unknown request id");
    return JNE_ERR_INVALID_ARG;
  }
  auto r = it->second;
  auto ready = [&] { return r->complete; };
  if (timeout_ms < 0)
    s->done_cv.wait(lk, ready);
  else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return JNE_ERR_TIMEOUT;
  s->all.erase(id);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
  if (r->status) {
    jne_host::set_error(r->error);
    return r->status;
  }
  if (quantiles) memcpy(quantiles, r->q, sizeof r->q);
  return JNE_OK;
}

int jne_server_counters(jne_server* s, uint64_t* walks, double* busy_seconds) {
  if (!s) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (walks) *walks = s->walks_done;
  if (busy_seconds) *busy_seconds = s->busy_s;
  return JNE_OK;
}

}  // extern "C"
