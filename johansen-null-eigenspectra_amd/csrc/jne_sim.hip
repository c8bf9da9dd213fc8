// jne_sim.hip -- the simulation driver: remaining seeds of a (model, dim,
// steps, runs) job computed on the device in batches and streamed to the
// .dat file, with resume (src/data_storage/parallel_compute.rs:150-232,
// src/data_storage/thread_manager.rs:25-92).
//
// Pipeline: batch b's seeds go up, its walks run, its eigenvalues come back
// into pinned host buffer b%2 on one stream; a host thread waits on the
// batch's event, checks it and encodes/writes it while the device computes
// batch b+1.
#include <errno.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jne.h"
#include "jne_host.h"

namespace {

using namespace jne_host;

constexpr size_t kBatchWalks = size_t(1) << 20;
const char* const kRule = "===============================\n";

constexpr int kMaxModels = 5;

struct Slot {
  uint32_t* h_seeds = nullptr;            // pinned
  double* h_out[kMaxModels] = {};         // pinned, per model of the group
  uint32_t* d_seeds = nullptr;
  double* d_out[kMaxModels] = {};
  hipEvent_t done = nullptr;
  size_t n = 0;
  bool busy = false;  // owned by the writer thread until written
};

struct Job {
  int slot;
  size_t n;
};

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return JNE_ERR_HIP;
}

// Detected ... completed ... Seeds range (parallel_compute.rs:79-103)
void report_progress(const char* path, uint64_t runs, uint64_t completed) {
  if (completed == 0) return;
  DatContents c;
  uint32_t lo = 0, hi = 0;
  if (read_dat(path, false, &c) == 0 && !c.seeds.empty()) {
    lo = *std::min_element(c.seeds.begin(), c.seeds.end());
    hi = *std::max_element(c.seeds.begin(), c.seeds.end());
  }
  printf("Detected %s completed out of %s calculations, Seeds range: %s-%s\n", commas(completed).c_str(),
         commas(runs).c_str(), commas(lo).c_str(), commas(hi).c_str());
}

void validate_output(const char* path, uint64_t expected) {  // parallel_compute.rs:44-66
  DatContents c;
  if (read_dat(path, false, &c) == 0) {
    if (c.seeds.size() == expected)
      printf("SUCCESS: append file validation successful\n");
    else
      printf("ERROR: append file validation failed: data length mismatch (expected: %llu, actual: %llu)\n",
             (unsigned long long)expected, (unsigned long long)c.seeds.size());
  } else {
    printf("ERROR: failed to read append file: %s\n", last_error().c_str());
  }
}

// The walks of `seeds` for the n_models models of a group (the same remaining
// seeds in every model's file), streamed to paths[k]: batches of seeds go up,
// one moment pass and each model's eigen solve run on the device
// (jne_eigenvalues_by_seed_models_device), the eigenvalues come back into
// pinned host buffers of slot b%2, and a host thread checks and appends each
// model's records while the device computes the next batch.
int run_group(jne_ctx* ctx, const int* models, const char* const* paths, int n_models, int dim, uint32_t steps,
              uint64_t runs, int precision, int quiet, const std::vector<uint32_t>& seeds) {
  const size_t remaining = seeds.size();
  int nf[kMaxModels] = {};
  for (int k = 0; k < n_models; ++k) {
    nf[k] = jne_num_eigenvalues(models[k], dim);
    if (nf[k] < 0) {
      set_error("invalid model or dim");
      return JNE_ERR_INVALID_ARG;
    }
  }
  int rc = 0;
  jne_dat_writer* w[kMaxModels] = {};
  for (int k = 0; k < n_models && rc == 0; ++k) {
    if ((rc = jne_dat_writer_open(paths[k], models[k], dim, steps, runs, quiet, &w[k])) == 0)
      jne_dat_writer_set_progress(w[k], runs);
  }
  if (rc) {
    for (int k = 0; k < n_models; ++k)
      if (w[k]) jne_dat_writer_abandon(w[k]);
    return rc;
  }

  const size_t batch = std::min<size_t>(kBatchWalks, remaining);
  Slot slot[2];
  hipStream_t st = nullptr;
  hipError_t he = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (int s = 0; s < 2 && he == hipSuccess; ++s) {
    if ((he = hipHostMalloc((void**)&slot[s].h_seeds, batch * sizeof(uint32_t), hipHostMallocDefault)) != hipSuccess) break;
    if ((he = hipMalloc((void**)&slot[s].d_seeds, batch * sizeof(uint32_t))) != hipSuccess) break;
    for (int k = 0; k < n_models && he == hipSuccess; ++k) {
      if ((he = hipHostMalloc((void**)&slot[s].h_out[k], batch * nf[k] * sizeof(double), hipHostMallocDefault)) !=
          hipSuccess)
        break;
      he = hipMalloc((void**)&slot[s].d_out[k], batch * nf[k] * sizeof(double));
    }
    if (he == hipSuccess) he = hipEventCreateWithFlags(&slot[s].done, hipEventDisableTiming);
  }

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  bool closing = false;
  int werr = 0;
  std::string werr_msg;
  std::thread writer([&] {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !jobs.empty() || closing; });
        if (jobs.empty()) return;
        j = jobs.front();
        jobs.pop_front();
      }
      Slot& s = slot[j.slot];
      int code = 0;
      std::string msg;
      hipError_t e = hipEventSynchronize(s.done);
      if (e != hipSuccess) {
        code = JNE_ERR_HIP;
        msg = std::string("hipEventSynchronize: ") + hipGetErrorString(e);
      }
      for (int k = 0; k < n_models; ++k) {
        for (size_t i = 0; code == 0 && i < j.n * (size_t)nf[k]; ++i)
          if (isnan(s.h_out[k][i])) {
            code = JNE_ERR_NUMERIC;
            msg = "seed " + std::to_string(s.h_seeds[i / nf[k]]) + ": S11 is not positive definite";
          }
        if (code == 0 && jne_dat_writer_append(w[k], s.h_seeds, s.h_out[k], j.n, nf[k]) != 0) {
          code = JNE_ERR_IO;
          msg = last_error();
        }
      }
      fflush(stdout);
      std::lock_guard<std::mutex> lk(mu);
      if (code && !werr) {
        werr = code;
        werr_msg = msg;
      }
      s.busy = false;
      cv.notify_all();
    }
  });

  int perr = he == hipSuccess ? 0 : hip_fail(he, "allocating the batch buffers");
  std::string perr_msg = perr ? last_error() : "";
  for (size_t off = 0, b = 0; perr == 0 && off < remaining; off += batch, ++b) {
    const int si = (int)(b & 1);
    Slot& s = slot[si];
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !s.busy || werr; });
      if (werr) break;
    }
    s.n = std::min<size_t>(batch, remaining - off);
    memcpy(s.h_seeds, seeds.data() + off, s.n * sizeof(uint32_t));
    if ((he = hipMemcpyAsync(s.d_seeds, s.h_seeds, s.n * sizeof(uint32_t), hipMemcpyHostToDevice, st)) != hipSuccess) {
      perr = hip_fail(he, "hipMemcpyAsync");
    } else if ((rc = jne_eigenvalues_by_seed_models_device(ctx, models, n_models, dim, steps, 0, s.d_seeds, s.n,
                                                           precision, s.d_out, st)) != 0) {
      perr = rc;
      set_error(jne_last_error(ctx));
    } else {
      for (int k = 0; k < n_models && he == hipSuccess; ++k)
        he = hipMemcpyAsync(s.h_out[k], s.d_out[k], s.n * nf[k] * sizeof(double), hipMemcpyDeviceToHost, st);
      if (he == hipSuccess) he = hipEventRecord(s.done, st);
      if (he != hipSuccess) perr = hip_fail(he, "hipMemcpyAsync/hipEventRecord");
    }
    if (perr) {
      perr_msg = last_error();
      break;
    }
    std::lock_guard<std::mutex> lk(mu);
    s.busy = true;
    jobs.push_back({si, s.n});
    cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    closing = true;
    cv.notify_all();
  }
  writer.join();
  if (st) (void)hipStreamSynchronize(st);
  for (Slot& s : slot) {
    if (s.done) (void)hipEventDestroy(s.done);
    for (int k = 0; k < n_models; ++k) {
      if (s.d_out[k]) (void)hipFree(s.d_out[k]);
      if (s.h_out[k]) (void)hipHostFree(s.h_out[k]);
    }
    if (s.d_seeds) (void)hipFree(s.d_seeds);
    if (s.h_seeds) (void)hipHostFree(s.h_seeds);
  }
  if (st) (void)hipStreamDestroy(st);

  if (perr || werr) {
    // keep what was written; the next run resumes from it
    for (int k = 0; k < n_models; ++k) jne_dat_writer_abandon(w[k]);
    set_error(perr ? perr_msg : werr_msg);
    return perr ? perr : werr;
  }
  for (int k = 0; k < n_models; ++k) {
    if ((rc = jne_dat_writer_finish(w[k]))) {
      for (int m = k + 1; m < n_models; ++m) jne_dat_writer_abandon(w[m]);
      return rc;
    }
  }
  return 0;
}

int run(jne_ctx* ctx, int model, int dim, uint32_t steps, uint64_t runs, int precision, const char* path, int quiet,
        int depth) {
  if (!quiet) printf("Using model: %s (supports resuming from checkpoint)\n", model_display(model).c_str());
  uint64_t remaining = 0, completed = 0;
  int rc = jne_dat_remaining_seeds(path, model, dim, steps, runs, nullptr, 0, &remaining, &completed);
  if (rc == JNE_ERR_FORMAT && depth == 0) {
    const std::string msg = last_error();
    if (!quiet) {
      printf("WARNING: Existing file has incompatible parameters:\n");
      printf("  %s\n", msg.c_str());
      printf("  The existing file will be removed and recreated with correct parameters.\n");
    }
    if (unlink(path) != 0 && !quiet) printf("WARNING: Failed to remove incompatible file: %s\n", strerror(errno));
    if (!quiet) printf("Starting fresh calculation with correct parameters...\n");
    return run(ctx, model, dim, steps, runs, precision, path, quiet, depth + 1);
  }
  if (rc) {
    set_error("Failed to check progress: " + last_error());
    return rc;
  }
  if (!quiet) report_progress(path, runs, completed);
  if (completed >= runs) {
    if (!quiet) printf("SUCCESS: calculation for this model already completed, skipping\n%s\n", kRule);
    return 0;
  }
  if (remaining == 0) {
    if (!quiet) printf("SUCCESS: calculation for this model already completed\n%s\n", kRule);
    return 0;
  }
  std::vector<uint32_t> seeds(remaining);
  uint64_t got = 0;
  if ((rc = jne_dat_remaining_seeds(path, model, dim, steps, runs, seeds.data(), remaining, &got, &completed)))
    return rc;
  if (!quiet) printf("Remaining %s calculations to complete\n", commas(remaining).c_str());
  fflush(stdout);

  if ((rc = run_group(ctx, &model, &path, 1, dim, steps, runs, precision, quiet, seeds))) return rc;
  if (!quiet) {
    printf("Saved to %s\n", path);
    validate_output(path, runs);
    printf("%s\n", kRule);
  }
  fflush(stdout);
  return 0;
}

// Quiet sweeps over several models of one dim: each model's remaining seeds,
// then the models with the same remaining seeds (all of them in a fresh
// sweep) run as one group, whose walks get one moment pass for every model.
// With messages (quiet = 0), or when two models would share a file, the
// models run one after another as the reference's loop does
// (src/main.rs:109-127), so the output is the reference's.
int run_models(jne_ctx* ctx, const int* models, int n_models, int dim, uint32_t steps, uint64_t runs, int precision,
               const char* const* paths, int quiet) {
  bool distinct = true;
  for (int k = 0; k < n_models; ++k)
    for (int m = k + 1; m < n_models; ++m)
      if (strcmp(paths[k], paths[m]) == 0) distinct = false;
  if (!quiet || !distinct) {
    for (int k = 0; k < n_models; ++k) {
      const int rc = run(ctx, models[k], dim, steps, runs, precision, paths[k], quiet, 0);
      if (rc) return rc;
    }
    return 0;
  }
  std::vector<std::vector<uint32_t>> rem(n_models);
  for (int k = 0; k < n_models; ++k) {
    uint64_t remaining = 0, completed = 0;
    int rc = jne_dat_remaining_seeds(paths[k], models[k], dim, steps, runs, nullptr, 0, &remaining, &completed);
    if (rc == JNE_ERR_FORMAT) {  // incompatible parameters: replaced, as in run()
      (void)unlink(paths[k]);
      rc = jne_dat_remaining_seeds(paths[k], models[k], dim, steps, runs, nullptr, 0, &remaining, &completed);
    }
    if (rc) {
      set_error("Failed to check progress: " + last_error());
      return rc;
    }
    if (completed >= runs || remaining == 0) continue;
    rem[k].resize(remaining);
    uint64_t got = 0;
    if ((rc = jne_dat_remaining_seeds(paths[k], models[k], dim, steps, runs, rem[k].data(), remaining, &got,
                                      &completed)))
      return rc;
    rem[k].resize(got);
  }
  std::vector<bool> taken(n_models, false);
  for (int k = 0; k < n_models; ++k) {
    if (taken[k] || rem[k].empty()) continue;
    int gm[kMaxModels];
    const char* gp[kMaxModels];
    int g = 0;
    for (int m = k; m < n_models; ++m)
      if (!taken[m] && rem[m] == rem[k]) {
        taken[m] = true;
        gm[g] = models[m];
        gp[g] = paths[m];
        ++g;
      }
    const int rc = run_group(ctx, gm, gp, g, dim, steps, runs, precision, 1, rem[k]);
    if (rc) return rc;
  }
  return 0;
}

}  // namespace

extern "C" int jne_run_simulation_models(jne_ctx* ctx, const int* models, int n_models, int dim, uint32_t steps,
                                         uint64_t runs, int precision, const char* const* paths, int quiet) {
  if (!ctx || !models || !paths || n_models < 1 || n_models > kMaxModels) {
    set_error("jne_run_simulation_models: ctx/models/paths must not be NULL and n_models must be in 1..5");
    return JNE_ERR_INVALID_ARG;
  }
  for (int k = 0; k < n_models; ++k)
    if (!paths[k] || models[k] < 0 || models[k] > 4) {
      set_error("jne_run_simulation_models: paths[" + std::to_string(k) + "] is NULL or model " +
                std::to_string(models[k]) + " is not in 0..4");
      return JNE_ERR_INVALID_ARG;
    }
  if (runs == 0 || runs > 0xFFFFFFFFull) {
    set_error("jne_run_simulation_models: runs must be in 1..2^32-1");
    return JNE_ERR_INVALID_ARG;
  }
  return run_models(ctx, models, n_models, dim, steps, runs, precision, paths, quiet);
}

extern "C" int jne_run_simulation(jne_ctx* ctx, int model, int dim, uint32_t steps, uint64_t runs, int precision,
                                  const char* path, int quiet) {
  if (!ctx || !path) {
    set_error("jne_run_simulation: ctx/path must not be NULL");
    return JNE_ERR_INVALID_ARG;
  }
  if (runs == 0 || runs > 0xFFFFFFFFull) {
    set_error("jne_run_simulation: runs must be in 1..2^32-1");
    return JNE_ERR_INVALID_ARG;
  }
  return run(ctx, model, dim, steps, runs, precision, path, quiet, 0);
}
