// jne_capi.hip -- the C ABI (include/jne.h): contexts, argument checks,
// chunked launches of the walk kernels, host/device copies.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jne.h"
#include "jne_internal.h"
#include "jne_rng.h"

namespace jne {

// Walk w0 + blockIdx.y; its Philox calls q = blockIdx.x*256 + tid, stepping by
// the grid width, each writing normals 4q..4q+3 (16-byte stores when `vec`:
// every walk's row of T*d floats starts 16-byte aligned).
__global__ __launch_bounds__(256) void walk_normals_kernel(const uint32_t* __restrict__ seeds, uint32_t rng_stream,
                                                           uint32_t w0, uint32_t total, uint32_t calls, bool vec,
                                                           float* __restrict__ out) {
  const uint32_t walk = w0 + blockIdx.y;
  const uint32_t seed = seeds[walk];
  float* const o = out + (size_t)walk * total;
  for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < calls; q += gridDim.x * 256u) {
    float v[4];
    normals4(q, seed, rng_stream, v);
    if (vec) {
      *reinterpret_cast<float4*>(o + 4 * (size_t)q) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (4 * q + k < total) o[4 * (size_t)q + k] = v[k];
    }
  }
}

hipError_t launch_walk_normals(int d, const uint32_t* seeds, uint32_t rng_stream, uint32_t T, uint32_t n,
                               float* out, hipStream_t st) {
  // the caller keeps T * d < 2^32; a launch covers at most 65,535 walks (grid y)
  const uint32_t total = T * (uint32_t)d;
  const uint32_t calls = (total + 3) / 4;
#ifndef JNE_NORMALS_GX
#define JNE_NORMALS_GX 512u
#endif
  const uint32_t gx = (calls + 255) / 256 < JNE_NORMALS_GX ? (calls + 255) / 256 : JNE_NORMALS_GX;
  const bool vec = (total & 3u) == 0u && (reinterpret_cast<uintptr_t>(out) & 15u) == 0u;
  for (uint32_t w0 = 0; w0 < n; w0 += 65535u) {
    const uint32_t gy = n - w0 < 65535u ? n - w0 : 65535u;
    hipLaunchKernelGGL(walk_normals_kernel, dim3(gx, gy), dim3(256), 0, st, seeds, rng_stream, w0, total, calls, vec,
                       out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace jne

// Device scratch of jne_eigenvalues_by_seed_device_noalloc: the per-walk
// moments and the M matrices of the two-kernel eigen solve of one call, in a
// fixed array that each GPU the library runs on holds until the process ends
// (JNE_NOALLOC_SCRATCH_BYTES).  In host code `scratch` names a host-side
// stand-in, not the device address: the address of each GPU's copy is taken
// once per device with hipGetSymbolAddress (jne_create) and only that
// pointer is passed to kernels.
static __device__ __align__(256) double scratch[JNE_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // per device: the address of its `scratch`
}  // namespace

struct jne_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* noalloc_scratch = nullptr;  // this device's `scratch`
  std::mutex mu;
  std::string err;
  // device scratch: per-walk moments of one chunk
  double* d_acc = nullptr;
  size_t acc_bytes = 0;
  // device scratch: M = W^T W of one chunk for the two-kernel eigen solve
  double* d_m = nullptr;
  size_t m_bytes = 0;
  // the scratch is shared by every call: a call on another stream than the
  // previous one first waits for the previous call's last use of it
  hipEvent_t scratch_free = nullptr;
  hipStream_t scratch_stream = nullptr;
  // device staging for the host-pointer entry points
  void* d_in = nullptr;
  size_t in_bytes = 0;
  void* d_out = nullptr;
  size_t out_bytes = 0;
  // device scratch of jne_percentiles_device
  jne::QuantileScratch qs;
  // kernel timing (jne_profile_*)
  struct TimedLaunch {
    int kernel;
    hipEvent_t start, stop;
    uint64_t walks;
  };
  bool profiling = false;
  std::vector<TimedLaunch> timed;
  std::vector<hipEvent_t> event_pool;
};

namespace {

thread_local std::string g_create_err;
constexpr size_t kChunkWalks = size_t(1) << 18;

int set_err(jne_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  ctx->err = buf;
  return code;
}

#define JNE_HIP_CHECK(ctx, expr)                                                                   \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return set_err((ctx), JNE_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int ensure(jne_ctx* ctx, void** ptr, size_t* have, size_t need) {
  if (*have >= need) return JNE_OK;
  if (*ptr) {
    (void)hipFree(*ptr);
    *ptr = nullptr;
    *have = 0;
  }
  hipError_t e = hipMalloc(ptr, need);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return set_err(ctx, JNE_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", need, hipGetErrorString(e));
  }
  *have = need;
  return JNE_OK;
}

int check_args(jne_ctx* ctx, int model, int dim, uint32_t steps, int precision) {
  if (model < 0 || model > 4) return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid model %d (expected 0..4)", model);
  if (dim < 1 || dim > JNE_MAX_DIM)
    return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid dim %d (expected 1..%d)", dim, JNE_MAX_DIM);
  if (steps < 2 || steps > jne::kMaxSteps)
    return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid steps %u (expected 2..2^30)", steps);
  switch (precision) {
    case JNE_PREC_FP64:
    case JNE_PREC_AUTO:
      return JNE_OK;
    case JNE_PREC_MFMA_FP16X2:
    case JNE_PREC_MFMA_BF16X2:
    case JNE_PREC_MFMA_BF16X3:
    case JNE_PREC_MFMA_FP32:
    case JNE_PREC_FP32_KAHAN:
      if (dim > JNE_MAX_MIXED_DIM)
        return set_err(ctx, JNE_ERR_UNSUPPORTED, "precision %d covers dims 1..%d; dim %d runs the FP64 path",
                       precision, JNE_MAX_MIXED_DIM, dim);
      return JNE_OK;
    default:
      return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid precision %d", precision);
  }
}

hipEvent_t pool_event(jne_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void release_timed(jne_ctx* ctx) {
  for (auto& t : ctx->timed) {
    ctx->event_pool.push_back(t.start);
    ctx->event_pool.push_back(t.stop);
  }
  ctx->timed.clear();
}

// Records events around `launch` on st when profiling is enabled.
template <class F>
hipError_t timed_launch(jne_ctx* ctx, int kernel, uint64_t walks, hipStream_t st, F&& launch) {
  if (!ctx->profiling) return launch();
  hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, st);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e != hipSuccess) {
    ctx->event_pool.push_back(a);
    ctx->event_pool.push_back(b);
    return e;
  }
  ctx->timed.push_back({kernel, a, b, walks});
  return hipSuccess;
}

// Enqueue the walk path for n walks (device pointers) on stream st.
// JNE_PREC_AUTO: the MFMA pass where it measured faster than the FP64 pass
// (walks/s of the whole path, model 4, T = 100..10^4): dim 12 from 200 steps,
// dims 10-11 from 300, dim 9 from 500, dim 8 from 1,000 and dim 7 from 3,000;
// short walks carry the MFMA pass's per-walk setup.  At dims 1-6 the FP64 pass
// keeps each walk in one lane's registers and runs 1.2-1.45x the MFMA pass
// (which packs 12 / dim walks per wave at 1, 2, 3, 4 and 6).
bool auto_uses_mfma(int dim, uint32_t steps) {
  if (dim > JNE_MAX_MIXED_DIM) return false;
  if (dim == 12) return steps >= 200;
  if (dim >= 10) return steps >= 300;
  if (dim == 9) return steps >= 500;
  if (dim == 8) return steps >= 1000;
  if (dim == 7) return steps >= 3000;
  return false;
}

// Enqueues the walk kernels for n walks on st, in chunks of at most `chunk`
// walks over the scratch `acc` (chunk * moment_count(dim) doubles) and `mbuf`
// (chunk * eigen_m_count(dim) doubles): per chunk one moment pass, then the
// eigen solve of each of the n_models models into its output (the moments of
// a walk do not depend on the model).  timed: record the launches when
// profiling is enabled.
int launch_walks(jne_ctx* ctx, const int* models, int n_models, int dim, uint32_t steps, uint32_t rng_stream,
                 const uint32_t* d_seeds, const double* d_z, size_t n, int precision, double* const* d_outs,
                 hipStream_t st, double* acc, double* mbuf, size_t chunk, bool timed) {
  if (precision == JNE_PREC_AUTO) precision = auto_uses_mfma(dim, steps) ? JNE_PREC_MFMA_FP16X2 : JNE_PREC_FP64;
  for (size_t off = 0; off < n; off += chunk) {
    const uint32_t cn = (uint32_t)std::min(chunk, n - off);
    const uint32_t* s = d_seeds ? d_seeds + off : nullptr;
    const double* z = d_z ? d_z + off * (size_t)steps * dim : nullptr;
    auto moments = [&] {
      if (dim > jne::kMaxDim) return jne::launch_walk_moments_wide(dim, s, z, rng_stream, steps, cn, acc, cn, st);
      if (precision == JNE_PREC_FP64) return jne::launch_walk_moments_fp64(dim, s, z, rng_stream, steps, cn, acc, cn, st);
      return jne::launch_walk_moments_mfma(dim, precision, s, z, rng_stream, steps, cn, acc, cn, st);
    };
    if (timed)
      JNE_HIP_CHECK(ctx, timed_launch(ctx, JNE_KERNEL_MOMENTS, cn, st, moments));
    else
      JNE_HIP_CHECK(ctx, moments());
    for (int k = 0; k < n_models; ++k) {
      const int model = models[k];
      double* const out = d_outs[k] + off * (size_t)jne::num_eigenvalues(model, dim);
      auto eigen = [&] {
        if (dim > jne::kMaxDim) return jne::launch_walk_eigen_wide(model, dim, acc, cn, steps, cn, out, st);
        return jne::launch_walk_eigen_fp64(model, dim, acc, cn, steps, cn, out, mbuf, st);
      };
      if (timed)
        JNE_HIP_CHECK(ctx, timed_launch(ctx, JNE_KERNEL_EIGEN, cn, st, eigen));
      else
        JNE_HIP_CHECK(ctx, eigen());
    }
  }
  return JNE_OK;
}

int run_walks(jne_ctx* ctx, const int* models, int n_models, int dim, uint32_t steps, uint32_t rng_stream,
              const uint32_t* d_seeds, const double* d_z, size_t n, int precision, double* const* d_outs,
              hipStream_t st) {
  const size_t mc = (size_t)jne::moment_count(dim);
  // at most 2^18 walks per launch, and at the wide dims about 1 GiB of moments
  const size_t chunk = std::min(n, std::min(kChunkWalks, std::max<size_t>(65536, (size_t(1) << 27) / mc)));
  int rc = ensure(ctx, (void**)&ctx->d_acc, &ctx->acc_bytes, chunk * mc * sizeof(double));
  if (rc) return rc;
  rc = ensure(ctx, (void**)&ctx->d_m, &ctx->m_bytes, chunk * (size_t)jne::eigen_m_count(dim) * sizeof(double));
  if (rc) return rc;
  if (!ctx->scratch_free) JNE_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->scratch_free, hipEventDisableTiming));
  if (ctx->scratch_stream && ctx->scratch_stream != st)
    JNE_HIP_CHECK(ctx, hipStreamWaitEvent(st, ctx->scratch_free, 0));
  if ((rc = launch_walks(ctx, models, n_models, dim, steps, rng_stream, d_seeds, d_z, n, precision, d_outs, st,
                         ctx->d_acc, ctx->d_m, chunk, true)))
    return rc;
  JNE_HIP_CHECK(ctx, hipEventRecord(ctx->scratch_free, st));
  ctx->scratch_stream = st;
  return JNE_OK;
}

// The stream a device entry point enqueues on: JNE_STREAM_CONTEXT is the
// context's own stream, anything else (NULL included: HIP's null stream) is
// the caller's.
hipStream_t stream_of(jne_ctx* ctx, void* hip_stream) {
  return hip_stream == JNE_STREAM_CONTEXT ? ctx->stream : (hipStream_t)hip_stream;
}

// n rows of nf eigenvalues from the context's device buffer; a NaN row (S11
// not positive definite) makes the call return JNE_ERR_NUMERIC, naming the
// first such walk (0-based row).
int finish_host_out(jne_ctx* ctx, double* out, size_t n, size_t nf) {
  const size_t count = n * nf;
  JNE_HIP_CHECK(ctx, hipMemcpyAsync(out, ctx->d_out, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  JNE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < count; ++i)
    if (isnan(out[i]))
      return set_err(ctx, JNE_ERR_NUMERIC, "walk %zu: S11 is not positive definite", i / nf);
  return JNE_OK;
}

}  // namespace

extern "C" {

int jne_create(int device, jne_ctx** out) {
  if (!out) {
    g_create_err = "jne_create: out is NULL";
    return JNE_ERR_INVALID_ARG;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    g_create_err = std::string("jne_create: no HIP device: ") + hipGetErrorString(e);
    return JNE_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_err = "jne_create: device index out of range";
    return JNE_ERR_INVALID_ARG;
  }
  if (device >= kMaxDevices) {
    g_create_err = "jne_create: device index above 63";
    return JNE_ERR_INVALID_ARG;
  }
  jne_ctx* ctx = new jne_ctx();
  ctx->device = device;
  if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
    g_create_err = std::string("jne_create: ") + hipGetErrorString(e);
    delete ctx;
    return JNE_ERR_HIP;
  }
  {
    // the address of this device's copy of the no-alloc scratch, once per
    // device (with the device current)
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        g_create_err = std::string("jne_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return JNE_ERR_HIP;
      }
      g_scratch_ptr[device] = static_cast<double*>(p);
    }
    ctx->noalloc_scratch = g_scratch_ptr[device];
  }
  *out = ctx;
  return JNE_OK;
}

void jne_destroy(jne_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  release_timed(ctx);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->d_acc) (void)hipFree(ctx->d_acc);
  if (ctx->d_m) (void)hipFree(ctx->d_m);
  if (ctx->scratch_free) (void)hipEventDestroy(ctx->scratch_free);
  if (ctx->d_in) (void)hipFree(ctx->d_in);
  if (ctx->d_out) (void)hipFree(ctx->d_out);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* jne_last_error(const jne_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int jne_num_eigenvalues(int model, int dim) {
  if (model < 0 || model > 4 || dim < 1 || dim > 255) return JNE_ERR_INVALID_ARG;
  return jne::num_eigenvalues(model, dim);
}

int jne_eigenvalues_by_seed(jne_ctx* ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                            const uint32_t* seeds, size_t n, int precision, double* out) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_args(ctx, model, dim, steps, precision);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!seeds || !out) return set_err(ctx, JNE_ERR_INVALID_ARG, "seeds/out must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t nf = (size_t)jne::num_eigenvalues(model, dim);
  if ((rc = ensure(ctx, &ctx->d_in, &ctx->in_bytes, n * sizeof(uint32_t)))) return rc;
  if ((rc = ensure(ctx, &ctx->d_out, &ctx->out_bytes, n * nf * sizeof(double)))) return rc;
  JNE_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_in, seeds, n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
  double* const outs[1] = {(double*)ctx->d_out};
  if ((rc = run_walks(ctx, &model, 1, dim, steps, rng_stream, (const uint32_t*)ctx->d_in, nullptr, n, precision, outs,
                      ctx->stream)))
    return rc;
  return finish_host_out(ctx, out, n, nf);
}

int jne_eigenvalues_by_seed_device(jne_ctx* ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                                   const uint32_t* d_seeds, size_t n, int precision, double* d_out,
                                   void* hip_stream) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_args(ctx, model, dim, steps, precision);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!d_seeds || !d_out) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_seeds/d_out must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream_of(ctx, hip_stream);
  double* const outs[1] = {d_out};
  return run_walks(ctx, &model, 1, dim, steps, rng_stream, d_seeds, nullptr, n, precision, outs, st);
}

int jne_eigenvalues_by_seed_models_device(jne_ctx* ctx, const int* models, int n_models, int dim, uint32_t steps,
                                          uint32_t rng_stream, const uint32_t* d_seeds, size_t n, int precision,
                                          double* const* d_outs, void* hip_stream) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  if (!models || !d_outs || n_models < 1 || n_models > 5)
    return set_err(ctx, JNE_ERR_INVALID_ARG, "models/d_outs must not be NULL and n_models must be in 1..5");
  for (int k = 0; k < n_models; ++k) {
    int rc = check_args(ctx, models[k], dim, steps, precision);
    if (rc) return rc;
    if (n > 0 && !d_outs[k]) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_outs[%d] must not be NULL", k);
  }
  if (n == 0) return JNE_OK;
  if (!d_seeds) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_seeds must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return run_walks(ctx, models, n_models, dim, steps, rng_stream, d_seeds, nullptr, n, precision, d_outs,
                   stream_of(ctx, hip_stream));
}

int jne_eigenvalues_by_seed_device_noalloc(jne_ctx* ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                                           const uint32_t* d_seeds, size_t n, int precision, double* d_out,
                                           void* hip_stream) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_args(ctx, model, dim, steps, precision);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!d_seeds || !d_out) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_seeds/d_out must not be NULL");
  const size_t mc = (size_t)jne::moment_count(dim);
  const size_t per_walk = mc + (size_t)jne::eigen_m_count(dim);
  const size_t cap = (JNE_NOALLOC_SCRATCH_BYTES / sizeof(double)) / per_walk;
  if (n > cap)
    return set_err(ctx, JNE_ERR_INVALID_ARG,
                   "jne_eigenvalues_by_seed_device_noalloc: %zu walks exceed the scratch of %zu walks at dim %d "
                   "(%zu bytes per walk in %d MiB)",
                   n, cap, dim, per_walk * sizeof(double), (int)(JNE_NOALLOC_SCRATCH_BYTES >> 20));
  if (!ctx->noalloc_scratch) return set_err(ctx, JNE_ERR_HIP, "the no-alloc scratch has no device address");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  double* const acc = ctx->noalloc_scratch;
  double* const outs[1] = {d_out};
  return launch_walks(ctx, &model, 1, dim, steps, rng_stream, d_seeds, nullptr, n, precision, outs,
                      stream_of(ctx, hip_stream), acc, acc + mc * n, n, false);
}

int jne_eigenvalues_from_normals(jne_ctx* ctx, int model, int dim, uint32_t steps, const double* z, size_t n,
                                 int precision, double* out) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_args(ctx, model, dim, steps, precision);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!z || !out) return set_err(ctx, JNE_ERR_INVALID_ARG, "z/out must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t nf = (size_t)jne::num_eigenvalues(model, dim);
  const size_t zcount = n * (size_t)steps * dim;
  if ((rc = ensure(ctx, &ctx->d_in, &ctx->in_bytes, zcount * sizeof(double)))) return rc;
  if ((rc = ensure(ctx, &ctx->d_out, &ctx->out_bytes, n * nf * sizeof(double)))) return rc;
  JNE_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_in, z, zcount * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  double* const outs[1] = {(double*)ctx->d_out};
  if ((rc = run_walks(ctx, &model, 1, dim, steps, 0, nullptr, (const double*)ctx->d_in, n, precision, outs,
                      ctx->stream)))
    return rc;
  return finish_host_out(ctx, out, n, nf);
}

namespace {
int check_normals_args(jne_ctx* ctx, int dim, uint32_t steps, size_t n) {
  if (dim < 1 || dim > 255) return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid dim %d", dim);
  if (steps < 1 || steps > (1u << 30)) return set_err(ctx, JNE_ERR_INVALID_ARG, "invalid steps %u", steps);
  if ((uint64_t)steps * dim > 0xFFFFFFFCull)
    return set_err(ctx, JNE_ERR_INVALID_ARG, "jne_walk_normals: steps * dim must stay below 2^32");
  if (n > 0xFFFFFFFFull) return set_err(ctx, JNE_ERR_INVALID_ARG, "jne_walk_normals: at most 2^32-1 walks");
  return JNE_OK;
}
}  // namespace

int jne_walk_normals(jne_ctx* ctx, int dim, uint32_t steps, uint32_t rng_stream, const uint32_t* seeds, size_t n,
                     float* out) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_normals_args(ctx, dim, steps, n);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!seeds || !out) return set_err(ctx, JNE_ERR_INVALID_ARG, "seeds/out must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t total = n * (size_t)steps * dim;
  if ((rc = ensure(ctx, &ctx->d_in, &ctx->in_bytes, n * sizeof(uint32_t)))) return rc;
  if ((rc = ensure(ctx, &ctx->d_out, &ctx->out_bytes, total * sizeof(float)))) return rc;
  JNE_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_in, seeds, n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
  JNE_HIP_CHECK(ctx, timed_launch(ctx, JNE_KERNEL_NORMALS, n, ctx->stream, [&] {
                  return jne::launch_walk_normals(dim, (const uint32_t*)ctx->d_in, rng_stream, steps, (uint32_t)n,
                                                  (float*)ctx->d_out, ctx->stream);
                }));
  JNE_HIP_CHECK(ctx, hipMemcpyAsync(out, ctx->d_out, total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  JNE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return JNE_OK;
}

int jne_walk_normals_device(jne_ctx* ctx, int dim, uint32_t steps, uint32_t rng_stream, const uint32_t* d_seeds,
                            size_t n, float* d_out, void* hip_stream) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  int rc = check_normals_args(ctx, dim, steps, n);
  if (rc) return rc;
  if (n == 0) return JNE_OK;
  if (!d_seeds || !d_out) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_seeds/d_out must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream_of(ctx, hip_stream);
  JNE_HIP_CHECK(ctx, timed_launch(ctx, JNE_KERNEL_NORMALS, n, st, [&] {
                  return jne::launch_walk_normals(dim, d_seeds, rng_stream, steps, (uint32_t)n, d_out, st);
                }));
  return JNE_OK;
}

int jne_percentiles_device(jne_ctx* ctx, const double* d_values, uint64_t n, const double* p, size_t np,
                           double* out, void* hip_stream) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->err.clear();
  if (!p || !out || np < 1 || np > (size_t)jne::kMaxPercentiles)
    return set_err(ctx, JNE_ERR_INVALID_ARG, "p/out must not be NULL and np must be 1..%d", jne::kMaxPercentiles);
  for (size_t i = 0; i < np; ++i)
    if (!(p[i] >= 0.0 && p[i] <= 1.0)) return set_err(ctx, JNE_ERR_INVALID_ARG, "p[%zu] is not in [0, 1]", i);
  if (n > 0x7FFFFFFFull) return set_err(ctx, JNE_ERR_INVALID_ARG, "n must be below 2^31");
  if (n == 0) {  // as jne_percentiles: no order statistics
    for (size_t i = 0; i < np; ++i) out[i] = NAN;
    return JNE_OK;
  }
  if (!d_values) return set_err(ctx, JNE_ERR_INVALID_ARG, "d_values must not be NULL");
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream_of(ctx, hip_stream);
  const int rc = jne::quantile_select(ctx->qs, &d_values, 1, n, p, (int)np, out, st, nullptr);
  if (rc == JNE_ERR_NUMERIC) return set_err(ctx, rc, "a value is NaN");
  if (rc) return set_err(ctx, rc, "percentile selection failed");
  return JNE_OK;
}

int jne_synchronize(jne_ctx* ctx) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  JNE_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return JNE_OK;
}

int jne_profile_enable(jne_ctx* ctx, int enable) {
  if (!ctx) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  for (auto& t : ctx->timed) {
    (void)hipEventSynchronize(t.stop);
  }
  release_timed(ctx);
  ctx->profiling = enable != 0;
  return JNE_OK;
}

int jne_profile_read(jne_ctx* ctx, int kernel, double* total_ms, uint64_t* launches, uint64_t* walks) {
  if (!ctx || !total_ms || !launches || !walks) return JNE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(ctx->mu);
  JNE_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  double ms = 0.0;
  uint64_t nl = 0, nw = 0;
  for (auto& t : ctx->timed) {
    if (t.kernel != kernel) continue;
    JNE_HIP_CHECK(ctx, hipEventSynchronize(t.stop));
    float e = 0.0f;
    JNE_HIP_CHECK(ctx, hipEventElapsedTime(&e, t.start, t.stop));
    ms += e;
    ++nl;
    nw += t.walks;
  }
  *total_ms = ms;
  *launches = nl;
  *walks = nw;
  return JNE_OK;
}

}  // extern "C"
