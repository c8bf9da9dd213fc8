// jne_quantile.hip -- percentiles of a device array by bucket selection.
//
// The percentile rule is the reference's get_percentile_value
// (src/simulation_analyzers.rs:4-18): rank p*(n-1) in the ascending array,
// linear interpolation between the neighbouring order statistics.  Only the
// order statistics at floor(rank) and ceil(rank) are needed, so instead of
// sorting the whole array:
//
//   1. min, max and a NaN flag (one pass);
//   2. a histogram of kBuckets equal-width value buckets between min and max
//      (one pass).  The bucket index floor((x - min) * scale) is monotone in
//      x, so the k-th smallest value lies in the bucket where the cumulative
//      count first exceeds k;
//   3. one workgroup finds the bucket (and the rank inside it) of every needed
//      order statistic;
//   4. the values of those few buckets are gathered (one pass);
//   5. one workgroup sorts them in LDS and reads the order statistics.
//
// When the gathered buckets hold more than kMaxCand values (heavy ties, or a
// few extreme values stretching the bucket width), or min/max are not finite,
// the order statistics come from a radix select instead: eight passes of
// 8-bit digits over the order-preserving 64-bit keys, MSB first.  Either way
// the results are the exact order statistics, identical to sorting.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/jne.h"
#include "jne_internal.h"

namespace jne {
namespace {

constexpr int kBuckets = 8192;
constexpr int kMaxCand = 4096;
constexpr int kMaxRanks = 2 * kMaxPercentiles;

// Order-preserving map of a double to an unsigned key (NaNs sort past the
// infinities and are caught separately).
__device__ __forceinline__ uint64_t ord_key(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Bucket selection state of one array, written by qs_select and read by
// qs_gather/qs_finish.
struct Sel {
  uint32_t status;  // kSelect, kRadix, kNan or kEqual
  uint32_t nb;      // distinct buckets holding needed order statistics
  uint32_t m;       // values gathered
  uint32_t bucket[kMaxRanks];
  uint32_t base[kMaxRanks];     // first slot of bucket j's values
  uint32_t count[kMaxRanks];    // values in bucket j
  uint32_t fill[kMaxRanks];     // gather cursors
  uint32_t pos[kMaxRanks];      // slot of rank k among the sorted gathered values
};
constexpr uint32_t kSelect = 0, kRadix = 1, kNan = 2, kEqual = 3;

// Everything a launch needs, passed by value: no host-to-device copies.
struct QsArgs {
  const double* x[kMaxArrays];
  uint64_t n;
  int np;
  uint64_t ranks[kMaxRanks];  // floor and ceil of rank i at 2i, 2i+1
  double prank[kMaxPercentiles];
};

// Per-array device scratch: mm[3] (min key, max key, NaN flag), the
// histogram, the selection state and the gathered values; then the results.
struct QsArray {
  uint64_t mm[4];
  uint32_t hist[kBuckets];
  Sel sel;
  double cand[kMaxCand];
};
struct QsResult {
  uint32_t status, pad;
  double out[kMaxPercentiles];
};

__global__ void qs_init(QsArray* arr) {
  QsArray& a = arr[blockIdx.y];
  const int t = threadIdx.x;
  if (t == 0) {
    a.mm[0] = ~0ull;
    a.mm[1] = 0ull;
    a.mm[2] = 0ull;
  }
  for (int b = t; b < kBuckets; b += blockDim.x) a.hist[b] = 0u;
}

__global__ __launch_bounds__(256) void qs_minmax(QsArgs args, QsArray* arr) {
  __shared__ uint64_t smin[256], smax[256];
  __shared__ int snan;
  const double* __restrict__ x = args.x[blockIdx.y];
  QsArray& a = arr[blockIdx.y];
  const int t = threadIdx.x;
  if (t == 0) snan = 0;
  uint64_t lo = ~0ull, hi = 0ull;
  bool nan = false;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < args.n; i += (uint64_t)gridDim.x * 256) {
    const double v = x[i];
    nan |= isnan(v);
    const uint64_t k = ord_key(v);
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  smin[t] = lo;
  smax[t] = hi;
  __syncthreads();
  if (nan) snan = 1;
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      smin[t] = smin[t + s] < smin[t] ? smin[t + s] : smin[t];
      smax[t] = smax[t + s] > smax[t] ? smax[t + s] : smax[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    atomicMin((unsigned long long*)&a.mm[0], (unsigned long long)smin[0]);
    atomicMax((unsigned long long*)&a.mm[1], (unsigned long long)smax[0]);
    if (snan) atomicOr((unsigned long long*)&a.mm[2], 1ull);
  }
}

__device__ __forceinline__ double from_key(uint64_t k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}

// The bucket grid of an array from its min and max: kSelect when it is
// usable, else the reason (every thread of every kernel decides the same).
__device__ __forceinline__ uint32_t grid_of(const QsArray& a, double& lo, double& scale) {
  lo = from_key(a.mm[0]);
  const double hi = from_key(a.mm[1]);
  if (a.mm[2]) return kNan;
  if (lo == hi) return kEqual;
  const double width = hi - lo;
  if (!(isfinite(lo) && isfinite(hi) && isfinite(width) && width > 0.0)) return kRadix;
  scale = (double)kBuckets / width;
  return kSelect;
}

__device__ __forceinline__ uint32_t bucket_of(double v, double lo, double scale) {
  const double f = (v - lo) * scale;  // >= 0, monotone in v
  return f < (double)(kBuckets - 1) ? (uint32_t)f : (uint32_t)(kBuckets - 1);
}

__global__ __launch_bounds__(256) void qs_hist(QsArgs args, QsArray* arr) {
  __shared__ uint32_t h[kBuckets];
  const double* __restrict__ x = args.x[blockIdx.y];
  QsArray& a = arr[blockIdx.y];
  double lo, scale;
  if (grid_of(a, lo, scale) != kSelect) return;  // the same for every thread
  const int t = threadIdx.x;
  for (int b = t; b < kBuckets; b += 256) h[b] = 0u;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < args.n; i += (uint64_t)gridDim.x * 256)
    atomicAdd(&h[bucket_of(x[i], lo, scale)], 1u);
  __syncthreads();
  for (int b = t; b < kBuckets; b += 256)
    if (h[b]) atomicAdd(&a.hist[b], h[b]);
}

// One workgroup of 256 threads per array, each owning kBuckets/256
// consecutive buckets.
__global__ __launch_bounds__(256) void qs_select(QsArgs args, QsArray* arr) {
  constexpr int kPer = kBuckets / 256;
  __shared__ uint64_t part[256];
  __shared__ uint32_t rb[kMaxRanks];  // bucket of rank k
  __shared__ uint64_t rl[kMaxRanks];  // rank of k inside its bucket
  QsArray& a = arr[blockIdx.y];
  Sel* sel = &a.sel;
  const int t = threadIdx.x;
  const int nr = 2 * args.np;
  double lo, scale;
  const uint32_t g = grid_of(a, lo, scale);
  if (g != kSelect) {
    if (t == 0) sel->status = g;
    return;
  }
  uint64_t s = 0;
  for (int j = 0; j < kPer; ++j) s += a.hist[t * kPer + j];
  part[t] = s;
  __syncthreads();
  if (t == 0) {  // exclusive prefix of the 256 chunk sums
    uint64_t acc = 0;
    for (int j = 0; j < 256; ++j) {
      const uint64_t v = part[j];
      part[j] = acc;
      acc += v;
    }
  }
  __syncthreads();
  uint64_t cum = part[t];
  for (int j = 0; j < kPer; ++j) {
    const int b = t * kPer + j;
    const uint64_t c = a.hist[b];
    for (int k = 0; k < nr; ++k) {
      if (args.ranks[k] >= cum && args.ranks[k] < cum + c) {
        rb[k] = (uint32_t)b;
        rl[k] = args.ranks[k] - cum;
      }
    }
    cum += c;
  }
  __syncthreads();
  if (t == 0) {
    // distinct buckets in increasing order, their slots among the gathered values
    uint32_t nb = 0, m = 0;
    bool over = false;
    for (int k = 0; k < nr; ++k) {
      bool seen = false;
      for (uint32_t j = 0; j < nb; ++j) seen |= sel->bucket[j] == rb[k];
      if (!seen) sel->bucket[nb++] = rb[k];
    }
    for (uint32_t i = 1; i < nb; ++i)  // insertion sort, nb <= kMaxRanks
      for (uint32_t j = i; j > 0 && sel->bucket[j - 1] > sel->bucket[j]; --j) {
        const uint32_t tmp = sel->bucket[j];
        sel->bucket[j] = sel->bucket[j - 1];
        sel->bucket[j - 1] = tmp;
      }
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t c = a.hist[sel->bucket[j]];
      sel->base[j] = m;
      sel->count[j] = c;
      sel->fill[j] = 0u;
      over |= (uint64_t)m + c > (uint64_t)kMaxCand;
      m += over ? 0u : c;
    }
    for (int k = 0; k < nr; ++k)
      for (uint32_t j = 0; j < nb; ++j)
        if (sel->bucket[j] == rb[k]) sel->pos[k] = sel->base[j] + (uint32_t)rl[k];
    sel->nb = nb;
    sel->m = m;
    sel->status = over ? kRadix : kSelect;
  }
}

__global__ __launch_bounds__(256) void qs_gather(QsArgs args, QsArray* arr) {
  __shared__ uint32_t bk[kMaxRanks];
  const double* __restrict__ x = args.x[blockIdx.y];
  QsArray& a = arr[blockIdx.y];
  Sel* sel = &a.sel;
  if (sel->status != kSelect) return;  // the same for every thread
  double lo, scale;
  (void)grid_of(a, lo, scale);
  const uint32_t nb = sel->nb;
  if (threadIdx.x < nb) bk[threadIdx.x] = sel->bucket[threadIdx.x];
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < args.n; i += (uint64_t)gridDim.x * 256) {
    const double v = x[i];
    const uint32_t b = bucket_of(v, lo, scale);
    for (uint32_t j = 0; j < nb; ++j)
      if (b == bk[j]) {
        const uint32_t slot = atomicAdd(&sel->fill[j], 1u);
        a.cand[sel->base[j] + slot] = v;
      }
  }
}

// The interpolation of get_percentile_value on the order statistics s(lo), s(hi).
__device__ __forceinline__ double interp(double rank, uint64_t lo, uint64_t hi, double slo, double shi) {
  if (lo == hi) return slo;
  const double w = rank - (double)lo;
  return slo * (1.0 - w) + shi * w;
}

// One workgroup per array: bitonic sort of the m <= kMaxCand gathered values
// in LDS, then the percentiles; the status and values into res.
__global__ __launch_bounds__(1024) void qs_finish(QsArgs args, const QsArray* arr, QsResult* res) {
  __shared__ double v[kMaxCand];
  const QsArray& a = arr[blockIdx.y];
  const Sel& sel = a.sel;
  QsResult& r = res[blockIdx.y];
  const int t = threadIdx.x;
  const uint32_t status = sel.status;
  if (t == 0) r.status = status;
  if (status == kEqual) {  // every order statistic is the minimum
    if (t < args.np) r.out[t] = from_key(a.mm[0]);
    return;
  }
  if (status != kSelect) return;  // the same for every thread
  const uint32_t m = sel.m;
  uint32_t len = 1;
  while (len < m) len <<= 1;
  for (uint32_t i = t; i < len; i += blockDim.x) v[i] = i < m ? a.cand[i] : INFINITY;
  __syncthreads();
  for (uint32_t size = 2; size <= len; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = t; i < len; i += blockDim.x) {
        const uint32_t j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const double x0 = v[i], x1 = v[j];
          if ((x0 > x1) == up) {
            v[i] = x1;
            v[j] = x0;
          }
        }
      }
      __syncthreads();
    }
  if (t < args.np)
    r.out[t] = interp(args.prank[t], args.ranks[2 * t], args.ranks[2 * t + 1], v[sel.pos[2 * t]],
                      v[sel.pos[2 * t + 1]]);
}

// ---- the fallback: MSB-first radix select on the keys of ord_key.  Pass L
// histograms digit L (8 bits) of the keys whose first L digits equal the
// prefix found so far for each needed rank, one histogram per rank; the pick
// step extends each prefix by the digit whose cumulative count passes the
// rank, and the rank becomes the rank among the keys of that digit.  After
// the eighth pass each prefix is the whole key of its order statistic, exact
// whatever the ties, outliers or infinities, with no sort of the array.
constexpr int kDigits = 8;

struct Radix {
  uint64_t prefix[kMaxRanks];  // the digits found so far
  uint64_t k[kMaxRanks];       // rank among the keys with that prefix
  uint32_t hist[kMaxRanks][256];
};

__global__ void qs_radix_init(QsArgs args, Radix* rx) {
  Radix& r = rx[blockIdx.y];
  const int t = threadIdx.x;
  if (t < 2 * args.np) {
    r.prefix[t] = 0;
    r.k[t] = args.ranks[t];
  }
  for (int e = t; e < kMaxRanks * 256; e += blockDim.x) (&r.hist[0][0])[e] = 0u;
}

__global__ __launch_bounds__(256) void qs_radix_hist(QsArgs args, Radix* rx, int level) {
  __shared__ uint32_t h[kMaxRanks][256];
  __shared__ uint64_t pre[kMaxRanks];
  const double* __restrict__ x = args.x[blockIdx.y];
  Radix& r = rx[blockIdx.y];
  const int t = threadIdx.x;
  const int nr = 2 * args.np;
  for (int e = t; e < nr * 256; e += 256) (&h[0][0])[e] = 0u;
  if (t < nr) pre[t] = r.prefix[t];
  __syncthreads();
  const int shift = 56 - 8 * level;  // this pass's digit
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < args.n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t key = ord_key(x[i]);
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const uint64_t high = level == 0 ? 0ull : key >> (shift + 8);
    for (int k = 0; k < nr; ++k)
      if (high == pre[k]) atomicAdd(&h[k][d], 1u);
  }
  __syncthreads();
  for (int e = t; e < nr * 256; e += 256) {
    const uint32_t c = (&h[0][0])[e];
    if (c) atomicAdd(&(&r.hist[0][0])[e], c);
  }
}

// One thread per needed rank; the histograms are then cleared for the next pass.
__global__ void qs_radix_pick(QsArgs args, Radix* rx) {
  Radix& r = rx[blockIdx.y];
  const int t = threadIdx.x;
  if (t < 2 * args.np) {
    const uint64_t k = r.k[t];
    uint64_t cum = 0;
    uint32_t d = 0;
    for (; d < 255u; ++d) {  // the keys with this prefix number more than k
      const uint64_t c = r.hist[t][d];
      if (k < cum + c) break;
      cum += c;
    }
    r.k[t] = k - cum;
    r.prefix[t] = (r.prefix[t] << 8) | d;
  }
  __syncthreads();
  for (int e = t; e < kMaxRanks * 256; e += blockDim.x) (&r.hist[0][0])[e] = 0u;
}

__global__ void qs_radix_finish(QsArgs args, const Radix* rx, QsResult* res) {
  const Radix& r = rx[blockIdx.y];
  const int t = threadIdx.x;
  if (t < args.np)
    res[blockIdx.y].out[t] = interp(args.prank[t], args.ranks[2 * t], args.ranks[2 * t + 1], from_key(r.prefix[2 * t]),
                                    from_key(r.prefix[2 * t + 1]));
}

hipError_t grow(void** p, size_t* have, size_t need) {
  if (*have >= need) return hipSuccess;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  const hipError_t e = hipMalloc(p, need);
  if (e == hipSuccess) *have = need;
  return e;
}

}  // namespace

QuantileScratch::~QuantileScratch() {
  if (small) (void)hipFree(small);
}

int quantile_select(QuantileScratch& sc, const double* const* xs, int nx, uint64_t n, const double* p, int np,
                    double* out, hipStream_t st, bool* used_fallback) {
  if (used_fallback) *used_fallback = false;
  if (nx < 1 || nx > kMaxArrays || np < 1 || np > kMaxPercentiles || n < 1 || n > 0x7FFFFFFFull)
    return JNE_ERR_INVALID_ARG;
  const size_t res_off = ((size_t)kMaxArrays * sizeof(QsArray) + 255) & ~size_t(255);
  const size_t radix_off = (res_off + kMaxArrays * sizeof(QsResult) + 255) & ~size_t(255);
  hipError_t e = grow(&sc.small, &sc.small_bytes, radix_off + kMaxArrays * sizeof(Radix));
  if (e != hipSuccess) return JNE_ERR_HIP;
  QsArray* arr = static_cast<QsArray*>(sc.small);
  QsResult* d_res = reinterpret_cast<QsResult*>(static_cast<char*>(sc.small) + res_off);
  Radix* radix = reinterpret_cast<Radix*>(static_cast<char*>(sc.small) + radix_off);

  // ranks of the needed order statistics (get_percentile_value's rank)
  QsArgs args = {};
  for (int a = 0; a < nx; ++a) args.x[a] = xs[a];
  args.n = n;
  args.np = np;
  for (int i = 0; i < np; ++i) {
    args.prank[i] = p[i] * (double)(n - 1);
    args.ranks[2 * i] = (uint64_t)floor(args.prank[i]);
    args.ranks[2 * i + 1] = (uint64_t)ceil(args.prank[i]);
  }
  const unsigned grid = (unsigned)std::min<uint64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(qs_init, dim3(1, nx), dim3(256), 0, st, arr);
  hipLaunchKernelGGL(qs_minmax, dim3(grid, nx), dim3(256), 0, st, args, arr);
  hipLaunchKernelGGL(qs_hist, dim3(std::min(grid, 512u), nx), dim3(256), 0, st, args, arr);
  hipLaunchKernelGGL(qs_select, dim3(1, nx), dim3(256), 0, st, args, arr);
  hipLaunchKernelGGL(qs_gather, dim3(grid, nx), dim3(256), 0, st, args, arr);
  hipLaunchKernelGGL(qs_finish, dim3(1, nx), dim3(1024), 0, st, args, arr, d_res);
  QsResult res[kMaxArrays];
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(res, d_res, nx * sizeof(QsResult), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return JNE_ERR_HIP;
  for (int a = 0; a < nx; ++a)
    if (res[a].status == kNan) return JNE_ERR_NUMERIC;
  for (int a = 0; a < nx; ++a) {
    if (res[a].status != kRadix) {
      memcpy(out + (size_t)a * np, res[a].out, np * sizeof(double));
      continue;
    }
    // the fallback: radix select of this array's order statistics
    if (used_fallback) *used_fallback = true;
    QsArgs one = args;
    one.x[0] = xs[a];
    Radix* const rx = radix + a;
    hipLaunchKernelGGL(qs_radix_init, dim3(1), dim3(256), 0, st, one, rx);
    for (int level = 0; level < kDigits; ++level) {
      hipLaunchKernelGGL(qs_radix_hist, dim3(std::min(grid, 512u)), dim3(256), 0, st, one, rx, level);
      hipLaunchKernelGGL(qs_radix_pick, dim3(1), dim3(256), 0, st, one, rx);
    }
    hipLaunchKernelGGL(qs_radix_finish, dim3(1), dim3(64), 0, st, one, rx, d_res + a);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(out + (size_t)a * np, d_res[a].out, np * sizeof(double), hipMemcpyDeviceToHost, st)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
      return JNE_ERR_HIP;
  }
  return 0;
}

}  // namespace jne
