// jne_internal.h -- shared definitions of the walk kernels (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace jne {

constexpr int kMaxDim = 12;      // the register-resident kernels (and the MFMA pass)
constexpr int kMaxWideDim = 32;  // the runtime-dim FP64 path (jne_walk_wide.hip), dims 13..32
constexpr uint32_t kMaxSteps = 1u << 30;  // steps per walk accepted by the C ABI
// A walk of dim <= kMaxDim draws at most kMaxDim * kMaxSteps / 4 Philox calls,
// so the counter's high word is zero in the register-resident kernels: they
// count calls in 32 bits, which keeps the first rounds' products of a walk
// whose seed is wave-uniform on the scalar unit.
static_assert((uint64_t)kMaxDim * kMaxSteps / 4 < (1ull << 32), "32-bit Philox counters");

// Per-dim shape of the FP64 moment kernel: lanes per walk (L) and target
// waves per SIMD (the register cap).  Measured on MI355X (model 4, T = 10^4).
// Dims with L = 1 run the register-only kernel (its own wave targets); one
// lane per walk at dims 7 and 8 would need a wave per SIMD and ran 11% and
// 36% slower than the LDS-staged shapes.
struct Fp64Shape {
  int lanes, waves;
};
#ifndef JNE_FP64_SHAPES
#define JNE_FP64_SHAPES \
  {1, 2}, {1, 2}, {1, 2}, {1, 2}, {1, 2}, {1, 2}, {7, 4}, {2, 2}, {9, 4}, {5, 3}, {4, 2}, {4, 2}
#endif
constexpr Fp64Shape kFp64Shapes[12] = {JNE_FP64_SHAPES};
constexpr int lanes_for_dim(int d) { return kFp64Shapes[d - 1].lanes; }
constexpr int waves_for_dim(int d) { return kFp64Shapes[d - 1].waves; }
constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

// Steps per round of the FP64 moment kernel: a multiple of 4/gcd(d,4) (so a
// round is a whole number of 4-normal Philox calls), at least 4, preferring a
// call count divisible by L while one wave's double-buffered LDS stays <= 24 KiB.
constexpr int steps_per_round(int d, int L) {
  const int s0 = 4 / gcd_int(d, 4);
  const int first = ((4 + s0 - 1) / s0) * s0;
  for (int s = first; s <= 8 * s0; s += s0) {
    const int calls = s * d / 4;
    const int lds = 2 * (64 / L) * s * d * 8;
    if (calls % L == 0 && lds <= 24 * 1024) return s;
  }
  return first;
}

template <int D>
struct WalkShape {
  static constexpr int L = lanes_for_dim(D);          // lanes per walk
  static constexpr int DP = ((D + L - 1) / L) * L;     // dim padded to L
  static constexpr int GPW = 64 / L;                   // walks per wave
  static constexpr int S = steps_per_round(D, L);      // steps per round
  static constexpr int C = S * D / 4;                  // Philox calls per round
  static_assert((S * D) % 4 == 0, "a round must be whole Philox calls");
};

// Step s and component i (n = s*D + i) of the normals n = 4c .. 4c+3 drawn by
// Philox call c of a round.  One division where the position is created, then
// constant strides: the loops over calls advance c by a compile-time step.
// A compile-time flag as a type, for generic lambdas that take a branch-free
// copy of a loop body (the tail form of a segment or batch).
template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

template <int D>
struct CallPos {
  int s, i;  // of normal 4c
  __device__ __forceinline__ explicit CallPos(int c) : s((4 * c) / D), i(4 * c - ((4 * c) / D) * D) {}
  template <int DC>
  __device__ __forceinline__ void advance() {
    constexpr int ds = (4 * DC) / D, di = (4 * DC) % D;
    s += ds;
    i += di;
    if (i >= D) {
      i -= D;
      ++s;
    }
  }
  // position of normal 4c + j, j = 0..3
  __device__ __forceinline__ void at(int j, int& sj, int& ij) const {
    // D % 4 == 0: i is a multiple of 4, so i + 3 < D; otherwise up to 3 wraps (D < 4)
    constexpr int kWraps = D % 4 == 0 ? 0 : D >= 4 ? 1 : 3;
    sj = s;
    ij = i + j;
#pragma unroll
    for (int w = 0; w < kWraps; ++w) {
      if (ij >= D) {
        ij -= D;
        ++sj;
      }
    }
  }
};

// Number of per-walk moment doubles written by the FP64 moment kernel:
// sum b b^T (D*D), sum e b^T (D*D), 6 moment vectors (6*D).
constexpr int moment_count(int d) { return 2 * d * d + 6 * d; }

constexpr int num_eigenvalues(int model, int d) { return (model == 1 || model == 3) ? d + 1 : d; }

hipError_t launch_walk_moments_fp64(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                    uint32_t n, double* acc, size_t ld, hipStream_t st);
// the segment-split pass: the MFMA precisions and the compensated FP32 VALU check
hipError_t launch_walk_moments_mfma(int d, int precision, const uint32_t* seeds, const double* zin,
                                    uint32_t rng_stream, uint32_t T, uint32_t n, double* acc, size_t ld,
                                    hipStream_t st);
// mbuf: (eigen_m_count(d)) * ld doubles of scratch for the two-kernel form
// (dims whose fused solve would hold fewer than 2 waves per SIMD)
// (none for the runtime-dim path, whose solve keeps M in LDS)
constexpr int eigen_m_count(int d) { return d <= kMaxDim ? d * (d + 1) / 2 + 1 : 0; }
hipError_t launch_walk_eigen_fp64(int model, int d, const double* acc, size_t ld, uint32_t T, uint32_t n,
                                  double* out, double* mbuf, hipStream_t st);
// the runtime-dim FP64 path, dims kMaxDim+1 .. kMaxWideDim
hipError_t launch_walk_moments_wide(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                    uint32_t n, double* acc, size_t ld, hipStream_t st);
hipError_t launch_walk_eigen_wide(int model, int d, const double* acc, size_t ld, uint32_t T, uint32_t n, double* out,
                                  hipStream_t st);
hipError_t launch_walk_normals(int d, const uint32_t* seeds, uint32_t rng_stream, uint32_t T, uint32_t n,
                               float* out, hipStream_t st);

// Percentiles of a device array (jne_quantile.hip): device scratch, grown on
// demand and kept by its owner (a context, the query server's finisher).
constexpr int kMaxPercentiles = 8;
constexpr int kMaxArrays = 2;  // arrays of one call (the server's trace and max-eig)
struct QuantileScratch {
  void* small = nullptr;  // histogram, selection state, gathered values, radix-select state
  size_t small_bytes = 0;
  ~QuantileScratch();
};
// out[a*np + i] = percentile p[i] (0 <= p[i] <= 1, np <= kMaxPercentiles) of
// xs[a][0..n), a < nx <= kMaxArrays, 1 <= n < 2^31, by the rule of
// get_percentile_value; synchronous on st.  Returns 0, JNE_ERR_NUMERIC (a NaN
// in an array) or another JNE_ERR_* code; *used_fallback tells whether an
// array's order statistics came from the radix select instead of the bucket
// selection.
int quantile_select(QuantileScratch& sc, const double* const* xs, int nx, uint64_t n, const double* p, int np,
                    double* out, hipStream_t st, bool* used_fallback);

}  // namespace jne
