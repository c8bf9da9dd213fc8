// jne_internal.h -- shared definitions of the walk kernels (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace jne {

constexpr int kMaxDim = 12;

// Lanes per walk (L) and dim padded to a multiple of L (DP) for the FP64
// moment kernel.  Chosen so that one lane holds <= ~90 double accumulators.
constexpr int lanes_for_dim(int d) {
  return d <= 5 ? 1 : d <= 8 ? 2 : d == 9 ? 3 : d == 10 ? 5 : 4;
}
template <int D>
struct WalkShape {
  static constexpr int L = lanes_for_dim(D);
  static constexpr int DP = ((D + L - 1) / L) * L;
};

// Number of per-walk moment doubles written by the FP64 moment kernel:
// sum b b^T (D*D), sum e b^T (D*D), 6 moment vectors (6*D).
constexpr int moment_count(int d) { return 2 * d * d + 6 * d; }

constexpr int num_eigenvalues(int model, int d) { return (model == 1 || model == 3) ? d + 1 : d; }

hipError_t launch_walk_moments_fp64(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream,
                                    uint32_t T, uint32_t n, double* acc, size_t ld, hipStream_t st);
hipError_t launch_walk_eigen_fp64(int model, int d, const double* acc, size_t ld, uint32_t T, uint32_t n,
                                  double* out, hipStream_t st);
hipError_t launch_walk_normals(int d, const uint32_t* seeds, uint32_t rng_stream, uint32_t T, uint32_t n,
                               float* out, hipStream_t st);

}  // namespace jne
