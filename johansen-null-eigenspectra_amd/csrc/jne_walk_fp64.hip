// jne_walk_fp64.hip -- FP64 walk path: one-pass moments + FP64 eigen epilogue.
//
// Replaces, per walk, the reference's
//   brownian_motion_matrix      src/rng_matrix.rs:57-141
//   calculate_eigenvalues       src/johansen_statistics.rs:59-85
//   construct_f_matrix          src/johansen_statistics.rs:102-197
//   sum_of_outer_products       src/matrix_utils.rs:67-85
//   GeneralizedEigen (dggev)    src/johansen_statistics.rs:35-45
// without materialising B, dB or F.  Kernel 1 streams the walk once and keeps
// the moments of x_t = [b_t, 1, u_t, u_t^2, e_t] (b_t = B_t before step t,
// e_t = B_{t+1} - B_t, u_t = (t+1)/T) in FP64; kernel 2 forms S11 and S01 of
// the requested model from those moments (demeaning and detrending are Schur
// complements on the {1, u} columns) and solves the eigenproblem -- as one
// kernel, or for the larger dims as two (algebra up to M = W^T W, then Jacobi).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "jne_internal.h"
#include "jne_rng.h"
#include "jne_tridiag.h"

// The smallest dim whose eigenvalues come from the tridiagonal QL solve
// (jne_tridiag.h) instead of cyclic Jacobi: 1.2x (dim 7) to 1.55x (dim 12)
// faster there; below dim 7 the two time alike (dim 5) or Jacobi leads, and
// Jacobi keeps the higher relative accuracy of the smallest eigenvalues of
// M = W^T W (QL's error is absolute, a few ulp of |M|, as with dggev).
#ifndef JNE_EIGEN_QL_MIN_DIM
#define JNE_EIGEN_QL_MIN_DIM 7
#endif
// Tuning: unroll the steps of a round (rolled and unrolled time alike).
#ifndef JNE_FP64_UNROLL_STEPS
#define JNE_FP64_UNROLL_STEPS 0
#endif
// The Philox calls of a lane in a round are unrolled, so their chains
// overlap: +3-6% at dims 8, 10, 11 and 12, with the same waves per SIMD (at
// dims 7 and 9 a lane draws one call per round).
#ifndef JNE_FP64_UNROLL_CALLS
#define JNE_FP64_UNROLL_CALLS 1
#endif
// dims with one lane per walk: the register-only kernel (0: the LDS-staged one)
#ifndef JNE_FP64_REG
#define JNE_FP64_REG 1
#endif


namespace jne {

// ---------------------------------------------------------------------------
// Kernel 1: moments.
//
// A walk is owned by a group of L lanes.  Every lane keeps the whole
// state vector b (dim doubles, in a lane-specific rotated order) and
// accumulates its share of the Gram entries:
//   rows k < R of  sum e b^T               (R = DP/L, DP = dim padded to L)
//   rows k < R, cyclic offsets 0..DP/2 of  sum b b^T
//   rows k < R of  sum u^p b, sum u^p e    (p = 0, 1, 2)
// Lane r's row k is component (r*R + k) mod DP, so the L lanes cover every
// entry; register indices are compile-time constants.
//
// The walk advances in rounds of S steps.  The S*dim normals of a round are
// S*dim/4 Philox calls, dealt round-robin to the group's lanes; each normal
// is scaled by sqrt(1/T) once and stored to LDS as a double, twice per step
// row (row = [w_0..w_{dim-1}, 0.., w_0..w_{dim-1}, 0..], 2*DP entries) so that
// a lane's rotated read of DP entries is one base address plus immediates.
// The walk index is the fastest LDS dimension (odd pitch) so that the lanes
// of a wave read from different banks.
// ---------------------------------------------------------------------------
// FP64 Gram accumulator.
struct GramAcc {
  double s;
  __device__ __forceinline__ void zero() { s = 0.0; }
  __device__ __forceinline__ void add_prod(double a, double b) { s = fma(a, b, s); }
  __device__ __forceinline__ void add(double x) { s += x; }
  __device__ __forceinline__ double value() const { return s; }
};

template <int D, bool FROM_NORMALS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(waves_for_dim(D), waves_for_dim(D))))
void walk_moments_fp64(const uint32_t* __restrict__ seeds, const double* __restrict__ zin, uint32_t rng_stream,
                       uint32_t T, uint32_t n_walks, double* __restrict__ acc, size_t ld) {
  using Acc = GramAcc;
  using Sh = WalkShape<D>;
  constexpr int L = Sh::L;
  constexpr int DP = Sh::DP;
  constexpr int R = DP / L;
  constexpr int H = DP / 2;
  constexpr int GPW = Sh::GPW;
  constexpr int S = Sh::S;
  constexpr int C = Sh::C;
  constexpr int IT = (C + L - 1) / L;
  constexpr int RW = 2 * DP;        // LDS row width per step
  constexpr int GP = GPW | 1;       // odd walk pitch: a wave's reads spread over the banks
  __shared__ double ws[S][RW][GP];

  const int lane = threadIdx.x;
  // walk fastest across the lanes (lane = r * GPW + g): the 16 lanes of a
  // 64-bit LDS access group read the same row of 16 different walks (odd walk
  // pitch: distinct banks), and the output stores of a lane row go to
  // consecutive walks.  One lane per walk: lane = walk.
  const int r = L == 1 ? 0 : lane / GPW;
  const int g = L == 1 ? lane : lane - r * GPW;
  const bool in_group = L == 1 || r < L;
  const int gs = in_group ? g : 0;
  const size_t walk = (size_t)blockIdx.x * GPW + gs;
  const bool valid = in_group && walk < n_walks;
  const size_t wc = walk < n_walks ? walk : (size_t)(n_walks - 1);
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS) seed = seeds[wc];
  const double* zw = FROM_NORMALS ? zin + wc * (size_t)T * D : nullptr;
  const int rot = r * R;
  const CallPos<D> pos0(r);  // normals of this lane's first Philox call in a round

  if constexpr (DP != D) {  // zero the padding columns once
    if (in_group)
      for (int s = 0; s < S; ++s)
        for (int i = D + r; i < DP; i += L) {
          ws[s][i][gs] = 0.0;
          ws[s][DP + i][gs] = 0.0;
        }
  }

  double b[DP];
  Acc geb[R][DP];
  Acc gbb[R][H + 1];
  Acc mom[6][R];
#pragma unroll
  for (int m = 0; m < DP; ++m) b[m] = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
#pragma unroll
    for (int m = 0; m < DP; ++m) geb[k][m].zero();
#pragma unroll
    for (int m = 0; m <= H; ++m) gbb[k][m].zero();
#pragma unroll
    for (int p = 0; p < 6; ++p) mom[p][k].zero();
  }

  // src/johansen_statistics.rs:70 and src/rng_matrix.rs:137: sqrt(1/steps)
  const double dt = 1.0 / (double)T;
  const double sqrt_dt = sqrt(dt);

  for (uint32_t t0 = 0; t0 < T; t0 += S) {
    // produce: this round's increments
    if (in_group) {
      const uint32_t q0 = (uint32_t)((uint64_t)t0 * D / 4);  // < 2^32: dim <= kMaxDim, steps <= kMaxSteps
      CallPos<D> pos = pos0;
#if JNE_FP64_UNROLL_CALLS
#pragma unroll
#else
#pragma unroll 1
#endif
      for (int it = 0; it < IT; ++it) {
        const int c = it * L + r;
        if (IT * L == C || c < C) {
          int sj[4], ij[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) pos.at(j, sj[j], ij[j]);
          double wv[4];
          if constexpr (!FROM_NORMALS) {
            float v[4];
            normals4(q0 + c, seed, rng_stream, v);
#pragma unroll
            for (int j = 0; j < 4; ++j) wv[j] = (double)v[j] * sqrt_dt;
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const uint32_t t = t0 + sj[j];
              wv[j] = t < T ? zw[(size_t)t * D + ij[j]] * sqrt_dt : 0.0;
            }
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            ws[sj[j]][ij[j]][gs] = wv[j];
            ws[sj[j]][DP + ij[j]][gs] = wv[j];
          }
        }
        pos.template advance<L>();
      }
    }
    __syncthreads();
    // consume: advance the walk S steps
    const int ns = (T - t0) < (uint32_t)S ? (int)(T - t0) : S;
#if JNE_FP64_UNROLL_STEPS
#pragma unroll
#else
#pragma unroll 1
#endif
    for (int s = 0; s < S; ++s) {
      if (s >= ns) break;
      const double* w = &ws[s][rot][gs];
      const uint32_t t = t0 + s;
      const double u = (double)(t + 1) * dt;
      const double u2 = u * u;
      // only the lane's own rows need B_{t+1} before the Gram update; the
      // other components are advanced in place afterwards
      double bnk[R], e[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        bnk[k] = b[k] + w[k * GP];  // B_{t+1} = B_t + z*sqrt(dt) (rng_matrix.rs:138-140)
        e[k] = bnk[k] - b[k];  // dbm = bm_current - bm_previous (johansen_statistics.rs:80-82)
      }
#pragma unroll
      for (int k = 0; k < R; ++k) {
#pragma unroll
        for (int m = 0; m < DP; ++m) geb[k][m].add_prod(e[k], b[m]);
#pragma unroll
        for (int dl = 0; dl <= H; ++dl) gbb[k][dl].add_prod(b[k], b[(k + dl) % DP]);
        mom[0][k].add(b[k]);
        mom[1][k].add_prod(u, b[k]);
        mom[2][k].add_prod(u2, b[k]);
        mom[3][k].add(e[k]);
        mom[4][k].add_prod(u, e[k]);
        mom[5][k].add_prod(u2, e[k]);
      }
#pragma unroll
      for (int k = 0; k < R; ++k) b[k] = bnk[k];
#pragma unroll
      for (int m = R; m < DP; ++m) b[m] = b[m] + w[m * GP];
    }
    __syncthreads();
  }

  if (!valid) return;
  // canonical layout, entry-major: [0, D^2) sum b b^T (full), [D^2, 2D^2)
  // sum e b^T (row e, col b), then 6 moment vectors of length D.
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int ci = (rot + k) % DP;
    if (ci >= D) continue;
#pragma unroll
    for (int m = 0; m < DP; ++m) {
      const int cj = (rot + m) % DP;
      if (cj < D) acc[(size_t)(D * D + ci * D + cj) * ld + walk] = geb[k][m].value();
    }
#pragma unroll
    for (int dl = 0; dl <= H; ++dl) {
      const int cj = (rot + k + dl) % DP;
      if (cj < D) {
        acc[(size_t)(ci * D + cj) * ld + walk] = gbb[k][dl].value();
        acc[(size_t)(cj * D + ci) * ld + walk] = gbb[k][dl].value();
      }
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) acc[(size_t)(2 * D * D + p * D + ci) * ld + walk] = mom[p][k].value();
  }
}

// ---------------------------------------------------------------------------
// Kernel 1, one lane per walk (the dims whose shape has L = 1): the same moments and the same
// arithmetic, step by step, as walk_moments_fp64 with L = 1, with the
// increments kept in registers instead of LDS.  A lane draws its walk's
// Philox calls in batches of NB calls (kFp64RegCalls: three independent
// calls in flight at dims 1-4 and 6, the five of a whole-step group at dim 5)
// and consumes the batch's G = 4 NB / dim steps from registers: no LDS, no
// barriers.
// ---------------------------------------------------------------------------
// target waves per SIMD of the register-only kernel by dim (the register cap:
// 512 / waves VGPRs per lane); it runs the dims whose shape has one lane per walk
#ifndef JNE_FP64_REG_WAVES
#define JNE_FP64_REG_WAVES 4, 4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2
#endif
constexpr int kFp64RegWaves[12] = {JNE_FP64_REG_WAVES};

// Philox calls per batch by dim (a multiple of the calls of one whole-step
// group: 1, 1, 3, 1, 5, 3 at dims 1..6)
#ifndef JNE_FP64_REG_CALLS
#define JNE_FP64_REG_CALLS 3, 3, 3, 3, 5, 3, 3, 3, 3, 3, 3, 3
#endif
constexpr int kFp64RegCalls[12] = {JNE_FP64_REG_CALLS};

template <int D>
struct RegBatch {
  static constexpr int NC = (4 / gcd_int(D, 4)) * D / 4;  // calls per whole-step group
  static constexpr int NB = kFp64RegCalls[D - 1];         // calls per batch
  static constexpr int G = 4 * NB / D;                     // steps per batch
  static_assert(NB % NC == 0, "a batch must be whole steps");
};

template <int D, bool FROM_NORMALS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kFp64RegWaves[D - 1], kFp64RegWaves[D - 1])))
void walk_moments_fp64_reg(const uint32_t* __restrict__ seeds, const double* __restrict__ zin, uint32_t rng_stream,
                           uint32_t T, uint32_t n_walks, double* __restrict__ acc, size_t ld) {
  using Acc = GramAcc;
  constexpr int H = D / 2;
  constexpr int NB = RegBatch<D>::NB, G = RegBatch<D>::G;
  const size_t walk = (size_t)blockIdx.x * 64 + threadIdx.x;
  const bool valid = walk < n_walks;
  const size_t wc = valid ? walk : (size_t)(n_walks - 1);
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS) seed = seeds[wc];
  const double* zw = FROM_NORMALS ? zin + wc * (size_t)T * D : nullptr;

  double b[D];
  Acc geb[D][D];
  Acc gbb[D][H + 1];
  Acc mom[6][D];
#pragma unroll
  for (int m = 0; m < D; ++m) b[m] = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
#pragma unroll
    for (int m = 0; m < D; ++m) geb[k][m].zero();
#pragma unroll
    for (int m = 0; m <= H; ++m) gbb[k][m].zero();
#pragma unroll
    for (int p = 0; p < 6; ++p) mom[p][k].zero();
  }

  // src/johansen_statistics.rs:70 and src/rng_matrix.rs:137: sqrt(1/steps)
  const double dt = 1.0 / (double)T;
  const double sqrt_dt = sqrt(dt);

  // one step of the walk from its scaled increments w[0..D) (the moment
  // updates of walk_moments_fp64 with one lane per walk)
  auto step = [&](const double* w, uint32_t t) {
    const double u = (double)(t + 1) * dt;
    const double u2 = u * u;
    double bn[D], e[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      bn[k] = b[k] + w[k];  // B_{t+1} = B_t + z*sqrt(dt) (rng_matrix.rs:138-140)
      e[k] = bn[k] - b[k];  // dbm = bm_current - bm_previous (johansen_statistics.rs:80-82)
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
#pragma unroll
      for (int m = 0; m < D; ++m) geb[k][m].add_prod(e[k], b[m]);
#pragma unroll
      for (int dl = 0; dl <= H; ++dl) gbb[k][dl].add_prod(b[k], b[(k + dl) % D]);
      mom[0][k].add(b[k]);
      mom[1][k].add_prod(u, b[k]);
      mom[2][k].add_prod(u2, b[k]);
      mom[3][k].add(e[k]);
      mom[4][k].add_prod(u, e[k]);
      mom[5][k].add_prod(u2, e[k]);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) b[k] = bn[k];
  };

  // a batch: G steps from t0; tail batches mask the steps at or beyond T
  auto batch = [&](uint32_t t0, auto tail_tag) {
    constexpr bool tail = decltype(tail_tag)::value;
    double w[G * D];
    if constexpr (!FROM_NORMALS) {
      const uint32_t q0 = (uint32_t)((uint64_t)t0 * D / 4);  // < 2^32: dim <= kMaxDim, steps <= kMaxSteps
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        float v[4];
        normals4(q0 + c, seed, rng_stream, v);
#pragma unroll
        for (int j = 0; j < 4; ++j) w[4 * c + j] = (double)v[j] * sqrt_dt;
      }
    } else {
#pragma unroll
      for (int n = 0; n < G * D; ++n) {
        const uint32_t t = t0 + (uint32_t)(n / D);
        w[n] = (!tail || t < T) ? zw[(size_t)t0 * D + n] * sqrt_dt : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < G; ++s) {
      if (tail && t0 + s >= T) break;
      step(&w[s * D], t0 + s);
    }
  };
  const uint32_t full = T / G;
  for (uint32_t q = 0; q < full; ++q) batch(q * G, BoolTag<false>{});
  if (full * G < T) batch(full * G, BoolTag<true>{});

  if (!valid) return;
  // canonical layout, entry-major: [0, D^2) sum b b^T (full), [D^2, 2D^2)
  // sum e b^T (row e, col b), then 6 moment vectors of length D.
#pragma unroll
  for (int k = 0; k < D; ++k) {
#pragma unroll
    for (int m = 0; m < D; ++m) acc[(size_t)(D * D + k * D + m) * ld + walk] = geb[k][m].value();
#pragma unroll
    for (int dl = 0; dl <= H; ++dl) {
      const int cj = (k + dl) % D;
      acc[(size_t)(k * D + cj) * ld + walk] = gbb[k][dl].value();
      acc[(size_t)(cj * D + k) * ld + walk] = gbb[k][dl].value();
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) acc[(size_t)(2 * D * D + p * D + k) * ld + walk] = mom[p][k].value();
  }
}

// ---------------------------------------------------------------------------
// Kernel 2: per-walk model algebra + generalized eigen solve (one thread per
// walk; split in two kernels where the fused one would fit too few waves).  F_t = X_t - P Z_t with P = (sum X Z^T)(sum Z Z^T)^-1 applied to the
// projected rows of X:
//   model 0: X = b,            no Z
//   model 1: X = [b; 1],       no Z
//   model 2: X = [b_0..b_{d-2}; y], Z = [1], y = u - 1/2 not projected
//   model 3: X = [b; y],       Z = [1], y not projected
//   model 4: X = [b_0..b_{d-2}; u^2], Z = [1; u], all rows projected
// S11 = dt (XX - sum_{projected} XZ ZZ^-1 ZX),  S01raw = eX - eZ ZZ^-1 ZX.
// The eigenvalues of S11^-1 S10 S01 are those of M = W^T W, W = L^-1 S10
// (S11 = L L^T), plus nF - dim structural zeros (models 1 and 3).
// ---------------------------------------------------------------------------
template <int N>
struct SymReg {
  // upper triangle (i <= j) of an N x N symmetric matrix in registers
  double v[N * (N + 1) / 2];
  __device__ __forceinline__ static constexpr int idx(int i, int j) {
    return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j);
  }
  __device__ __forceinline__ double& operator()(int i, int j) { return v[idx(i, j)]; }
};

// 1/x and 1/sqrt(x) in FP64: the hardware estimate and one Newton step
// (error a few ulp; a rotation needs orthogonality, not correct rounding)
__device__ __forceinline__ double rcp_nr(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  return r * fma(-0.5 * x * r, r, 1.5);
}

template <int N>
__device__ __forceinline__ void jacobi_rotate(SymReg<N>& a, int p, int q) {
  const double apq = a(p, q);
  const double app = a(p, p), aqq = a(q, q);
  // threshold Jacobi: below 1e-18 sqrt(|app aqq|), zeroing a_pq moves the
  // eigenvalues by at most ~1e-18 of themselves (a(p, q) = 0 is such a case)
  if (apq * apq <= 1e-36 * fabs(app * aqq)) {
    a(p, q) = 0.0;
    return;
  }
  // theta clamped so theta^2 stays finite (t ~ 1/(2 theta) there anyway)
  const double theta = fmin(fmax((aqq - app) * rcp_nr(2.0 * apq), -1e150), 1e150);
  const double h = fma(theta, theta, 1.0);
  double t = rcp_nr(fabs(theta) + h * rsq_nr(h));
  if (theta < 0.0) t = -t;
  const double c = rsq_nr(fma(t, t, 1.0));
  const double s = t * c;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k == p || k == q) continue;
    const double akp = a(k, p), akq = a(k, q);
    a(k, p) = c * akp - s * akq;
    a(k, q) = s * akp + c * akq;
  }
  a(p, p) = app - t * apq;
  a(q, q) = aqq + t * apq;
  a(p, q) = 0.0;
}

template <int N>
__device__ __forceinline__ void jacobi_eigvals(SymReg<N>& a) {
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dia = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      dia = fma(a(i, i), a(i, i), dia);
#pragma unroll
      for (int j = i + 1; j < N; ++j) off = fma(a(i, j), a(i, j), off);
    }
    if (!(off > 1e-34 * dia)) break;
#pragma unroll
    for (int p = 0; p < N - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) jacobi_rotate<N>(a, p, q);
    }
  }
}

template <int MODEL, int D>
struct EigenShape {
  static constexpr int NX = (MODEL == 1 || MODEL == 3) ? D + 1 : D;  // nF
  static constexpr int NB = (MODEL == 2 || MODEL == 4) ? D - 1 : D;  // b rows in X
  static constexpr int NZ = MODEL == 4 ? 2 : ((MODEL == 2 || MODEL == 3) ? 1 : 0);
  // one workgroup's W in LDS: NX*D doubles per walk.  Above 20 KiB fewer than
  // 8 workgroups (2 waves per SIMD) fit a CU, and the solve runs as two
  // kernels: the algebra up to M = W^T W, then Jacobi on M without LDS.
  static constexpr bool kSplit = NX * D * 64 * 8 > 20 * 1024;
};

// Model algebra, Cholesky of S11 and M = W^T W for walk wc (the prefix of the
// solve; W goes through the workgroup's LDS buffer wsh).
template <int MODEL, int D>
__device__ __forceinline__ void eigen_prep(const double* __restrict__ acc, size_t ld, uint32_t T, size_t wc,
                                           int tid, double (*wsh)[64], SymReg<D>& M, bool& pd_out) {
  constexpr int NX = EigenShape<MODEL, D>::NX;
  constexpr int NB = EigenShape<MODEL, D>::NB;
  constexpr int NZ = EigenShape<MODEL, D>::NZ;
  auto A = [&](int e) -> double { return acc[(size_t)e * ld + wc]; };
  auto GBB = [&](int i, int j) { return A(i * D + j); };
  auto GEB = [&](int k, int j) { return A(D * D + k * D + j); };
  auto MB = [&](int p, int i) { return A(2 * D * D + p * D + i); };
  auto ME = [&](int p, int k) { return A(2 * D * D + (3 + p) * D + k); };

  const double Tn = (double)T;
  const double dt = 1.0 / Tn;
  // closed forms of sum_t u_t^p, u_t = (t+1)/T, t = 0..T-1
  const double c1 = (Tn + 1.0) * 0.5;
  const double c2 = (Tn + 1.0) * (2.0 * Tn + 1.0) / (6.0 * Tn);
  const double c3 = (Tn + 1.0) * (Tn + 1.0) / (4.0 * Tn);
  const double c4 = (Tn + 1.0) * (2.0 * Tn + 1.0) * (3.0 * Tn * Tn + 3.0 * Tn - 1.0) / (30.0 * Tn * Tn * Tn);

  // XZ row x, column z (z = 0: 1, z = 1: u)
  auto XZ = [&](int x, int z) -> double {
    if (x < NB) return MB(z, x);
    if constexpr (MODEL == 2 || MODEL == 3) return c1 - 0.5 * Tn;  // sum y
    else if constexpr (MODEL == 4) return z == 0 ? c2 : c3;        // sum u^2, sum u^3
    else return 0.0;
  };
  auto projected = [&](int x) -> bool {
    if constexpr (MODEL == 4) return true;
    else if constexpr (MODEL == 2 || MODEL == 3) return x < NB;
    else return false;
  };
  auto XX = [&](int i, int j) -> double {
    if (i < NB && j < NB) return GBB(i, j);
    if (i >= NB && j >= NB) {
      if constexpr (MODEL == 1) return Tn;
      else if constexpr (MODEL == 2 || MODEL == 3) return c2 - c1 + 0.25 * Tn;  // sum (u - 1/2)^2
      else return c4;
    }
    const int bi = i < NB ? i : j;
    if constexpr (MODEL == 1) return MB(0, bi);
    else if constexpr (MODEL == 2 || MODEL == 3) return MB(1, bi) - 0.5 * MB(0, bi);  // sum b y
    else return MB(2, bi);                                                            // sum b u^2
  };
  auto EX = [&](int k, int j) -> double {
    if (j < NB) return GEB(k, j);
    if constexpr (MODEL == 1) return ME(0, k);
    else if constexpr (MODEL == 2 || MODEL == 3) return ME(1, k) - 0.5 * ME(0, k);
    else return ME(2, k);
  };

  // ZZ^-1 (closed form): model 4: [[T, c1], [c1, c2]]^-1, det = (T^2 - 1)/12
  double zi00 = 0.0, zi01 = 0.0, zi11 = 0.0;
  if constexpr (NZ == 1) zi00 = dt;
  if constexpr (NZ == 2) {
    const double det = (Tn * Tn - 1.0) / 12.0;
    zi00 = c2 / det;
    zi01 = -c1 / det;
    zi11 = Tn / det;
  }
  // P_x = XZ_x ZZ^-1 for every X row
  double P0[NX], P1[NX];
#pragma unroll
  for (int x = 0; x < NX; ++x) {
    P0[x] = 0.0;
    P1[x] = 0.0;
    if constexpr (NZ == 1) P0[x] = XZ(x, 0) * zi00;
    if constexpr (NZ == 2) {
      const double a0 = XZ(x, 0), a1 = XZ(x, 1);
      P0[x] = a0 * zi00 + a1 * zi01;
      P1[x] = a0 * zi01 + a1 * zi11;
    }
  }
  auto proj = [&](int i, int j) -> double {  // XZ_i ZZ^-1 ZX_j with projection mask
    if constexpr (NZ == 0) return 0.0;
    else {
      if (projected(i)) return NZ == 2 ? P0[i] * XZ(j, 0) + P1[i] * XZ(j, 1) : P0[i] * XZ(j, 0);
      if (projected(j)) return NZ == 2 ? P0[j] * XZ(i, 0) + P1[j] * XZ(i, 1) : P0[j] * XZ(i, 0);
      return 0.0;
    }
  };

  // S11 (lower triangle in SymReg form) and its Cholesky factor in place;
  // dinv[j] = 1 / L(j, j), which the substitutions below multiply by
  SymReg<NX> Lf;
  double dinv[NX];
  bool pd = true;
#pragma unroll
  for (int i = 0; i < NX; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) Lf(i, j) = (XX(i, j) - proj(i, j)) * dt;
  }
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    double s = Lf(j, j);
#pragma unroll
    for (int k = 0; k < j; ++k) s -= Lf(j, k) * Lf(j, k);
    pd = pd && (s > 0.0);
    const double ljj = sqrt(s > 0.0 ? s : 1.0);
    const double inv = 1.0 / ljj;
    Lf(j, j) = ljj;
    dinv[j] = inv;
#pragma unroll
    for (int i = j + 1; i < NX; ++i) {
      double v = Lf(i, j);
#pragma unroll
      for (int k = 0; k < j; ++k) v -= Lf(i, k) * Lf(j, k);
      Lf(i, j) = v * inv;
    }
  }

  // W column k = L^-1 (row k of S01raw), stored to LDS
#pragma unroll 1
  for (int k = 0; k < D; ++k) {
    double wv[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      double v = EX(k, i);
      if constexpr (NZ > 0) {
        if (projected(i)) {
          const double ez0 = ME(0, k);
          v -= NZ == 2 ? P0[i] * ez0 + P1[i] * ME(1, k) : P0[i] * ez0;
        }
      }
#pragma unroll
      for (int j = 0; j < i; ++j) v -= Lf(i, j) * wv[j];
      wv[i] = v * dinv[i];
      wsh[i * D + k][tid] = wv[i];
    }
  }

  // M = W^T W (D x D)
#pragma unroll
  for (int k = 0; k < D * (D + 1) / 2; ++k) M.v[k] = 0.0;
#pragma unroll 1
  for (int i = 0; i < NX; ++i) {
    double row[D];
#pragma unroll
    for (int k = 0; k < D; ++k) row[k] = wsh[i * D + k][tid];
#pragma unroll
    for (int k = 0; k < D; ++k) {
#pragma unroll
      for (int l = k; l < D; ++l) M(k, l) = fma(row[k], row[l], M(k, l));
    }
  }
  pd_out = pd;
}

// sqrt, 1/x and 1/sqrt(x) for jne_tridiag.h: the hardware estimates and one
// Newton step
struct TriMath {
  __device__ __forceinline__ static double sqrt(double x) { return x > 0.0 ? x * rsq_nr(x) : 0.0; }
  __device__ __forceinline__ static double rcp(double x) { return rcp_nr(x); }
  __device__ __forceinline__ static double rsqrt(double x) { return rsq_nr(x); }
};

// The eigenvalues of M (tridiagonal reduction + implicit QL from dim
// JNE_EIGEN_QL_MIN_DIM, cyclic Jacobi below it), |lambda| sorted descending,
// NX - D structural zeros; NaN rows for a walk whose S11 was not positive
// definite.
template <int D, int NX>
__device__ __forceinline__ void eigen_finish(SymReg<D>& M, bool pd, double* __restrict__ o) {
  double ev[D];
  if constexpr (D >= JNE_EIGEN_QL_MIN_DIM) {
    tridiag_eigvals<D, TriMath>(M.v, ev);
#pragma unroll
    for (int k = 0; k < D; ++k) ev[k] = fabs(ev[k]);
  } else {
    jacobi_eigvals<D>(M);
#pragma unroll
    for (int k = 0; k < D; ++k) ev[k] = fabs(M(k, k));
  }
  // sort descending (odd-even transposition, fixed network)
#pragma unroll
  for (int pass = 0; pass < D; ++pass) {
#pragma unroll
    for (int k = pass & 1; k + 1 < D; k += 2) {
      const double hi = fmax(ev[k], ev[k + 1]);
      const double lo = fmin(ev[k], ev[k + 1]);
      ev[k] = hi;
      ev[k + 1] = lo;
    }
  }
  const double nan = __longlong_as_double(0x7FF8000000000000LL);
#pragma unroll
  for (int k = 0; k < D; ++k) o[k] = pd ? ev[k] : nan;
  if constexpr (NX > D) o[D] = pd ? 0.0 : nan;
}

// One thread per walk: the whole solve in one kernel (small D).
template <int MODEL, int D>
__global__ __launch_bounds__(64) void walk_eigen_fp64(const double* __restrict__ acc, size_t ld, uint32_t T,
                                                      uint32_t n_walks, double* __restrict__ out) {
  constexpr int NX = EigenShape<MODEL, D>::NX;
  __shared__ double wsh[NX * D][64];
  const int tid = threadIdx.x;
  const size_t w = (size_t)blockIdx.x * 64 + tid;
  const bool valid = w < n_walks;
  const size_t wc = valid ? w : (size_t)(n_walks - 1);
  SymReg<D> M;
  bool pd;
  eigen_prep<MODEL, D>(acc, ld, T, wc, tid, wsh, M, pd);
  if (!valid) return;
  eigen_finish<D, NX>(M, pd, out + w * NX);
}

// Split form, first kernel: M and the definiteness flag, entry-major
// (m[e * ld + walk], e < D(D+1)/2; e = D(D+1)/2 holds 1 or 0).
template <int MODEL, int D>
__global__ __launch_bounds__(64) void walk_eigen_prep(const double* __restrict__ acc, size_t ld, uint32_t T,
                                                      uint32_t n_walks, double* __restrict__ m) {
  constexpr int NX = EigenShape<MODEL, D>::NX;
  __shared__ double wsh[NX * D][64];
  const int tid = threadIdx.x;
  const size_t w = (size_t)blockIdx.x * 64 + tid;
  const bool valid = w < n_walks;
  const size_t wc = valid ? w : (size_t)(n_walks - 1);
  SymReg<D> M;
  bool pd;
  eigen_prep<MODEL, D>(acc, ld, T, wc, tid, wsh, M, pd);
  if (!valid) return;
#pragma unroll
  for (int e = 0; e < D * (D + 1) / 2; ++e) m[(size_t)e * ld + w] = M.v[e];
  m[(size_t)(D * (D + 1) / 2) * ld + w] = pd ? 1.0 : 0.0;
}

// Split form, second kernel: the eigenvalues of M in registers, no LDS
// (eigen_finish: tridiagonal QL from dim JNE_EIGEN_QL_MIN_DIM, Jacobi below).
template <int D, int NX>
__global__ __launch_bounds__(64) void walk_eigen_jacobi(const double* __restrict__ m, size_t ld, uint32_t n_walks,
                                                        double* __restrict__ out) {
  const size_t w = (size_t)blockIdx.x * 64 + threadIdx.x;
  if (w >= n_walks) return;
  SymReg<D> M;
#pragma unroll
  for (int e = 0; e < D * (D + 1) / 2; ++e) M.v[e] = m[(size_t)e * ld + w];
  const bool pd = m[(size_t)(D * (D + 1) / 2) * ld + w] != 0.0;
  eigen_finish<D, NX>(M, pd, out + w * NX);
}


// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int D>
static hipError_t launch_moments_d(const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                   uint32_t n, double* acc, size_t ld, hipStream_t st) {
  constexpr int GPW = WalkShape<D>::GPW;
  const unsigned grid = (unsigned)((n + GPW - 1) / GPW);
#if JNE_FP64_REG
  if constexpr (WalkShape<D>::L == 1) {
    if (zin != nullptr)
      hipLaunchKernelGGL((walk_moments_fp64_reg<D, true>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream, T, n,
                         acc, ld);
    else
      hipLaunchKernelGGL((walk_moments_fp64_reg<D, false>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream, T, n,
                         acc, ld);
    return hipGetLastError();
  }
#endif
  if (zin != nullptr)
    hipLaunchKernelGGL((walk_moments_fp64<D, true>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream, T, n, acc,
                       ld);
  else
    hipLaunchKernelGGL((walk_moments_fp64<D, false>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream, T, n,
                       acc, ld);
  return hipGetLastError();
}

template <int MODEL, int D>
static hipError_t launch_eigen_md(const double* acc, size_t ld, uint32_t T, uint32_t n, double* out, double* mbuf,
                                  hipStream_t st) {
  const unsigned grid = (unsigned)((n + 63) / 64);
  if constexpr (EigenShape<MODEL, D>::kSplit) {
    if (mbuf == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL((walk_eigen_prep<MODEL, D>), dim3(grid), dim3(64), 0, st, acc, ld, T, n, mbuf);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((walk_eigen_jacobi<D, EigenShape<MODEL, D>::NX>), dim3(grid), dim3(64), 0, st, mbuf, ld, n,
                       out);
  } else {
    hipLaunchKernelGGL((walk_eigen_fp64<MODEL, D>), dim3(grid), dim3(64), 0, st, acc, ld, T, n, out);
  }
  return hipGetLastError();
}

hipError_t launch_walk_moments_fp64(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                    uint32_t n, double* acc, size_t ld, hipStream_t st) {
  switch (d) {
#define JNE_CASE(DD) \
  case DD:           \
    return launch_moments_d<DD>(seeds, zin, rng_stream, T, n, acc, ld, st);
    JNE_CASE(1) JNE_CASE(2) JNE_CASE(3) JNE_CASE(4) JNE_CASE(5) JNE_CASE(6)
    JNE_CASE(7) JNE_CASE(8) JNE_CASE(9) JNE_CASE(10) JNE_CASE(11) JNE_CASE(12)
#undef JNE_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <int MODEL>
static hipError_t launch_eigen_m(int d, const double* acc, size_t ld, uint32_t T, uint32_t n, double* out,
                                 double* mbuf, hipStream_t st) {
  switch (d) {
#define JNE_CASE(DD) \
  case DD:           \
    return launch_eigen_md<MODEL, DD>(acc, ld, T, n, out, mbuf, st);
    JNE_CASE(1) JNE_CASE(2) JNE_CASE(3) JNE_CASE(4) JNE_CASE(5) JNE_CASE(6)
    JNE_CASE(7) JNE_CASE(8) JNE_CASE(9) JNE_CASE(10) JNE_CASE(11) JNE_CASE(12)
#undef JNE_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_walk_eigen_fp64(int model, int d, const double* acc, size_t ld, uint32_t T, uint32_t n,
                                  double* out, double* mbuf, hipStream_t st) {
  switch (model) {
    case 0: return launch_eigen_m<0>(d, acc, ld, T, n, out, mbuf, st);
    case 1: return launch_eigen_m<1>(d, acc, ld, T, n, out, mbuf, st);
    case 2: return launch_eigen_m<2>(d, acc, ld, T, n, out, mbuf, st);
    case 3: return launch_eigen_m<3>(d, acc, ld, T, n, out, mbuf, st);
    case 4: return launch_eigen_m<4>(d, acc, ld, T, n, out, mbuf, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace jne
