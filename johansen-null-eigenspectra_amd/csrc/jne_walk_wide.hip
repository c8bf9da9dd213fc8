// jne_walk_wide.hip -- FP64 walk path for dims 13..32 (runtime dim).
//
// The reference's per-walk operator takes any dim
// (calculate_eigenvalues, src/johansen_statistics.rs:59-85; the CLI checks
// only dim > 0, src/cli.rs:201-225).  The register-resident kernels of
// jne_walk_fp64.hip cover dims 1..12; above that a walk's Gram and its
// eigenproblem no longer fit one lane's registers, so here one wave owns one
// walk and keeps both in registers spread over its lanes and in LDS:
//
//   walk_moments_wide  the same one-pass moments as walk_moments_fp64 (the
//                      canonical entry-major layout of jne_internal.h): per
//                      round of S steps the wave draws the S*dim normals
//                      (Philox calls dealt over the lanes, the stream of
//                      jne_rng.h), lanes 0..dim-1 run the FP64 recurrence
//                      B_{t+1} = B_t + z sqrt(dt) of their component, and every
//                      lane adds its tile of sum [b; e] b^T (components padded
//                      to 16, 24 or 32) and one row of the moment vectors,
//                      FP64 fma;
//   walk_eigen_wide    model algebra (S11 and S01 as Schur complements on the
//                      moments), Cholesky of S11, W = L^-1 S10, M = W^T W,
//                      with lanes working on matrix entries, rows or columns
//                      in parallel; then the Householder reduction of M to
//                      tridiagonal form (lane i holding row i in registers)
//                      and eigenvalue k by a group of lanes, by multisection
//                      on Sturm counts (jne_tridiag.h).
//
// The arithmetic is FP64 throughout, as in walk_moments_fp64 /
// walk_eigen_fp64, so the results meet the same 1e-9 parity with the
// reference's dggev solve.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "jne_internal.h"
#include "jne_rng.h"
#include "jne_tridiag.h"

namespace jne {

namespace {

constexpr int kPadDim = kMaxWideDim;            // components padded to 32
constexpr int kRoundMax = 16;                   // most steps in one round (dim 13..16)
static_assert(kPadDim + kRoundMax <= 64, "lanes past the components form the steps' u and u^2");

// Steps per round: a multiple of 4 (whole Philox calls for any dim) with at
// most 64 calls, so each lane draws at most one call per round.
__host__ __device__ constexpr int wide_round_steps(int d) { return ((256 / d) & ~3) < 4 ? 4 : ((256 / d) & ~3); }

}  // namespace

// One wave per walk (blockIdx.x).  The components are padded to DP = 16, 24
// or 32 (dims 13-16, 17-24, 25-32), and lane l owns an RB x CB tile of the
// stacked rows y = [b (DP, zero-padded); e (DP)] against the columns b:
// rows RB (l / NCB) .. + RB - 1, columns CB (l % NCB) .. + CB - 1, so that the
// 64 tiles cover sum y b^T = [sum b b^T; sum e b^T] (DP = 16: 2 x 4 tiles,
// 24: 3 x 6, 32: 8 x 4); lanes l < 2 DP also add sum y_l, sum u y_l and
// sum u^2 y_l.
template <int DP>
struct WideTile {
  static constexpr int CB = DP == 24 ? 6 : 4;             // columns per lane
  static constexpr int NCB = DP / CB;                     // column blocks
  static constexpr int RB = 2 * DP / (64 / NCB);          // rows per lane
  static_assert((2 * DP / RB) * NCB == 64 && RB * (64 / NCB) == 2 * DP, "64 tiles");
};

template <int DP, bool FROM_NORMALS>
__global__ __launch_bounds__(64) void walk_moments_wide(const uint32_t* __restrict__ seeds,
                                                        const double* __restrict__ zin, uint32_t rng_stream,
                                                        uint32_t T, uint32_t n_walks, int D,
                                                        double* __restrict__ acc, size_t ld) {
  using Tl = WideTile<DP>;
  constexpr int RB = Tl::RB, CB = Tl::CB, NCB = Tl::NCB;
  __shared__ __align__(16) double ys[kRoundMax][2 * DP];  // per step of the round: b, then e
  __shared__ double wz[kRoundMax * kPadDim];              // the round's increments z sqrt(dt), t-major
  // u_t and u_t^2 of the round's steps, formed once per round by spare lanes
  // at DP = 32 (3% faster there; at DP = 16 and 24, whose steps carry fewer
  // FMAs, the extra LDS read per step costs more than the 3 VALU it saves)
  constexpr bool kUTable = DP == 32;
  __shared__ __align__(16) double us[kUTable ? kRoundMax : 1][2];
  const int l = threadIdx.x;
  const size_t walk = blockIdx.x;
  if (walk >= n_walks) return;  // whole wave exits together
  const int S = wide_round_steps(D);
  const int calls = S * D / 4;
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS) seed = seeds[walk];
  const double* zw = FROM_NORMALS ? zin + walk * (size_t)T * D : nullptr;

  // padding rows and columns stay zero
  for (int e = l; e < kRoundMax * 2 * DP; e += 64) (&ys[0][0])[e] = 0.0;
  __syncthreads();

  const int rb = l / NCB, cb = l % NCB;
  double g[RB][CB];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) g[r][c] = 0.0;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;  // sum y_l, sum u y_l, sum u^2 y_l (lanes l < 2 DP)
  const int lm = l < 2 * DP ? l : 0;

  // src/johansen_statistics.rs:70 and src/rng_matrix.rs:137: sqrt(1/steps)
  const double dt = 1.0 / (double)T;
  const double sqrt_dt = sqrt(dt);
  double b = 0.0;  // B_t of component l (lanes l < D)

  for (uint32_t t0 = 0; t0 < T; t0 += (uint32_t)S) {
    const int ns = (T - t0) < (uint32_t)S ? (int)(T - t0) : S;
    // this round's increments: call l draws normals 4l..4l+3 of the round
    if (l < calls) {
      float v[4];
      if constexpr (!FROM_NORMALS) normals4((uint64_t)t0 * D / 4 + l, seed, rng_stream, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = 4 * l + j;  // = s*D + i
        double w;
        if constexpr (FROM_NORMALS) {
          const uint32_t t = t0 + (uint32_t)(q / D);
          w = t < T ? zw[(size_t)t0 * D + q] * sqrt_dt : 0.0;
        } else {
          w = (double)v[j] * sqrt_dt;
        }
        wz[q] = w;
      }
    }
    __syncthreads();
    // the recurrence of component l, with the reference's rounding:
    // B_{t+1} = B_t + z sqrt(dt) (rng_matrix.rs:138-140), e = B_{t+1} - B_t
    // (johansen_statistics.rs:80-82)
    if (l < D) {
      for (int s = 0; s < ns; ++s) {
        const double bn = b + wz[s * D + l];
        ys[s][l] = b;
        ys[s][DP + l] = bn - b;
        b = bn;
      }
    } else if (kUTable && l >= kPadDim && l - kPadDim < ns) {
      // lanes 32.. (never a component's): u_t = (t+1)/T and u_t^2 of step
      // l - 32, the same for every lane of the moment sums below
      const int s = l - kPadDim;
      const double u = (double)(t0 + (uint32_t)s + 1u) * dt;
      us[s][0] = u;
      us[s][1] = u * u;
    }
    __syncthreads();
    for (int s = 0; s < ns; ++s) {
      double y[RB], bb[CB];
#pragma unroll
      for (int k = 0; k < RB; ++k) y[k] = ys[s][RB * rb + k];
#pragma unroll
      for (int k = 0; k < CB; ++k) bb[k] = ys[s][CB * cb + k];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c) g[r][c] = fma(y[r], bb[c], g[r][c]);
      double u, u2;
      if constexpr (kUTable) {
        u = us[s][0];
        u2 = us[s][1];
      } else {
        u = (double)(t0 + (uint32_t)s + 1u) * dt;
        u2 = u * u;
      }
      const double yl = ys[s][lm];
      m0 += yl;
      m1 = fma(u, yl, m1);
      m2 = fma(u2, yl, m2);
    }
    __syncthreads();
  }

  // canonical layout, entry-major: [0, D^2) sum b b^T, [D^2, 2D^2) sum e b^T
  // (row e, column b), then sum b, sum u b, sum u^2 b, sum e, sum u e, sum u^2 e
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const int row = RB * rb + r;
    const bool is_e = row >= DP;
    const int i = is_e ? row - DP : row;
    if (i >= D) continue;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int j = CB * cb + c;
      if (j < D) acc[(size_t)((is_e ? D * D : 0) + i * D + j) * ld + walk] = g[r][c];
    }
  }
  if (l < 2 * DP) {
    const bool is_e = l >= DP;
    const int i = is_e ? l - DP : l;
    if (i < D) {
      const int base = 2 * D * D + (is_e ? 3 * D : 0) + i;
      acc[(size_t)base * ld + walk] = m0;
      acc[(size_t)(base + D) * ld + walk] = m1;
      acc[(size_t)(base + 2 * D) * ld + walk] = m2;
    }
  }
}

namespace {

__device__ __forceinline__ double rcp_nr_w(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

struct WideMath {
  static __device__ __forceinline__ double rcp(double x) { return rcp_nr_w(x); }
};

}  // namespace

// A row-major matrix in LDS with a runtime pitch: m[i][j].
struct LdsMat {
  double* p;
  int pitch;
  __device__ __forceinline__ double* operator[](int i) const { return p + i * pitch; }
};

// LDS bytes of walk_eigen_wide for (model, dim): L (NX x (NX + 1)), W (NX x
// (D + 1)) and P0, P1 (NX each), NX = D + 1 for models 1 and 3, else D.
size_t wide_eigen_lds_bytes(int model, int D) {
  const size_t nx = (model == 1 || model == 3) ? (size_t)D + 1 : (size_t)D;
  return sizeof(double) * (nx * (nx + 1) + nx * ((size_t)D + 1) + 2 * nx);
}

// One wave per walk: the model algebra of walk_eigen_fp64 (F = X - P Z with
// P = (sum X Z^T)(sum Z Z^T)^-1 on the projected rows; models 0..4 as in
// construct_f_matrix, src/johansen_statistics.rs:102-197), then the
// generalized eigenproblem of src/johansen_statistics.rs:24-47 as the
// symmetric one of M = W^T W, W = L^-1 S10, S11 = L L^T.
template <int DP>
__global__ __launch_bounds__(64) void walk_eigen_wide(int model, int D, const double* __restrict__ acc, size_t ld,
                                                      uint32_t T, uint32_t n_walks, double* __restrict__ out) {
  // LDS sized by the launch for this (model, dim) (wide_eigen_lds_bytes):
  // S11 and then its Cholesky factor L (NX x NX, pitch NX + 1); S10 and then
  // W (NX x D, pitch D + 1); the projections P0, P1.  Once W is formed the
  // bytes of L hold M (D x D, pitch D + 1), which needs only W.
  extern __shared__ double wide_lds[];
  const int NX = (model == 1 || model == 3) ? D + 1 : D;
  const LdsMat Ls{wide_lds, NX + 1};
  const LdsMat Ws{wide_lds + NX * (NX + 1), D + 1};
  const LdsMat Ms{wide_lds, D + 1};
  double* const P0 = wide_lds + NX * (NX + 1) + NX * (D + 1);
  double* const P1 = P0 + NX;
  const int l = threadIdx.x;
  const size_t walk = blockIdx.x;
  if (walk >= n_walks) return;
  const int NB = (model == 2 || model == 4) ? D - 1 : D;
  const int NZ = model == 4 ? 2 : ((model == 2 || model == 3) ? 1 : 0);
  auto A = [&](int e) -> double { return acc[(size_t)e * ld + walk]; };
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
  auto XZ = [&](int x, int z) -> double {
    if (x < NB) return MB(z, x);
    if (model == 2 || model == 3) return c1 - 0.5 * Tn;  // sum y, y = u - 1/2
    if (model == 4) return z == 0 ? c2 : c3;             // sum u^2, sum u^3
    return 0.0;
  };
  auto projected = [&](int x) -> bool { return model == 4 || ((model == 2 || model == 3) && x < NB); };
  auto XX = [&](int i, int j) -> double {
    if (i < NB && j < NB) return GBB(i, j);
    if (i >= NB && j >= NB) {
      if (model == 1) return Tn;
      if (model == 2 || model == 3) return c2 - c1 + 0.25 * Tn;  // sum (u - 1/2)^2
      return c4;
    }
    const int bi = i < NB ? i : j;
    if (model == 1) return MB(0, bi);
    if (model == 2 || model == 3) return MB(1, bi) - 0.5 * MB(0, bi);  // sum b y
    return MB(2, bi);                                                  // sum b u^2
  };
  auto EX = [&](int k, int j) -> double {
    if (j < NB) return GEB(k, j);
    if (model == 1) return ME(0, k);
    if (model == 2 || model == 3) return ME(1, k) - 0.5 * ME(0, k);
    return ME(2, k);
  };
  // ZZ^-1 in closed form: model 4: [[T, c1], [c1, c2]]^-1, det = (T^2 - 1)/12
  double zi00 = 0.0, zi01 = 0.0, zi11 = 0.0;
  if (NZ == 1) zi00 = dt;
  if (NZ == 2) {
    const double det = (Tn * Tn - 1.0) / 12.0;
    zi00 = c2 / det;
    zi01 = -c1 / det;
    zi11 = Tn / det;
  }
  if (l < NX) {
    double p0 = 0.0, p1 = 0.0;
    if (NZ == 1) p0 = XZ(l, 0) * zi00;
    if (NZ == 2) {
      const double a0 = XZ(l, 0), a1 = XZ(l, 1);
      p0 = a0 * zi00 + a1 * zi01;
      p1 = a0 * zi01 + a1 * zi11;
    }
    P0[l] = p0;
    P1[l] = p1;
  }
  __syncthreads();
  auto proj = [&](int i, int j) -> double {  // XZ_i ZZ^-1 ZX_j on the projected rows
    if (NZ == 0) return 0.0;
    if (projected(i)) return NZ == 2 ? P0[i] * XZ(j, 0) + P1[i] * XZ(j, 1) : P0[i] * XZ(j, 0);
    if (projected(j)) return NZ == 2 ? P0[j] * XZ(i, 0) + P1[j] * XZ(i, 1) : P0[j] * XZ(i, 0);
    return 0.0;
  };
  // S11 (lower triangle) and S10 = S01raw^T: S01raw = eX - eZ ZZ^-1 ZX
  for (int e = l; e < NX * NX; e += 64) {
    const int i = e / NX, j = e - (e / NX) * NX;
    if (j <= i) Ls[i][j] = (XX(i, j) - proj(i, j)) * dt;
  }
  for (int e = l; e < NX * D; e += 64) {
    const int i = e / D, k = e - (e / D) * D;
    double v = EX(k, i);
    if (NZ > 0 && projected(i)) v -= NZ == 2 ? P0[i] * ME(0, k) + P1[i] * ME(1, k) : P0[i] * ME(0, k);
    Ws[i][k] = v;
  }
  __syncthreads();

  // Cholesky S11 = L L^T, one column at a time: every lane forms the pivot
  // (the same value), lanes below it their entry of the column.  The
  // reciprocal pivots go to P0's bytes (P0 and P1 are not read again) for the
  // substitution below.
  double* const dinv = P0;
  bool pd = true;
  for (int j = 0; j < NX; ++j) {
    double s = Ls[j][j];
    for (int k = 0; k < j; ++k) s -= Ls[j][k] * Ls[j][k];
    pd = pd && (s > 0.0);
    const double ljj = sqrt(s > 0.0 ? s : 1.0);
    const double inv = 1.0 / ljj;
    const int i = j + 1 + l;
    double v = 0.0;
    if (i < NX) {
      v = Ls[i][j];
      for (int k = 0; k < j; ++k) v -= Ls[i][k] * Ls[j][k];
    }
    __syncthreads();
    if (l == 0) {
      Ls[j][j] = ljj;
      dinv[j] = inv;
    }
    if (i < NX) Ls[i][j] = v * inv;
    __syncthreads();
  }

  // W = L^-1 S10: lane k solves column k by forward substitution, in place
  if (l < D) {
    for (int i = 0; i < NX; ++i) {
      double v = Ws[i][l];
      for (int j = 0; j < i; ++j) v -= Ls[i][j] * Ws[j][l];
      Ws[i][l] = v * dinv[i];
    }
  }
  __syncthreads();
  // M = W^T W (D x D, symmetric: both triangles)
  for (int e = l; e < D * D; e += 64) {
    const int k = e / D, m = e - (e / D) * D;
    if (m < k) continue;
    double v = 0.0;
    for (int i = 0; i < NX; ++i) v = fma(Ws[i][k], Ws[i][m], v);
    Ms[k][m] = v;
    Ms[m][k] = v;
  }
  __syncthreads();

  // ---- Householder reduction of M to tridiagonal form: lane i keeps row i
  // of M in registers (columns padded to DP with zeros).  Step k broadcasts
  // column k below the diagonal (x) through LDS, every lane forms the same
  // reflector v = x - alpha e_{k+1}, tau = 2 / |v|^2 from it, lane i its entry
  // of p = tau M v, which is broadcast in turn; then w = p - (tau/2)(p.v) v and
  // lane i updates its row, M -= v w^T + w v^T on the trailing block.  d_k is
  // final once step k starts and e_k = alpha.  The scratch vectors (5 DP
  // doubles) use the bytes of W and P0, P1, at least D (D + 3) >= 5 DP doubles
  // since DP <= D + 7.
  double* const Xs = wide_lds + NX * (NX + 1);
  double* const Ps = Xs + DP;
  double* const Dg = Ps + DP;
  double* const Eg = Dg + DP;
  double* const EVs = Eg + DP;
  double row[DP];
  if (l < D) {
#pragma unroll
    for (int j = 0; j < DP; ++j) row[j] = j < D ? Ms[l][j < D ? j : 0] : 0.0;
  } else {
#pragma unroll
    for (int j = 0; j < DP; ++j) row[j] = 0.0;
  }
  if (l >= D && l < DP) {  // padding entries of the broadcast vectors stay 0
    Xs[l] = 0.0;
    Ps[l] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < DP - 2; ++k) {
    if (k < D - 2) {
      const double xi = l > k ? row[k] : 0.0;
      if (l > k && l < D) Xs[l] = xi;
      __syncthreads();
      double v[DP];
      double sigma = 0.0;
#pragma unroll
      for (int j = k + 2; j < DP; ++j) {
        v[j] = Xs[j];
        sigma = fma(v[j], v[j], sigma);
      }
      const double x0 = Xs[k + 1];
      const double nrm = sqrt(fma(x0, x0, sigma));
      const double alpha = x0 > 0.0 ? -nrm : nrm;
      const double v0 = x0 - alpha;
      const double vn = fma(v0, v0, sigma);
      // tau = 0 when x = 0: the reflection is the identity, e_k = 0
      const double tau = vn > 0.0 ? 2.0 / vn : 0.0;
      v[k + 1] = v0;
      const double vi = l == k + 1 ? v0 : xi;
      double s = 0.0;
#pragma unroll
      for (int j = k + 1; j < DP; ++j) s = fma(row[j], v[j], s);
      const double pi = l > k ? tau * s : 0.0;
      if (l > k && l < D) Ps[l] = pi;
      __syncthreads();
      double pj[DP];
      double pv = 0.0;
#pragma unroll
      for (int j = k + 1; j < DP; ++j) {
        pj[j] = Ps[j];
        pv = fma(pj[j], v[j], pv);
      }
      const double K = 0.5 * tau * pv;
      const double wi = fma(-K, vi, pi);
#pragma unroll
      for (int j = k + 1; j < DP; ++j) {
        const double wj = fma(-K, v[j], pj[j]);
        row[j] = fma(-vi, wj, fma(-wi, v[j], row[j]));
      }
      if (l == 0) Eg[k] = tau > 0.0 ? alpha : x0;
      if (l == k) Dg[k] = row[k];
    }
  }
  // the last 2 x 2 block: d_{D-2}, d_{D-1} and e_{D-2} = M[D-1][D-2]
  double dii = 0.0, sub = 0.0;
#pragma unroll
  for (int j = 0; j < DP; ++j) {
    if (l == j) dii = row[j];
    if (l == j + 1) sub = row[j];
  }
  if (l == D - 2 || l == D - 1) Dg[l] = dii;
  if (l == D - 1) Eg[D - 2] = sub;
  __syncthreads();

  // ---- eigenvalue g (ascending) of the tridiagonal matrix by the P = 64 / D
  // lanes g, g + D, .., g + (P - 1) D: from the Gershgorin interval, each round
  // lane g + jD takes the Sturm count at the (j+1)-th of P equally spaced
  // interior points, the counts meet in LDS, and every lane of the group keeps
  // the sub-interval whose upper point is the first count past g: (P + 1)-fold
  // per round, where bisection halves (P = 2 at D = 22..32, 4 at D <= 16).
  double dg[DP], e2[DP];
#pragma unroll
  for (int j = 0; j < DP; ++j) {
    dg[j] = j < D ? Dg[j < D ? j : 0] : 0.0;
    const double e = j < D - 1 ? Eg[j < D - 1 ? j : 0] : 0.0;
    e2[j] = e * e;
  }
  double lo, hi, pivmin;
  sturm_bounds<DP>(dg, e2, D, lo, hi, pivmin);
  const int P = 64 / D;  // >= 2 (D <= 32)
  const int grp = l % D, jp = l / D;  // lanes with jp >= P follow their group
  const double inv_sections = 1.0 / (double)(P + 1);
  int rounds = 0;  // (P + 1)^rounds >= 2^54: at most the width 54 halvings leave
  for (double w = 1.0; w > 0x1p-54; w *= inv_sections) ++rounds;
  int* const cnt = reinterpret_cast<int*>(Xs);  // 64 ints over Xs, Ps (not read again)
  for (int it = 0; it < rounds; ++it) {
    const double step = (hi - lo) * inv_sections;
    int c = 0;
    if (jp < P) c = sturm_count<DP, WideMath>(dg, e2, D, lo + (double)(jp + 1) * step, pivmin);
    cnt[l] = c;
    __syncthreads();
    double nlo = lo, nhi = hi;
    bool found = false;
    for (int q = 0; q < P; ++q) {
      const double xq = lo + (double)(q + 1) * step;
      if (!found) {
        if (cnt[grp + D * q] > grp) {
          nhi = xq;
          found = true;
        } else {
          nlo = xq;
        }
      }
    }
    lo = nlo;
    hi = nhi;
    __syncthreads();  // every lane has read the counts before the next round writes them
  }
  const double lam = 0.5 * (lo + hi);

  // |lambda| sorted descending (rank of each by lane), NX - D structural zeros
  if (l < D) EVs[l] = fabs(lam);
  __syncthreads();
  const double nan = __longlong_as_double(0x7FF8000000000000LL);
  double* o = out + walk * (size_t)NX;
  if (l < D) {
    const double ev = EVs[l];
    int rank = 0;
    for (int j = 0; j < D; ++j) {
      const double x = EVs[j];
      rank += (x > ev || (x == ev && j < l)) ? 1 : 0;
    }
    o[rank] = pd ? ev : nan;
  }
  if (l == 0 && NX > D) o[D] = pd ? 0.0 : nan;
}

hipError_t launch_walk_moments_wide(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                    uint32_t n, double* acc, size_t ld, hipStream_t st) {
  if (d <= kMaxDim || d > kMaxWideDim) return hipErrorInvalidValue;
  auto go = [&](auto dp_tag) {
    constexpr int DP = decltype(dp_tag)::value;
    if (zin != nullptr)
      hipLaunchKernelGGL((walk_moments_wide<DP, true>), dim3(n), dim3(64), 0, st, seeds, zin, rng_stream, T, n, d, acc,
                         ld);
    else
      hipLaunchKernelGGL((walk_moments_wide<DP, false>), dim3(n), dim3(64), 0, st, seeds, zin, rng_stream, T, n, d,
                         acc, ld);
  };
  if (d <= 16)
    go(std::integral_constant<int, 16>{});
  else if (d <= 24)
    go(std::integral_constant<int, 24>{});
  else
    go(std::integral_constant<int, 32>{});
  return hipGetLastError();
}

hipError_t launch_walk_eigen_wide(int model, int d, const double* acc, size_t ld, uint32_t T, uint32_t n, double* out,
                                  hipStream_t st) {
  if (d <= kMaxDim || d > kMaxWideDim || model < 0 || model > 4) return hipErrorInvalidValue;
  const size_t lds = wide_eigen_lds_bytes(model, d);
  if (d <= 16)
    hipLaunchKernelGGL(walk_eigen_wide<16>, dim3(n), dim3(64), lds, st, model, d, acc, ld, T, n, out);
  else if (d <= 24)
    hipLaunchKernelGGL(walk_eigen_wide<24>, dim3(n), dim3(64), lds, st, model, d, acc, ld, T, n, out);
  else
    hipLaunchKernelGGL(walk_eigen_wide<32>, dim3(n), dim3(64), lds, st, model, d, acc, ld, T, n, out);
  return hipGetLastError();
}

}  // namespace jne
