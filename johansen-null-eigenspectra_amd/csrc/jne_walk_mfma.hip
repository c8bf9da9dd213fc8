// jne_walk_mfma.hip -- mixed-precision moment pass on the matrix cores.
//
// Same output as walk_moments_fp64 (the canonical per-walk moments consumed by
// walk_eigen_fp64), computed as a split between FP64 per-segment terms and an
// MFMA Gram of small within-segment quantities.  With s = sqrt(1/T), z_t the
// unscaled normals, a segment of n <= 64 steps [ts, ts + n) starting at the
// FP64 state B0 = B_ts, k = t - ts and c_t = sum_{ts <= r < t} z_r (so
// b_t = B0 + s c_t and e_t = s z_t):
//
//   sum b b^T   = sum_seg [n B0 B0^T + s (B0 sig^T + sig B0^T)] + s^2 sum c c^T
//   sum e b^T   = sum_seg [s eps B0^T]                          + s^2 sum z c^T
//   sum u^p b   = sum_seg [B0 sum_seg u^p]                      + s   sum u^p c
//   sum u^p e   =                                                 s   sum u^p z
//
// (eps = sum_seg z, sig = sum_seg c).  The bracketed terms are FP64, once per
// segment.  The rest comes from the Gram of the 32 rows
//
//   X = [z_0..z_{D-1}, 0.., c_0..c_{D-1} (from row 12), 0.., 1, k, k(k-1)/2, 0..]
//
// (the last three from row 24) over the segment's steps, accumulated by
// v_mfma_f32_32x32x16_{f16,bf16} on hi/lo (or hi/mid/lo) splits, or by
// v_mfma_f32_32x32x2f32 on FP32 values.  The rows 1, k, k(k-1)/2 are small
// integers, exact in every operand format, so sum u^p x with u_t = (ts+1+k)/T
// is formed without rounding u: the MFMA gives sum x, sum k x and
// sum k(k-1)/2 x per segment and the FP64 flush weights them by the segment's
// start ts.  The FP32 accumulators are flushed into FP64 once per segment, so
// FP32 rounding never spans more than 64 steps; the segment sums of z that
// drive the state B0 are formed in FP64 from the stored normals.
//
// One wave per walk.  Lane l holds row l%32 of X for 8 of every 16 steps
// (half l/32), which is both the A and the B operand of X X^T.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jne.h"
#include "jne_internal.h"
#include "jne_rng.h"

// Tuning: unroll the Philox calls of a segment (independent chains for the
// scheduler, at the cost of registers).
#ifndef JNE_MFMA_UNROLL_CALLS
#define JNE_MFMA_UNROLL_CALLS 0
#endif

namespace jne {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

constexpr int kSegSteps = 64;  // steps per FP64 segment and per RNG round
constexpr int kChunks = 4;     // 16-step MFMA chunks per segment
constexpr int kRowC = 12;      // first row of the c block
constexpr int kRowDet = 24;    // rows 24, 25, 26: 1, k, k(k-1)/2

template <int D>
struct MfmaShape {
  static constexpr int CALLS = kSegSteps * D / 4;  // Philox calls per segment
  static constexpr int IT = (CALLS + 63) / 64;     // call iterations per lane
  static constexpr int ZP = kSegSteps + 4;         // LDS pitch of one component (floats)
  static constexpr int NG = (D * D + 63) / 64;     // (i, j) grid entries per lane
  static constexpr int EPS_P = 64 / D;             // lanes per component of the FP64 eps sum
  static constexpr int ZBYTES = D * ZP * 4;
  static constexpr int PBBYTES = 32 * 33 * 8;
  static constexpr int SMEM = ZBYTES > PBBYTES ? ZBYTES : PBBYTES;  // normals, later the Gram
  static_assert(D <= kRowC, "dim must fit the z block");
};

// Chunk MFMA of one 16-step chunk.  x[8] are this lane's 8 row values (steps
// 8h + k, or 2k + h for the interleaved FP32 form).
template <int PREC>
struct ChunkMfma;

template <>
struct ChunkMfma<JNE_PREC_MFMA_FP16X2> {
  static constexpr bool kInterleaved = false;
  static constexpr bool kPackedConst = true;  // rows 1, k, k(k-1)/2 (< 2048) are exact fp16: lo = 0
  __device__ __forceinline__ static void split(const float (&x)[8], u32x4v& hi, u32x4v& lo) {
    half8 h, l;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      h[k] = (_Float16)x[k];
      l[k] = (_Float16)(x[k] - (float)h[k]);
    }
    hi = __builtin_bit_cast(u32x4v, h);
    lo = __builtin_bit_cast(u32x4v, l);
  }
  __device__ __forceinline__ static void mma(const u32x4v& hi, const u32x4v& lo, floatx16& acc, floatx16& acc2) {
    const half8 h = __builtin_bit_cast(half8, hi), l = __builtin_bit_cast(half8, lo);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(h, h, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(h, l, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(l, h, acc2, 0, 0, 0);
  }
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    u32x4v hi, lo;
    split(x, hi, lo);
    mma(hi, lo, acc, acc2);
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_BF16X2> {
  static constexpr bool kInterleaved = false;
  static constexpr bool kPackedConst = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    bf16x8 hi, lo;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      hi[k] = (__bf16)x[k];
      lo[k] = (__bf16)(x[k] - (float)hi[k]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc2, 0, 0, 0);
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_BF16X3> {
  static constexpr bool kInterleaved = false;
  static constexpr bool kPackedConst = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    bf16x8 hi, mi, lo;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      hi[k] = (__bf16)x[k];
      const float r1 = x[k] - (float)hi[k];
      mi[k] = (__bf16)r1;
      lo[k] = (__bf16)(r1 - (float)mi[k]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, mi, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mi, hi, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mi, mi, acc2, 0, 0, 0);
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_FP32> {
  static constexpr bool kInterleaved = true;
  static constexpr bool kPackedConst = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16&) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[m], x[m], acc, 0, 0, 0);
  }
};

template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

// The value held by lane l ^ 32 (the other half-wave).  v_permlane32_swap
// exchanges the upper half of its first operand with the lower half of its
// second: with both operands v, the lower half finds the partner in the second
// result and the upper half in the first.
__device__ __forceinline__ float partner_half(float v, int h) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(h ? r[0] : r[1]);
}

template <int D, int PREC, bool FROM_NORMALS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void walk_moments_mfma(const uint32_t* __restrict__ seeds,
                                                        const double* __restrict__ zin, uint32_t rng_stream,
                                                        uint32_t T, uint32_t n_walks, double* __restrict__ acc,
                                                        size_t ld) {
  using Sh = MfmaShape<D>;
  constexpr int ZP = Sh::ZP, NG = Sh::NG, EPS_P = Sh::EPS_P;
  using Op = ChunkMfma<PREC>;

  __shared__ __align__(16) unsigned char smem[Sh::SMEM];
  float(*zbuf)[ZP] = reinterpret_cast<float(*)[ZP]>(smem);  // one segment's normals, component-major
  double(*pb)[33] = reinterpret_cast<double(*)[33]>(smem);  // the final Gram (after the last segment)
  __shared__ double eps_part[EPS_P][D];
  __shared__ double seg_eps[D];
  __shared__ double seg_sig[D];
  __shared__ double b0s[D];

  const int l = threadIdx.x;
  const int row = l & 31;  // also the accumulator column
  const int h = l >> 5;
  const size_t walk = blockIdx.x;
  if (walk >= n_walks) return;  // whole wave exits together
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS) seed = seeds[walk];
  const double* zw = FROM_NORMALS ? zin + walk * (size_t)T * D : nullptr;

  const bool is_z = row < D;
  const bool is_c = row >= kRowC && row < kRowC + D;
  const bool use_const = !is_z && !is_c;  // rows 1, k, k(k-1)/2 and zero rows
  const int comp = is_z ? row : is_c ? row - kRowC : 0;
  const float w1 = row == kRowDet ? 1.0f : 0.0f;
  const float wk = row == kRowDet + 1 ? 1.0f : 0.0f;
  const float wkk = row == kRowDet + 2 ? 1.0f : 0.0f;
  const float hf = (float)h;

  // step index (within the segment) of value k of chunk ch on this lane
  auto kidx = [&](int ch, int k) -> int { return 16 * ch + (Op::kInterleaved ? 2 * k + h : 8 * h + k); };
  // the value of row 1 / k / k(k-1)/2 (0 on other rows) at segment step kk
  auto det_value = [&](int kk) -> float {
    const float kf = (float)kk;
    return __builtin_fmaf(wkk, kf * (kf - 1.0f) * 0.5f, __builtin_fmaf(wk, kf, w1));
  };

  // fp16: the constant rows' hi operands of the 4 chunks, packed, in LDS (their
  // lo part is 0); slot (row - 24) * 2 + h for rows 24..26, slot 6 all zero
  __shared__ u32x4v cst[kChunks][7];
  const int cslot = row >= kRowDet && row < kRowDet + 3 ? (row - kRowDet) * 2 + h : 6;
  if constexpr (Op::kPackedConst) {
#pragma unroll
    for (int ch = 0; ch < kChunks; ++ch) {
      float xv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[k] = det_value(kidx(ch, k));
      u32x4v hi, lo;
      Op::split(xv, hi, lo);
      cst[ch][cslot] = hi;  // lanes sharing slot 6 all write zeros
    }
  }

  // (i, j) grid entries of this lane: e = l + 64 g
  int gi[NG], gj[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int e = l + 64 * g;
    gi[g] = e < D * D ? e / D : 0;
    gj[g] = e < D * D ? e % D : 0;
  }
  double bbA[NG], ebA[NG], mbA = 0.0;
#pragma unroll
  for (int g = 0; g < NG; ++g) bbA[g] = ebA[g] = 0.0;
  const int mp = l / D, mi = l % D;  // moment vector entry (power mp, component mi) for l < 3D
  const CallPos<D> pos0(l);          // normals of this lane's first Philox call in a segment

  if (l < D) b0s[l] = 0.0;
  const double dt = 1.0 / (double)T;
  const double s = sqrt(dt);
  const uint32_t rounds = (T + kSegSteps - 1) / kSegSteps;

  floatx16 accm = {}, accc = {};
  double g64[16];  // FP64 Gram, flushed from the MFMA accumulators once per segment
#pragma unroll
  for (int v = 0; v < 16; ++v) g64[v] = 0.0;

  // one segment; the partial last segment (T not a multiple of 64) runs a copy
  // with the step masks, so the loop over full segments carries none of them
  auto segment = [&](uint32_t r, auto tail_tag) {
    constexpr bool tail = decltype(tail_tag)::value;
    const uint32_t ts = r * kSegSteps;
    // ---- normals of the segment -> LDS
    CallPos<D> pos = pos0;
#if JNE_MFMA_UNROLL_CALLS
#pragma unroll
#else
#pragma unroll 1
#endif
    for (int j = 0; j < Sh::IT; ++j) {
      const int c = l + 64 * j;
      if (Sh::IT * 64 != Sh::CALLS && c >= Sh::CALLS) break;
      float v[4];
      if constexpr (!FROM_NORMALS) normals4((uint64_t)ts * D / 4 + c, seed, rng_stream, v);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        int tl, i;
        pos.at(jj, tl, i);
        const uint32_t t = ts + tl;
        float z;
        if constexpr (FROM_NORMALS)
          z = t < T ? (float)zw[(size_t)t * D + i] : 0.0f;
        else
          z = (!tail || t < T) ? v[jj] : 0.0f;
        zbuf[i][tl] = z;
      }
      pos.template advance<64>();
    }
    __syncthreads();

    // ---- 4 chunks of 16 steps on the matrix cores
    float carry = 0.0f;
    const float* zrow = zbuf[comp];
#pragma unroll 1  // unrolled, the four chunks' loads and operands exceed the register budget
    for (int ch = 0; ch < kChunks; ++ch) {
      const int off = 16 * ch;
      float zm[8], cv[8];
      if constexpr (Op::kInterleaved) {  // steps 2k + h
        float zo[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          zm[k] = zrow[off + 2 * k + h];
          zo[k] = zrow[off + 2 * k + 1 - h];
        }
        float p = carry;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float za = h ? zo[k] : zm[k], zb = h ? zm[k] : zo[k];  // z_2k, z_2k+1
          cv[k] = __builtin_fmaf(hf, za, p);
          p = p + za + zb;
        }
        carry = p;
      } else {  // steps 8h + k
#pragma unroll
        for (int k = 0; k < 8; ++k) zm[k] = zrow[off + 8 * h + k];
        const float sm = ((zm[0] + zm[1]) + (zm[2] + zm[3])) + ((zm[4] + zm[5]) + (zm[6] + zm[7]));
        const float so = partner_half(sm, h);
        float p = __builtin_fmaf(hf, so, carry);  // + sum of steps 0-7 on the upper half
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          cv[k] = p;
          p += zm[k];
        }
        carry = carry + (sm + so);
      }
      float x[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = is_c ? cv[k] : zm[k];
      if constexpr (!tail) {
        if constexpr (Op::kPackedConst) {
          u32x4v hi, lo;
          Op::split(x, hi, lo);
          const u32x4v cv4 = cst[ch][cslot];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            hi[q] = use_const ? cv4[q] : hi[q];
            lo[q] = use_const ? 0u : lo[q];
          }
          Op::mma(hi, lo, accm, accc);
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) x[k] = use_const ? det_value(kidx(ch, k)) : x[k];
          Op::run(x, accm, accc);
        }
      } else {  // steps at or beyond T contribute nothing
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int kk = kidx(ch, k);
          const float xv = use_const ? det_value(kk) : x[k];
          x[k] = ts + (uint32_t)kk < T ? xv : 0.0f;
        }
        Op::run(x, accm, accc);
      }
    }

    // ---- segment end
    // eps = sum of the segment's normals in FP64; lane (i, p) sums steps p, p+P, ... of component i
    if (l < EPS_P * D) {
      const int i = l % D, p = l / D;
      double e = 0.0;
      for (int t = p; t < kSegSteps; t += EPS_P) e += (double)zbuf[i][t];
      eps_part[p][i] = e;
    }
    float f[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      f[v] = accm[v] + accc[v];
      accm[v] = 0.0f;
      accc[v] = 0.0f;
    }
    // accumulator v of lane l is Gram element [8(v/4) + 4h + v%4][l%32]; v = 12 on
    // the lower half is row 24 (the ones row): f[12] there is sum_seg x_col
    if (h == 0 && is_c) seg_sig[comp] = (double)f[12];
    __syncthreads();
    if (l < D) {
      double e = 0.0;
#pragma unroll
      for (int p = 0; p < EPS_P; ++p) e += eps_part[p][l];
      seg_eps[l] = e;
    }
    __syncthreads();
    const uint32_t te = tail ? T : ts + kSegSteps;
    const double nseg = (double)(te - ts);
    {
      // rows 24..26 (v = 12..14, lower half): sum x, sum k x, sum k(k-1)/2 x ->
      // sum x, sum u x, sum u^2 x with u = (a + k)/T, a = ts + 1
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if (v < 12 || v > 14) g64[v] += (double)f[v];
      const double a = (double)ts + 1.0;
      const double P0 = (h == 0 && is_z) ? seg_eps[comp] : (double)f[12];
      const double P1 = (double)f[13], P2 = (double)f[14];
      g64[12] += P0;
      g64[13] += (a * P0 + P1) * dt;
      g64[14] += ((a * a) * P0 + (2.0 * a + 1.0) * P1 + 2.0 * P2) * (dt * dt);
    }
    // FP64 per-segment terms on the (i, j) grid
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      if (l + 64 * g < D * D) {
        const double bi = b0s[gi[g]], bj = b0s[gj[g]];
        const double si = seg_sig[gi[g]], sj = seg_sig[gj[g]];
        bbA[g] += nseg * (bi * bj) + s * ((bi * sj) + (si * bj));
        ebA[g] += (s * seg_eps[gi[g]]) * bj;
      }
    }
    if (l < 3 * D) {
      // sum over the segment of u^p, u = (t+1)/T: closed forms
      const double a = (double)ts, bnd = (double)te;
      const double up = mp == 0   ? nseg
                        : mp == 1 ? (bnd * (bnd + 1.0) - a * (a + 1.0)) * 0.5 * dt
                                  : (bnd * (bnd + 1.0) * (2.0 * bnd + 1.0) - a * (a + 1.0) * (2.0 * a + 1.0)) /
                                        6.0 * dt * dt;
      mbA += b0s[mi] * up;
    }
    __syncthreads();
    if (l < D) b0s[l] += s * seg_eps[l];
    __syncthreads();
  };
  const uint32_t full = T / kSegSteps;
  for (uint32_t r = 0; r < full; ++r) segment(r, BoolTag<false>{});
  if (full < rounds) segment(full, BoolTag<true>{});

  // the Gram to LDS (over the normals buffer)
#pragma unroll
  for (int v = 0; v < 16; ++v) pb[8 * (v / 4) + 4 * h + (v % 4)][row] = g64[v];
  __syncthreads();
  const double s2 = s * s;
  // canonical layout: [0, D^2) sum b b^T, [D^2, 2D^2) sum e b^T, then
  // sum b, sum u b, sum u^2 b, sum e, sum u e, sum u^2 e
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int e = l + 64 * g;
    if (e < D * D) {
      const int i = gi[g], j = gj[g];
      const int lo = i < j ? i : j, hi = i < j ? j : i;  // read one triangle: exactly symmetric
      acc[(size_t)e * ld + walk] = bbA[g] + s2 * pb[kRowC + lo][kRowC + hi];
      acc[(size_t)(D * D + e) * ld + walk] = ebA[g] + s2 * pb[i][kRowC + j];
    }
  }
  if (l < 3 * D) {
    acc[(size_t)(2 * D * D + mp * D + mi) * ld + walk] = mbA + s * pb[kRowDet + mp][kRowC + mi];
    acc[(size_t)(2 * D * D + (3 + mp) * D + mi) * ld + walk] = s * pb[kRowDet + mp][mi];
  }
}

template <int D, int PREC>
static hipError_t launch_mfma_dp(const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                 uint32_t n, double* acc, size_t ld, hipStream_t st) {
  if (zin != nullptr)
    hipLaunchKernelGGL((walk_moments_mfma<D, PREC, true>), dim3(n), dim3(64), 0, st, seeds, zin, rng_stream, T, n,
                       acc, ld);
  else
    hipLaunchKernelGGL((walk_moments_mfma<D, PREC, false>), dim3(n), dim3(64), 0, st, seeds, zin, rng_stream, T, n,
                       acc, ld);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_mfma_p(int d, const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                uint32_t n, double* acc, size_t ld, hipStream_t st) {
  switch (d) {
#define JNE_CASE(DD) \
  case DD:           \
    return launch_mfma_dp<DD, PREC>(seeds, zin, rng_stream, T, n, acc, ld, st);
    JNE_CASE(1) JNE_CASE(2) JNE_CASE(3) JNE_CASE(4) JNE_CASE(5) JNE_CASE(6)
    JNE_CASE(7) JNE_CASE(8) JNE_CASE(9) JNE_CASE(10) JNE_CASE(11) JNE_CASE(12)
#undef JNE_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_walk_moments_mfma(int d, int precision, const uint32_t* seeds, const double* zin,
                                    uint32_t rng_stream, uint32_t T, uint32_t n, double* acc, size_t ld,
                                    hipStream_t st) {
  switch (precision) {
    case JNE_PREC_MFMA_FP16X2: return launch_mfma_p<JNE_PREC_MFMA_FP16X2>(d, seeds, zin, rng_stream, T, n, acc, ld, st);
    case JNE_PREC_MFMA_BF16X2: return launch_mfma_p<JNE_PREC_MFMA_BF16X2>(d, seeds, zin, rng_stream, T, n, acc, ld, st);
    case JNE_PREC_MFMA_BF16X3: return launch_mfma_p<JNE_PREC_MFMA_BF16X3>(d, seeds, zin, rng_stream, T, n, acc, ld, st);
    case JNE_PREC_MFMA_FP32: return launch_mfma_p<JNE_PREC_MFMA_FP32>(d, seeds, zin, rng_stream, T, n, acc, ld, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace jne
