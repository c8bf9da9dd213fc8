// jne_walk_mfma.hip -- mixed-precision moment pass on the matrix cores.
//
// Same output as walk_moments_fp64 (the canonical per-walk moments consumed by
// walk_eigen_fp64), computed as a split between FP64 per-segment terms and an
// MFMA Gram of small within-segment quantities.  With s = sqrt(1/T), z_t the
// unscaled normals, a segment of n <= 64 steps [ts, ts + n) starting at the
// FP64 state B0 = B_ts, k = t - ts and c_t = sum_{ts <= r < t} z_r - m with
// m = eps/2 (so b_t = B0 + s m + s c_t and e_t = s z_t):
//
//   sum b b^T   = sum_seg [n B0' B0'^T + s (B0' sig^T + sig B0'^T)] + s^2 sum c c^T
//   sum e b^T   = sum_seg [s eps B0'^T]                            + s^2 sum z c^T
//   sum u^p b   = sum_seg [B0' sum_seg u^p]                        + s   sum u^p c
//   sum u^p e   =                                                    s   sum u^p z
//
// (eps = sum_seg z, sig = sum_seg c, B0' = B0 + s m).  The bracketed terms are FP64, once per
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
// scheduler, at the cost of registers); target waves per SIMD.
#ifndef JNE_MFMA_UNROLL_CALLS
#define JNE_MFMA_UNROLL_CALLS 1
#endif
#ifndef JNE_MFMA_WAVES
#define JNE_MFMA_WAVES 4
#endif
// fp16 split: round-to-nearest hi (else truncated hi); add the lo*lo product
#ifndef JNE_MFMA_SPLIT_ROUND
#define JNE_MFMA_SPLIT_ROUND 1
#endif
#ifndef JNE_MFMA_LOLO
#define JNE_MFMA_LOLO 0
#endif

namespace jne {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));

constexpr int kSegSteps = 64;  // steps per FP64 segment and per RNG round
constexpr int kChunks = 4;     // 16-step MFMA chunks per segment
constexpr int kRowC = 12;      // first row of the c block
constexpr int kRowDet = 24;    // rows 24, 25, 26: 1, k, k(k-1)/2

// The FP64 Gram entries the output needs, compact in LDS:
//   column c in [12, 24) (the c block): rows 0..26  at (c - 12) * 28 + r
//   column c < 12 (the z block):        rows 24..26 at 336 + 3 c + (r - 24)
// The c-block pitch of 28 keeps every group of 4 rows 16-byte aligned.
constexpr int kGramPitch = 28;
constexpr int kGramZ = 12 * kGramPitch;
constexpr int kGramSize = kGramZ + 12 * 3;

template <int D>
struct MfmaShape {
  static constexpr int CALLS = kSegSteps * D / 4;  // Philox calls per segment
  static constexpr int IT = (CALLS + 63) / 64;     // call iterations per lane
  static constexpr int ZP = kSegSteps + 4;         // LDS pitch of one row (floats)
  static constexpr int ZROWS = D + 4;              // normals, then 1, k, k(k-1)/2, 0
  static constexpr int NG = (D * D + 63) / 64;     // (i, j) grid entries per lane
  static constexpr int EPS_P = 64 / D;             // lanes per component of the FP64 eps sum
  static_assert(D <= kRowC, "dim must fit the z block");
};

// hi/lo splits by truncation: hi keeps the leading bits (exact in the narrow
// format), lo = x - hi is exact in FP32 and rounded once to the narrow format.
__device__ __forceinline__ float trunc_bits(float x, uint32_t mask) {
  return __uint_as_float(__float_as_uint(x) & mask);
}

// MFMA of one 16-step chunk.  x[8] are this lane's 8 row values (steps 8h + k,
// or 2k + h for the interleaved FP32 form).
template <int PREC>
struct ChunkMfma;

template <>
struct ChunkMfma<JNE_PREC_MFMA_FP16X2> {
  static constexpr bool kInterleaved = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    half8 hi, lo;
#if JNE_MFMA_SPLIT_ROUND
    // hi = x rounded to fp16 (packed pairs); x - hi is exact in fp32, and
    // v_dot2_f32_f16 forms it from the packed pair directly: (hi0, hi1).(-1, 0) + x0
    const fp16x2 m0 = {(__fp16)-1.0f, (__fp16)0.0f}, m1 = {(__fp16)0.0f, (__fp16)-1.0f};
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
      fp16x2 hp;
      hp[0] = (__fp16)x[k];
      hp[1] = (__fp16)x[k + 1];
      hi[k] = (_Float16)hp[0];
      hi[k + 1] = (_Float16)hp[1];
      lo[k] = (_Float16)__builtin_amdgcn_fdot2(hp, m0, x[k], false);
      lo[k + 1] = (_Float16)__builtin_amdgcn_fdot2(hp, m1, x[k + 1], false);
    }
#else
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float h = trunc_bits(x[k], 0xFFFFE000u);  // 11 significant bits
      hi[k] = (_Float16)h;
      lo[k] = (_Float16)(x[k] - h);
    }
#endif
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi, hi, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi, lo, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, hi, acc2, 0, 0, 0);
#if JNE_MFMA_LOLO
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, lo, acc2, 0, 0, 0);
#endif
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_BF16X2> {
  static constexpr bool kInterleaved = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    bf16x8 hi, lo;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float h = trunc_bits(x[k], 0xFFFF0000u);  // 8 significant bits
      hi[k] = (__bf16)h;
      lo[k] = (__bf16)(x[k] - h);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc2, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc2, 0, 0, 0);
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_BF16X3> {
  static constexpr bool kInterleaved = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16& acc2) {
    bf16x8 hi, mi, lo;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float h = trunc_bits(x[k], 0xFFFF0000u);
      const float r1 = x[k] - h;
      const float m = trunc_bits(r1, 0xFFFF0000u);
      hi[k] = (__bf16)h;
      mi[k] = (__bf16)m;
      lo[k] = (__bf16)(r1 - m);
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
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16&) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[m], x[m], acc, 0, 0, 0);
  }
};

// Philox calls of a segment unrolled: by-seed form only (the from-normals
// form has no Philox chains to overlap), whole iterations only, and not with
// the three-way bf16 split, whose operands leave no registers for it
template <int D, int PREC, bool FROM_NORMALS>
struct CallUnroll {
  static constexpr int value = (JNE_MFMA_UNROLL_CALLS && !FROM_NORMALS && PREC != JNE_PREC_MFMA_BF16X3 &&
                                MfmaShape<D>::CALLS % 64 == 0)
                                   ? MfmaShape<D>::IT
                                   : 1;
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

// Workgroups are dealt to the 8 XCDs round-robin by index, and each XCD has
// its own L2.  With one walk per workgroup, walks b and b+1 would run on
// different XCDs and their 8-byte stores into one 32-byte sector of the
// entry-major output would reach HBM as separate partial-sector writes.  This
// bijection gives XCD x the contiguous walk range [start_x, start_x + count_x)
// instead, so neighbouring walks share an L2 and their stores merge there.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_walk(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

template <int D, int PREC, bool FROM_NORMALS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(JNE_MFMA_WAVES)))
void walk_moments_mfma(const uint32_t* __restrict__ seeds, const double* __restrict__ zin, uint32_t rng_stream,
                       uint32_t T, uint32_t n_walks, double* __restrict__ acc, size_t ld) {
  using Sh = MfmaShape<D>;
  constexpr int ZP = Sh::ZP, NG = Sh::NG, EPS_P = Sh::EPS_P;
  using Op = ChunkMfma<PREC>;

  // rows 0..D-1: the segment's normals (component-major); rows D..D+3: the
  // constant rows 1, k, k(k-1)/2 and a zero row, written once
  __shared__ __align__(16) float zbuf[Sh::ZROWS][ZP];
  __shared__ __align__(16) double gram[kGramSize];  // FP64 Gram (compact), flushed into every segment
  __shared__ double eps_part[EPS_P][D];
  __shared__ double seg_eps[D];
  __shared__ double seg_sig[D];
  __shared__ double b0s[D];      // B at the segment start (FP64)
  __shared__ double seg_b0c[D];  // B0 + s m: the state the centred c rows are relative to
  __shared__ float seg_m[D];

  const int l = threadIdx.x;
  const int row = l & 31;  // also the accumulator column
  const int h = l >> 5;
  if (blockIdx.x >= n_walks) return;  // whole wave exits together
  const size_t walk = xcd_walk(blockIdx.x, n_walks);
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS) seed = seeds[walk];
  const double* zw = FROM_NORMALS ? zin + walk * (size_t)T * D : nullptr;

  const bool is_z = row < D;
  const bool is_c = row >= kRowC && row < kRowC + D;
  // LDS row feeding this lane's row of X (the c rows scan the z rows)
  const int zr = is_z ? row : is_c ? row - kRowC : (row >= kRowDet && row < kRowDet + 3) ? D + row - kRowDet : D + 3;
  const float hf = (float)h;
  // c rows take the scan, the others the LDS row.  A bitwise blend: written as
  // a select, the compiler turns it into a lane-indexed stack array
  const uint32_t cmask = is_c ? 0xFFFFFFFFu : 0u;

  // constant rows, exact small integers (k(k-1)/2 <= 2016)
  for (int t = l; t < ZP; t += 64) {
    const float kf = (float)t;
    zbuf[D][t] = 1.0f;
    zbuf[D + 1][t] = kf;
    zbuf[D + 2][t] = kf * (kf - 1.0f) * 0.5f;
    zbuf[D + 3][t] = 0.0f;
  }
  for (int e = l; e < kGramSize; e += 64) gram[e] = 0.0;

  // accumulator v of lane l is Gram element [8(v/4) + 4h + v%4][l%32].  The
  // output needs, on a c column, rows 0..26 (v = 0..11 on both halves, v =
  // 12..14 on the lower half: rows 24..26) and, on a z column, rows 24..26
  // (v = 12..14, lower half).  Rows of unused components (dim < 12) hold
  // zeros and land in their own, unread slots of the column.
  double* const gcol = gram + (is_c ? (row - kRowC) * kGramPitch + 4 * h : 0);  // c lanes: v = 0..11
  double* const gdet = is_c ? gram + (row - kRowC) * kGramPitch + kRowDet : gram + kGramZ + 3 * row;
  const bool det_lane = h == 0 && (is_c || is_z);

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

  // one segment; the partial last segment (T not a multiple of 64) runs a copy
  // with the step masks, so the loop over full segments carries none of them
  auto segment = [&](uint32_t r, auto tail_tag) {
    constexpr bool tail = decltype(tail_tag)::value;
    const uint32_t ts = r * kSegSteps;
    // ---- normals of the segment -> LDS
    CallPos<D> pos = pos0;
#pragma unroll CallUnroll<D, PREC, FROM_NORMALS>::value
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

    // ---- eps = sum of the segment's normals in FP64; lane (i, p) sums steps
    // p, p+P, ... of component i.  The c rows are centred on m = eps/2 (an FP32
    // value): c'_t = c_t - m halves their mean square, so the FP32 accumulation
    // rounds less, and b_t = (B0 + s m) + s c'_t
    if (l < EPS_P * D) {
      const int i = l % D, p = l / D;
      double e = 0.0;
      for (int t = p; t < kSegSteps; t += EPS_P) e += (double)zbuf[i][t];
      eps_part[p][i] = e;
    }
    __syncthreads();
    if (l < D) {
      double e = 0.0;
#pragma unroll
      for (int p = 0; p < EPS_P; ++p) e += eps_part[p][l];
      seg_eps[l] = e;
      const float m = (float)(0.5 * e);
      seg_m[l] = m;
      seg_b0c[l] = b0s[l] + s * (double)m;
    }
    __syncthreads();

    // ---- 4 chunks of 16 steps on the matrix cores
    floatx16 accm = {}, accc = {};
    float carry = is_c ? -seg_m[row - kRowC] : 0.0f;
    const float* zrow = zbuf[zr];
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
      for (int k = 0; k < 8; ++k) {
        x[k] = __uint_as_float((__float_as_uint(cv[k]) & cmask) | (__float_as_uint(zm[k]) & ~cmask));
        if constexpr (tail) {  // steps at or beyond T contribute nothing
          const int kk = off + (Op::kInterleaved ? 2 * k + h : 8 * h + k);
          x[k] = ts + (uint32_t)kk < T ? x[k] : 0.0f;
        }
      }
      Op::run(x, accm, accc);
    }

    // ---- segment end
    float f[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) f[v] = accm[v] + accc[v];
    // f[12] on the lower half is row 24 (the ones row): sum_seg x_col
    if (h == 0 && is_c) seg_sig[row - kRowC] = (double)f[12];
    __syncthreads();
    const uint32_t te = tail ? T : ts + kSegSteps;
    const double nseg = (double)(te - ts);
    {
      // flush into the FP64 Gram.  Rows 24..26 (v = 12..14, lower half) hold
      // sum x, sum k x, sum k(k-1)/2 x: stored as sum x, sum u x, sum u^2 x with
      // u = (a + k)/T, a = ts + 1; on the z columns sum x is the FP64 eps.
      const double a = (double)ts + 1.0;
      const double P0 = is_z ? seg_eps[row] : (double)f[12];
      const double P1 = (double)f[13], P2 = (double)f[14];
      double add[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) add[v] = (double)f[v];
      add[12] = P0;
      add[13] = (a * P0 + P1) * dt;
      add[14] = ((a * a) * P0 + (2.0 * a + 1.0) * P1 + 2.0 * P2) * (dt * dt);
      // two uniform regions instead of a branch per entry: the reads of a
      // region issue together and the groups of 4 rows move as 16-byte pairs
      if (is_c) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double2* p = reinterpret_cast<double2*>(gcol + 8 * q);
          double2 lo = p[0], hi2 = p[1];
          lo.x += add[4 * q];
          lo.y += add[4 * q + 1];
          hi2.x += add[4 * q + 2];
          hi2.y += add[4 * q + 3];
          p[0] = lo;
          p[1] = hi2;
        }
      }
      if (det_lane) {
        gdet[0] += add[12];
        gdet[1] += add[13];
        gdet[2] += add[14];
      }
    }
    // FP64 per-segment terms on the (i, j) grid
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      if (l + 64 * g < D * D) {
        const double bi = seg_b0c[gi[g]], bj = seg_b0c[gj[g]];
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
      mbA += seg_b0c[mi] * up;
    }
    __syncthreads();
    // the next reads of b0s come two barriers into the next segment
    if (l < D) b0s[l] += s * seg_eps[l];
  };
  const uint32_t full = T / kSegSteps;
  for (uint32_t r = 0; r < full; ++r) segment(r, BoolTag<false>{});
  if (full < rounds) segment(full, BoolTag<true>{});

  const double s2 = s * s;
  // canonical layout: [0, D^2) sum b b^T, [D^2, 2D^2) sum e b^T, then
  // sum b, sum u b, sum u^2 b, sum e, sum u e, sum u^2 e
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int e = l + 64 * g;
    if (e < D * D) {
      const int i = gi[g], j = gj[g];
      const int lo = i < j ? i : j, hi = i < j ? j : i;  // one triangle: exactly symmetric
      acc[(size_t)e * ld + walk] = bbA[g] + s2 * gram[hi * kGramPitch + kRowC + lo];
      acc[(size_t)(D * D + e) * ld + walk] = ebA[g] + s2 * gram[j * kGramPitch + i];
    }
  }
  if (l < 3 * D) {
    acc[(size_t)(2 * D * D + mp * D + mi) * ld + walk] = mbA + s * gram[mi * kGramPitch + kRowDet + mp];
    acc[(size_t)(2 * D * D + (3 + mp) * D + mi) * ld + walk] = s * gram[kGramZ + 3 * mi + mp];
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
