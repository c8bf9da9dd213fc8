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
// One wave per walk.  A segment runs in three phases: the wave draws its
// normals into LDS rows (64 Philox calls per iteration; where a segment's
// calls are not whole iterations, the last one's spare lanes draw the next
// segment's first calls); a scan lays the c rows beside them (lane (i, p)
// takes a block of steps of component i, sums it in FP64 for eps and writes
// each c_t as the FP32 rounding of an FP64 running sum); then lane l feeds
// row l%32 of X for 8 of every 16 steps (half l/32), read straight from its
// LDS row, as both the A and the B operand of X X^T.  Each lane keeps its
// share of the FP64 Gram in registers; it goes to LDS once, after the last
// segment, for the output.
//
// At dims 1, 2, 3, 4 and 6 one walk would leave most of the 32 x 32 tile and
// of the lanes idle, so a wave carries W = 12 / dim walks as the 12
// components of one walk of the kernel: walk w's components are rows
// w dim .. w dim + dim - 1 of each block, drawn from its own Philox stream,
// and the output keeps only the blocks on the diagonal (the products across
// walks are formed and dropped).  Every quantity a walk's output depends on is
// formed from that walk's own rows, so its results are those of a wave of its
// own.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jne.h"
#include "jne_internal.h"
#include "jne_rng.h"

// Tuning: how many of a lane's Philox calls of a segment to unroll
// (independent chains for the scheduler, at the cost of registers: 2 is the
// most that fits beside the FP64 Gram without spills at dim 12, and runs
// 0.4% faster than 1); target waves per SIMD.
#ifndef JNE_MFMA_UNROLL_CALLS
#define JNE_MFMA_UNROLL_CALLS 2
#endif
#ifndef JNE_MFMA_UNROLL_CALLS_FP16X2
#define JNE_MFMA_UNROLL_CALLS_FP16X2 3
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
// unroll of the loop over a segment's 4 MFMA chunks
#ifndef JNE_MFMA_CHUNK_UNROLL
#define JNE_MFMA_CHUNK_UNROLL 4
#endif
// dims 1, 2, 3, 4 and 6: 12 / dim walks per wave, as the components of one
// 12-component walk of the kernel (0: one walk per wave at every dim)
#ifndef JNE_MFMA_PACK
#define JNE_MFMA_PACK 1
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
//   symmetric form only, column c < 12: rows 12..23 at 372 + 12 c + (r - 12)
constexpr int kGramPitch = 28;
constexpr int kGramZ = 12 * kGramPitch;
constexpr int kGramT = kGramZ + 12 * 3;
constexpr int kGramSize = kGramT + 12 * 12;

template <int D>
struct MfmaShape {
  static constexpr int CALLS = kSegSteps * D / 4;       // Philox calls per segment
  static constexpr int IT = (CALLS + 63) / 64;          // call iterations per lane
  // By-seed form, when a segment's calls are not whole 64-call iterations:
  // the spare lanes of a segment's last iteration draw the next segment's
  // first calls (the same Philox counters), which wait in an LDS buffer of
  // at most OVF calls, so the segments draw CALLS calls each on average
  // instead of 64 IT.
  static constexpr bool CARRY = CALLS > 64 && CALLS % 64 != 0;
  static constexpr int OVF = CARRY ? 64 - gcd_int(CALLS, 64) : 1;
  static constexpr int ZP = kSegSteps + 4;              // LDS pitch of one row (floats)
  static constexpr int NG = (D * D + 63) / 64;          // (i, j) grid entries per lane
  static constexpr int SP = D <= 8 ? 8 : 4;             // scan lanes per component
  static constexpr int BS = kSegSteps / SP;             // steps per scan lane: whole 16-byte units
  // LDS rows, each at its row index of X: z (0..D-1), c (12..12+D-1), 1, k,
  // k(k-1)/2 (24..26); the compact FP64 Gram takes the same bytes after the
  // last segment
  static constexpr int ROW_BYTES = (kRowDet + 3) * ZP * 4;
  static constexpr int RAW_BYTES = ROW_BYTES > kGramSize * 8 ? ROW_BYTES : kGramSize * 8;
  static_assert(D <= kRowC, "dim must fit the z block");
};

// Per-component FP64 values of a segment (see the kernel).
struct alignas(16) SegVec {
  double P, b, Q, E;
};

// hi/lo splits by truncation: hi keeps the leading bits (exact in the narrow
// format), lo = x - hi is exact in FP32 and rounded once to the narrow format.
__device__ __forceinline__ float trunc_bits(float x, uint32_t mask) {
  return __uint_as_float(__float_as_uint(x) & mask);
}

// A wave-uniform double in scalar registers (VALU instructions read them as
// operands), keeping two vector registers free.
__device__ __forceinline__ double uniform_f64(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// MFMA of one 16-step chunk.  x[8] are this lane's 8 row values (steps 8h + k,
// or 2k + h for the interleaved FP32 form).
//
// Symmetric form (kSymmetric): of the cross products of the split pieces only
// one of each transposed pair is formed -- acc2 gets hi lo^T but not
// lo hi^T = (hi lo^T)^T -- and the wave's FP64 Gram accumulates
// F = acc + 2 acc2 for rows 0..23.  The Gram is then (F + F^T) / 2, formed once
// per walk from the two transposed entries, which restores
// hi hi^T + hi lo^T + lo hi^T: one MFMA per chunk fewer.  Rows 24..26 (the
// constants 1, k, k(k-1)/2 with k < 64, at most 1,953) are exact in fp16, so
// their lo is zero, acc + acc2 is already their full product and the
// per-segment sums read it directly.  In bf16 (8 significant bits) k(k-1)/2 is
// not exact, so the bf16 forms keep both cross products.
#ifndef JNE_MFMA_SYM
#define JNE_MFMA_SYM 1
#endif
template <int PREC>
struct ChunkMfma;

template <>
struct ChunkMfma<JNE_PREC_MFMA_FP16X2> {
  static constexpr bool kInterleaved = false;
  static constexpr bool kSymmetric = JNE_MFMA_SYM;
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
    if constexpr (!kSymmetric) acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, hi, acc2, 0, 0, 0);
#if JNE_MFMA_LOLO
    // lo lo^T is symmetric: in the symmetric form it joins hi hi^T
    if constexpr (kSymmetric)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, lo, acc, 0, 0, 0);
    else
      acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, lo, acc2, 0, 0, 0);
#endif
  }
};

template <>
struct ChunkMfma<JNE_PREC_MFMA_BF16X2> {
  static constexpr bool kInterleaved = false;
  static constexpr bool kSymmetric = false;
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
  static constexpr bool kSymmetric = false;
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
  static constexpr bool kSymmetric = false;
  __device__ __forceinline__ static void run(const float (&x)[8], floatx16& acc, floatx16&) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[m], x[m], acc, 0, 0, 0);
  }
};

// The compensated FP32 VALU check (JNE_PREC_FP32_KAHAN) runs the same
// segment split with the MFMA Gram replaced by FP32 products of the same X
// rows (z, c, 1, k, k(k-1)/2) added into Kahan-compensated FP32 sums, one
// per Gram entry the flush reads: for each c column the z, c and constant
// rows, for each z column the rows k and k(k-1)/2.  Lane l owns entries
// l, l + 64, ...; each segment's sums go to an LDS tile of doubles
// (sum - compensation, formed in FP64) that every lane reads in the MFMA
// output layout, so the flush is shared with the MFMA precisions.
template <int D>
struct KahanShape {
  static constexpr int PER_C = 2 * D + 3;            // rows per c column
  static constexpr int NE = D * PER_C + 2 * D;       // entries per segment
  static constexpr int NQ = (NE + 63) / 64;          // entries per lane
  static constexpr int COLS = kRowC + D;             // tile columns (X rows 0..11+D)
};

// X rows (i, j) of entry e of the compensated Gram (see KahanShape).
template <int D>
__device__ __forceinline__ void kahan_entry(int e, int& i, int& j) {
  using K = KahanShape<D>;
  if (e < D * K::PER_C) {
    const int jc = e / K::PER_C, r = e - jc * K::PER_C;
    j = kRowC + jc;
    i = r < D ? r : r < 2 * D ? kRowC + (r - D) : kRowDet + (r - 2 * D);
  } else {
    const int ez = e - D * K::PER_C;
    j = ez >> 1;
    i = kRowDet + 1 + (ez & 1);
  }
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
#pragma clang fp contract(off)
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Philox calls of a segment unrolled: by-seed form only (the from-normals
// form has no Philox chains to overlap), whole iterations only, and not with
// the three-way bf16 split, whose operands leave no registers for it.  The
// fp16x2 pass with one walk per wave keeps three calls in flight (all of a
// dim-12 segment's three iterations) within its 128 VGPRs; the other operand
// splits and the packed waves (W walks) hold more registers and keep two.
template <int D, int PREC, bool FROM_NORMALS, int W = 1>
struct CallUnroll {
  static constexpr int U =
      PREC == JNE_PREC_MFMA_FP16X2 && W == 1 ? JNE_MFMA_UNROLL_CALLS_FP16X2 : JNE_MFMA_UNROLL_CALLS;
  static constexpr int value = (U > 1 && !FROM_NORMALS && PREC != JNE_PREC_MFMA_BF16X3 && MfmaShape<D>::CALLS % 64 == 0)
                                   ? (MfmaShape<D>::IT < U ? MfmaShape<D>::IT : U)
                                   : 1;
};

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

// D: components of the wave (W walks of D / W each)
template <int D, int PREC, bool FROM_NORMALS, int W = 1>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(PREC == JNE_PREC_FP32_KAHAN ? 2 : JNE_MFMA_WAVES)))
void walk_moments_mfma(const uint32_t* __restrict__ seeds, const double* __restrict__ zin, uint32_t rng_stream,
                       uint32_t T, uint32_t n_walks, double* __restrict__ acc, size_t ld) {
  constexpr int DW = D / W;  // components of one walk
  static_assert(D % W == 0 && (W == 1 || MfmaShape<D>::CALLS % 64 == 0), "packed walks need whole iterations");
  using Sh = MfmaShape<D>;
  constexpr int ZP = Sh::ZP, NG = Sh::NG, SP = Sh::SP, BS = Sh::BS;
  constexpr bool kKahan = PREC == JNE_PREC_FP32_KAHAN;
  using Op = ChunkMfma<kKahan ? JNE_PREC_MFMA_FP32 : PREC>;
  constexpr bool kSym = !kKahan && Op::kSymmetric;
  using KS = KahanShape<D>;

  // Row r of X lives in LDS row r: the segment's normals z (rows 0..D-1), its
  // c rows (12..12+D-1) and the constant rows 1, k, k(k-1)/2 (24..26), written
  // once.  With a pitch of 17 16-byte units, the 8 lanes of a 128-bit read
  // group (8 consecutive rows) fall in distinct banks.  After the last
  // segment the same bytes hold the FP64 Gram.
  __shared__ __align__(16) unsigned char rows_raw[Sh::RAW_BYTES];
  float(*const xs)[ZP] = reinterpret_cast<float(*)[ZP]>(rows_raw);
  float(*const zbuf)[ZP] = xs;
  float(*const cbuf)[ZP] = xs + kRowC;
  double* const gram = reinterpret_cast<double*>(rows_raw);
  __shared__ double eps_part[SP][D];
  // (not with the three-way bf16 split, whose operands leave no registers for it)
  constexpr bool kCarry = Sh::CARRY && !FROM_NORMALS && PREC != JNE_PREC_MFMA_BF16X3 && PREC != JNE_PREC_FP32_KAHAN;
  __shared__ float4 ovf[kCarry ? Sh::OVF : 1];  // calls drawn ahead for the next segment
  __shared__ double seg_eps[D];
  __shared__ double b0s[D];  // B at the segment start (FP64)
  // per component j of a segment: b = B0 + s m (the state the centred c rows
  // are relative to), P = n b + s sig, Q = s sig, E = s eps, so that the
  // segment adds b_i P_j + Q_i b_j to sum b b^T and E_i b_j to sum e b^T
  __shared__ SegVec seg[D];
  __shared__ double mb[3 * D];   // sum over segments of B0' sum_seg u^p (lanes l < 3D)
  __shared__ double gd[2 * D][3];  // Gram rows 24..26 of the c columns, then of the z columns
  // the compensated check's per-segment Gram entries, [X row][X column]
  __shared__ double gk[kKahan ? kRowDet + 3 : 1][kKahan ? KS::COLS : 1];

  const int l = threadIdx.x;
  const int row = l & 31;  // also the accumulator column
  const int h = l >> 5;
  const uint32_t n_groups = (n_walks + W - 1) / W;
  if (blockIdx.x >= n_groups) return;  // whole wave exits together
  const size_t walk = (size_t)xcd_walk(blockIdx.x, n_groups) * W;  // the wave's first walk
  uint32_t seed = 0;
  if constexpr (!FROM_NORMALS && W == 1) seed = seeds[walk];
  const double* zw = FROM_NORMALS ? zin + walk * (size_t)T * DW : nullptr;
  // packed walks: a segment's Philox call c belongs to walk c / (16 DW) (a
  // walk draws 16 DW calls per segment); the walks' seeds, in LDS
  __shared__ uint32_t wseed[W];
  if constexpr (W > 1 && !FROM_NORMALS) {
    if (l < W) wseed[l] = seeds[walk + l < n_walks ? walk + l : n_walks - 1];
    __syncthreads();
  }

  const bool is_z = row < D;
  const bool is_c = row >= kRowC && row < kRowC + D;
  // LDS row feeding this lane's row of X: its own; rows 27..31 read rows
  // 19..23 (distinct banks from rows 24..26 in their read group).  Their Gram
  // entries, and those of the rows of unused components (dim < 12, zero), are
  // never read.
  const float* const xrow = xs[row < kRowDet + 3 ? row : row - 8];

  // the rows of unused components (dim < 12) zero, then the constant rows,
  // exact small integers (k(k-1)/2 <= 2016)
  if constexpr (D < kRowC) {
    for (int r = D; r < kRowC; ++r)
      for (int t = l; t < ZP; t += 64) {
        xs[r][t] = 0.0f;
        xs[kRowC + r][t] = 0.0f;
      }
  }
  for (int t = l; t < ZP; t += 64) {
    const float kf = (float)t;
    xs[kRowDet][t] = 1.0f;
    xs[kRowDet + 1][t] = kf;
    xs[kRowDet + 2][t] = kf * (kf - 1.0f) * 0.5f;
  }

  // FP64 Gram in registers: g[v] sums accumulator v of every segment, Gram
  // element [8(v/4) + 4h + v%4][l%32].  The output needs, on a c column, rows
  // 0..23 (v = 0..11 on both halves) and rows 24..26 (v = 12..14 on the
  // lower half), and on a z column rows 24..26; rows 24..26 are summed in LDS.
  double g[12];
#pragma unroll
  for (int v = 0; v < 12; ++v) g[v] = 0.0;
  // rows 24..26 (v = 12..14 on the lower half) of the c and z columns: in LDS
  const bool det_lane = h == 0 && (is_c || is_z);
  double* const gdl = gd[is_c ? row - kRowC : D + (row < D ? row : 0)];
  for (int e = l; e < 6 * D; e += 64) gd[e / 3][e % 3] = 0.0;

  // (i, j) grid entries of this lane: e = l + 64 g
  int gi[NG], gj[NG];
#pragma unroll
  for (int q = 0; q < NG; ++q) {
    const int e = l + 64 * q;
    gi[q] = e < D * D ? e / D : 0;
    gj[q] = e < D * D ? e % D : 0;
  }
  double bbA[NG], ebA[NG];
#pragma unroll
  for (int q = 0; q < NG; ++q) bbA[q] = ebA[q] = 0.0;
  // the compensated check: this lane's Gram entries as LDS row offsets; the
  // entries no lane computes (rows of unused components) stay zero
  int kei[kKahan ? KS::NQ : 1], kej[kKahan ? KS::NQ : 1];
  if constexpr (kKahan) {
    for (int e = l; e < (kRowDet + 3) * KS::COLS; e += 64) (&gk[0][0])[e] = 0.0;
#pragma unroll
    for (int q = 0; q < KS::NQ; ++q) {
      int i = 0, j = 0;
      if (l + 64 * q < KS::NE) kahan_entry<D>(l + 64 * q, i, j);
      kei[q] = i;
      kej[q] = j;
    }
  }
  const int mp = l / D, mi = l % D;  // moment vector entry (power mp, component mi) for l < 3D
  const CallPos<D> pos0(l);          // normals of this lane's first Philox call in a segment
  // scan lane (component si, block sp): steps [sp BS, sp BS + BS) of a segment
  const int si = l % D, sp = l / D;
  const bool scan_lane = l < SP * D;

  if (l < D) b0s[l] = 0.0;
  if (l < 3 * D) mb[l] = 0.0;
  // walk constants: uniform, held in scalar registers
  const double dt = uniform_f64(1.0 / (double)T);
  const double s = uniform_f64(sqrt(dt));
  const double dt2 = uniform_f64(dt * dt);
  const uint32_t rounds = (T + kSegSteps - 1) / kSegSteps;
  uint32_t drawn = 0;  // calls drawn so far (carry form): a multiple of 64

  // one segment; the partial last segment (T not a multiple of 64) runs a copy
  // with the step masks, so the loop over full segments carries none of them
  auto segment = [&](uint32_t r, auto tail_tag) {
    constexpr bool tail = decltype(tail_tag)::value;
    const uint32_t ts = r * kSegSteps;
    // ---- normals of the segment -> LDS
    // Wave priority: low while drawing (long runs of independent VALU work),
    // high from the scan to the flush, whose LDS round trips and MFMA
    // dependences leave gaps that the other waves' drawing then fills.
    __builtin_amdgcn_s_setprio(0);
    if constexpr (kCarry) {
      // calls [r CALLS, drawn) came with the previous segment's last iteration
      const int pending = (int)(drawn - r * (uint32_t)Sh::CALLS);
      if (l < pending) {
        const float4 q = ovf[l];
        const float qv[4] = {q.x, q.y, q.z, q.w};
        const CallPos<D> p(l);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          int tl, i;
          p.at(jj, tl, i);
          zbuf[i][tl] = (!tail || ts + tl < T) ? qv[jj] : 0.0f;
        }
      }
      // the rest, 64 calls per iteration; calls past the segment (spare lanes
      // of the last iteration) wait in ovf for the next one
      const uint32_t need = (r + 1) * (uint32_t)Sh::CALLS - drawn;
      const int nit = (int)((need + 63) / 64);
      int c = pending + l;  // this lane's call, counted from the segment's first
      CallPos<D> pos(c);
      for (int j = 0; j < nit; ++j) {
        float v[4];
        normals4(drawn + 64u * j + l, seed, rng_stream, v);  // 32-bit counter (kMaxSteps)
        if (c < Sh::CALLS) {
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            int tl, i;
            pos.at(jj, tl, i);
            zbuf[i][tl] = (!tail || ts + tl < T) ? v[jj] : 0.0f;
          }
        } else if (!tail) {
          ovf[c - Sh::CALLS] = make_float4(v[0], v[1], v[2], v[3]);
        }
        c += 64;
        pos.template advance<64>();
      }
      drawn += 64u * (uint32_t)nit;
    } else if constexpr (W > 1) {
      // call c: local call cl of walk w, whose normals 4 cl .. 4 cl + 3 are
      // (step, component) of that walk, component i -> row w DW + i
#pragma unroll CallUnroll<D, PREC, FROM_NORMALS, W>::value
      for (int j = 0; j < Sh::IT; ++j) {
        const int c = l + 64 * j;
        const int w = c / (16 * DW), cl = c - w * (16 * DW);
        float v[4];
        if constexpr (!FROM_NORMALS) normals4((uint32_t)((uint64_t)ts * DW / 4) + cl, wseed[w], rng_stream, v);
        const CallPos<DW> p(cl);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          int tl, i;
          p.at(jj, tl, i);
          const uint32_t t = ts + tl;
          float z;
          if constexpr (FROM_NORMALS) {
            const size_t wk = walk + w < n_walks ? (size_t)w : 0;
            z = t < T ? (float)zw[wk * (size_t)T * DW + (size_t)t * DW + i] : 0.0f;
          } else {
            z = (!tail || t < T) ? v[jj] : 0.0f;
          }
          zbuf[w * DW + i][tl] = z;
        }
      }
    } else {
    CallPos<D> pos = pos0;
#pragma unroll CallUnroll<D, PREC, FROM_NORMALS>::value
    for (int j = 0; j < Sh::IT; ++j) {
      const int c = l + 64 * j;
      if (Sh::IT * 64 != Sh::CALLS && c >= Sh::CALLS) break;
      float v[4];
      if constexpr (!FROM_NORMALS) normals4((uint32_t)((uint64_t)ts * D / 4) + c, seed, rng_stream, v);
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
    }
    __syncthreads();

    __builtin_amdgcn_s_setprio(2);
    // ---- scan, part 1: lane (si, sp) sums its block of component si in FP64:
    // the partial sums give the segment's eps, which drives the FP64 state,
    // and the offsets of the c rows in the blocks after it
    // (128-bit LDS reads and writes: the 8 lanes of a read group cover
    // distinct banks)
    float zb[BS];
    if (scan_lane) {
      const float4* const src = reinterpret_cast<const float4*>(&zbuf[si][sp * BS]);
#pragma unroll
      for (int q = 0; q < BS / 4; ++q) {
        const float4 v = src[q];
        zb[4 * q] = v.x;
        zb[4 * q + 1] = v.y;
        zb[4 * q + 2] = v.z;
        zb[4 * q + 3] = v.w;
      }
      double e = (double)zb[0];  // (no 0.0 + first term: one add fewer)
#pragma unroll
      for (int j = 1; j < BS; ++j) e += (double)zb[j];
      eps_part[sp][si] = e;
    }
    __syncthreads();

    // ---- scan, part 2: eps of component si; the c rows are centred on m =
    // eps/2 (an FP32 value): c_t = sum_{ts <= q < t} z_q - m halves their mean
    // square, so the FP32 accumulation rounds less, and b_t = (B0 + s m) + s c_t.
    // Steps at or beyond T get c = 0 (their z are 0), so they add nothing.
    if (scan_lane) {
      double e = eps_part[0][si];
#pragma unroll
      for (int p = 1; p < SP; ++p) e += eps_part[p][si];
      const float m = (float)(0.5 * e);
      // running sum in FP64 (exact but for FP64 rounding): each c_t is then
      // rounded to FP32 once
      double c = -(double)m;
#pragma unroll
      for (int p = 0; p < SP - 1; ++p) c += p < sp ? eps_part[p][si] : 0.0;
      float4* const dst = reinterpret_cast<float4*>(&cbuf[si][sp * BS]);
#pragma unroll
      for (int q = 0; q < BS / 4; ++q) {
        float cq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t t = ts + (uint32_t)(sp * BS + 4 * q + k);
          cq[k] = (!tail || t < T) ? (float)c : 0.0f;
          c += (double)zb[4 * q + k];
        }
        dst[q] = make_float4(cq[0], cq[1], cq[2], cq[3]);
      }
      if (sp == 0) {
        seg_eps[si] = e;
        seg[si].b = b0s[si] + s * (double)m;
        seg[si].E = s * e;
      }
    }
    __syncthreads();

    // ---- the segment's Gram of X: f[v] is element [8(v/4) + 4h + v%4][l%32]
    // (the 32x32 MFMA output layout)
    double f[16];
    if constexpr (kKahan) {
      // FP32 products in Kahan-compensated FP32 sums, then to LDS
      float ks[KS::NQ], kc[KS::NQ];
#pragma unroll
      for (int q = 0; q < KS::NQ; ++q) ks[q] = kc[q] = 0.0f;
#pragma unroll 2
      for (int k4 = 0; k4 < kSegSteps / 4; ++k4) {
#pragma unroll
        for (int q = 0; q < KS::NQ; ++q) {
          const float4 a = *reinterpret_cast<const float4*>(&xs[kei[q]][4 * k4]);
          const float4 b = *reinterpret_cast<const float4*>(&xs[kej[q]][4 * k4]);
          kahan_add(ks[q], kc[q], a.x * b.x);
          kahan_add(ks[q], kc[q], a.y * b.y);
          kahan_add(ks[q], kc[q], a.z * b.z);
          kahan_add(ks[q], kc[q], a.w * b.w);
        }
      }
#pragma unroll
      for (int q = 0; q < KS::NQ; ++q)
        if (l + 64 * q < KS::NE) gk[kei[q]][kej[q]] = (double)ks[q] - (double)kc[q];
      __syncthreads();
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = 8 * (v / 4) + 4 * h + v % 4;
        f[v] = (i < kRowDet + 3 && row < KS::COLS) ? gk[i][row] : 0.0;
      }
    } else {
      // 4 chunks of 16 steps on the matrix cores
      floatx16 accm = {}, accc = {};
#pragma unroll JNE_MFMA_CHUNK_UNROLL
      for (int ch = 0; ch < kChunks; ++ch) {
        const int off = 16 * ch;
        float x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = xrow[off + (Op::kInterleaved ? 2 * k + h : 8 * h + k)];
        Op::run(x, accm, accc);
      }
#pragma unroll
      for (int v = 0; v < 16; ++v)
        f[v] = (double)((kSym && v < 12) ? __builtin_fmaf(2.0f, accc[v], accm[v]) : accm[v] + accc[v]);
    }

    // ---- segment end
    // f[12] on the lower half is row 24 (the ones row): sum_seg x_col
    const uint32_t te = tail ? T : ts + kSegSteps;
    const double nseg = (double)(te - ts);
    if (h == 0 && is_c) {
      SegVec& v = seg[row - kRowC];
      const double sig = f[12];
      v.P = nseg * v.b + s * sig;
      v.Q = s * sig;
    }
    __syncthreads();
    {
      // flush into the FP64 Gram.  Rows 24..26 (v = 12..14, lower half) hold
      // sum x, sum k x, sum k(k-1)/2 x: stored as sum x, sum u x, sum u^2 x with
      // u = (a + k)/T, a = ts + 1; on the z columns sum x is the FP64 eps.
      const double a = (double)ts + 1.0;
      const double P0 = is_z ? seg_eps[row] : f[12];
      const double P1 = f[13], P2 = f[14];
#pragma unroll
      for (int v = 0; v < 12; ++v) g[v] += f[v];
      if (det_lane) {
        gdl[0] += P0;
        gdl[1] += (a * P0 + P1) * dt;
        gdl[2] += ((a * a) * P0 + (2.0 * a + 1.0) * P1 + 2.0 * P2) * dt2;
      }
    }
    // FP64 per-segment terms on the (i, j) grid
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      if (l + 64 * q < D * D) {
        const SegVec& vi = seg[gi[q]];
        const SegVec& vj = seg[gj[q]];
        bbA[q] = fma(vi.b, vj.P, fma(vi.Q, vj.b, bbA[q]));
        ebA[q] = fma(vi.E, vj.b, ebA[q]);
      }
    }
    if (l < 3 * D) {
      // sum over the segment of u^p, u = (t+1)/T, in closed form:
      // sum_{k<n} (A + k)^p with A = ts + 1, n = te - ts <= 64:
      // n A^p + p A^(p-1) n(n-1)/2 (+ n(n-1)(2n-1)/6 for p = 2).  Positive
      // terms and exact integers: no division and no cancellation, unlike the
      // difference of the two cubic prefix sums
      const uint32_t n = te - ts;
      const double A = (double)ts + 1.0;
      const double k1 = (double)(n * (n - 1) / 2), k2 = (double)(n * (n - 1) * (2 * n - 1) / 6);
      const double s1 = (nseg * A + k1) * dt;
      const double s2 = ((nseg * A) * A + (2.0 * A) * k1 + k2) * dt2;
      const double up = mp == 0 ? nseg : mp == 1 ? s1 : s2;
      mb[l] += seg[mi].b * up;
    }
    __syncthreads();
    // the next reads of b0s come two barriers into the next segment
    if (l < D) b0s[l] += s * seg_eps[l];
  };
  const uint32_t full = T / kSegSteps;
  for (uint32_t r = 0; r < full; ++r) segment(r, BoolTag<false>{});
  if (full < rounds) segment(full, BoolTag<true>{});

  // Lane indices of the output, recomputed from an opaque copy of the lane id
  // (an empty asm statement: no instruction), so that the compiler does not
  // keep the ones computed before the segment loop alive across it.
  int lx = l;
  asm volatile("" : "+v"(lx));
  const int xrow_ = lx & 31, xh = lx >> 5;
  const bool xc = xrow_ >= kRowC && xrow_ < kRowC + D, xz = xrow_ < D;

  // the Gram entries the output reads, into LDS over the row buffers (their
  // last reads came before the last segment's barriers): c lanes rows 0..23 of
  // their column, the lower half also rows 24..26; z lanes rows 24..26, and in
  // the symmetric form also rows 12..23 (the transposes of the z rows of the c
  // columns)
  if (xc) {
    double2* const gcol = reinterpret_cast<double2*>(gram + (xrow_ - kRowC) * kGramPitch + 4 * xh);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      gcol[4 * q] = make_double2(g[4 * q], g[4 * q + 1]);
      gcol[4 * q + 1] = make_double2(g[4 * q + 2], g[4 * q + 3]);
    }
  }
  if constexpr (kSym) {
    if (xz) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int r0 = 8 * q + 4 * xh;  // this lane's rows r0..r0+3 of the column
        if (r0 >= kRowC) {
          double2* const gt = reinterpret_cast<double2*>(gram + kGramT + 12 * xrow_ + (r0 - kRowC));
          gt[0] = make_double2(g[4 * q], g[4 * q + 1]);
          gt[1] = make_double2(g[4 * q + 2], g[4 * q + 3]);
        }
      }
    }
  }
  if (xh == 0 && (xc || xz)) {
    double* const gdet = xc ? gram + (xrow_ - kRowC) * kGramPitch + kRowDet : gram + kGramZ + 3 * xrow_;
    const double* const src = gd[xc ? xrow_ - kRowC : D + xrow_];
    gdet[0] = src[0];
    gdet[1] = src[1];
    gdet[2] = src[2];
  }
  __syncthreads();

  const double s2 = s * s;
  // canonical layout: [0, D^2) sum b b^T, [D^2, 2D^2) sum e b^T, then
  // sum b, sum u b, sum u^2 b, sum e, sum u e, sum u^2 e
#pragma unroll
  for (int q = 0; q < NG; ++q) {
    const int e = lx + 64 * q;
    if (e < D * D) {
      const int i = e / D, j = e % D;
      const int lo = i < j ? i : j, hi = i < j ? j : i;  // one triangle: exactly symmetric
      // symmetric form: the mean of the two transposed entries
      const double gbb = kSym ? 0.5 * (gram[hi * kGramPitch + kRowC + lo] + gram[lo * kGramPitch + kRowC + hi])
                              : gram[hi * kGramPitch + kRowC + lo];
      const double geb = kSym ? 0.5 * (gram[j * kGramPitch + i] + gram[kGramT + 12 * i + j]) : gram[j * kGramPitch + i];
      const double vbb = bbA[q] + s2 * gbb;
      const double veb = ebA[q] + s2 * geb;
      if constexpr (W == 1) {
        acc[(size_t)e * ld + walk] = vbb;
        acc[(size_t)(D * D + e) * ld + walk] = veb;
      } else {
        // packed walks: the blocks on the diagonal, in each walk's own layout
        const int w = i / DW, iw = i - w * DW, jw = j - w * DW;
        if (j / DW == w && walk + w < n_walks) {
          acc[(size_t)(iw * DW + jw) * ld + walk + w] = vbb;
          acc[(size_t)(DW * DW + iw * DW + jw) * ld + walk + w] = veb;
        }
      }
    }
  }
  if (lx < 3 * D) {
    const int xp = lx / D, xi = lx % D;
    const double vb = mb[lx] + s * gram[xi * kGramPitch + kRowDet + xp];
    const double ve = s * gram[kGramZ + 3 * xi + xp];
    const int w = xi / DW, iw = xi - w * DW;
    if (walk + w < n_walks) {
      acc[(size_t)(2 * DW * DW + xp * DW + iw) * ld + walk + w] = vb;
      acc[(size_t)(2 * DW * DW + (3 + xp) * DW + iw) * ld + walk + w] = ve;
    }
  }
}

template <int D, int PREC>
static hipError_t launch_mfma_dp(const uint32_t* seeds, const double* zin, uint32_t rng_stream, uint32_t T,
                                 uint32_t n, double* acc, size_t ld, hipStream_t st) {
  // dims dividing 12: 12 / D walks per wave as one 12-component walk
  constexpr int W = (JNE_MFMA_PACK && 12 % D == 0) ? 12 / D : 1;
  constexpr int DS = D * W;
  const unsigned grid = (n + W - 1) / W;
  if (zin != nullptr)
    hipLaunchKernelGGL((walk_moments_mfma<DS, PREC, true, W>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream, T,
                       n, acc, ld);
  else
    hipLaunchKernelGGL((walk_moments_mfma<DS, PREC, false, W>), dim3(grid), dim3(64), 0, st, seeds, zin, rng_stream,
                       T, n, acc, ld);
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
    case JNE_PREC_FP32_KAHAN: return launch_mfma_p<JNE_PREC_FP32_KAHAN>(d, seeds, zin, rng_stream, T, n, acc, ld, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace jne
