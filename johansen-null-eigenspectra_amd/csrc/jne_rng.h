// jne_rng.h -- device random stream: Philox4x32-10 + Box-Muller normals.
//
// Replaces the reference's Xoshiro256++/Ziggurat draw (src/rng_matrix.rs:11-37),
// whose stream depends on the host's physical core count, with a
// counter-based stream that depends only on (rng_stream, seed, step, component):
//
//   normal n = t*dim + i of walk `seed` is output n%4 of the Philox4x32-10 call
//   with counter (n/4 low word, n/4 high word, seed, 0) and key
//   (rng_stream, 0x4A4E4531).
//
// Outputs (w0,w1,w2,w3) of one call give normals 4q..4q+3 as
// normal_pair(w0,w1) and normal_pair(w2,w3).  normal_pair is Box-Muller with
// the logarithm and the sine/cosine replaced by fixed float32 polynomials
// (coefficients from derive_rng_coeffs.py) and a correctly rounded square
// root, so that every result is that of correctly rounded IEEE operations;
// the stream is therefore bit-identical on any IEEE-754 host that performs
// the same operation sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jne {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kKey1 = 0x4A4E4531u;

struct u32x4 {
  uint32_t x, y, z, w;
};

// Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one v_mad_u64_u32 per product gives both halves
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    // three-input XOR (truth table 0x96) in one v_bitop3_b32
    c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    c1 = (uint32_t)p1;
    c2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c3 = (uint32_t)p0;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return u32x4{c0, c1, c2, c3};
}

// ln(u) for u in [2^-33, 1]: u = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// ln(m) = f + f^2 Q(f), f = m - 1 (exact).  The split comes from
// v_frexp_mant/v_frexp_exp (mant in [1/2, 1)): m = mant, or 2 mant below
// sqrt(1/2) with e - 1 -- the same (m, e) as taking the exponent field.
__device__ __forceinline__ float ln_poly(float u) {
#pragma clang fp contract(off)
  const float mant = __builtin_amdgcn_frexp_mantf(u);
  const int ex = __builtin_amdgcn_frexp_expf(u);
  // 0x1.6a09e6p-1 is sqrt(1/2) rounded down: m = 2 mant whenever 2 mant is at
  // most sqrt(2) rounded down, as with the exponent-field split
  const bool lo = mant <= 0x1.6a09e6p-1f;
  const float m = lo ? mant + mant : mant;
  const int e = lo ? ex - 1 : ex;
  const float f = m - 1.0f;
  float q = 0x1.15425cp-4f;
  q = __builtin_fmaf(q, f, -0x1.d962f2p-4f);
  q = __builtin_fmaf(q, f, 0x1.e3d33ap-4f);
  q = __builtin_fmaf(q, f, -0x1.fcc1aep-4f);
  q = __builtin_fmaf(q, f, 0x1.237a56p-3f);
  q = __builtin_fmaf(q, f, -0x1.5559e6p-3f);
  q = __builtin_fmaf(q, f, 0x1.99a13p-3f);
  q = __builtin_fmaf(q, f, -0x1.000008p-2f);
  q = __builtin_fmaf(q, f, 0x1.55554ep-2f);
  q = __builtin_fmaf(q, f, -0x1p-1f);
  const float f2 = f * f;
  const float lnm = __builtin_fmaf(f2, q, f);
  const float ef = (float)e;
  return __builtin_fmaf(ef, 0x1.62e43p-1f, __builtin_fmaf(ef, -0x1.05c61p-29f, lnm));
}

// Correctly rounded sqrt(y) for y = +-0 or y in [2^-96, 2^126]: v_sqrt_f32
// (within 1 ulp) and one correction step to the neighbour whose residual
// has the right sign.  A general sqrtf adds scaling for tiny inputs and a
// special-value check, neither needed on this range.
__device__ __forceinline__ float sqrt_cr(float y) {
#pragma clang fp contract(off)
  const float s = __builtin_amdgcn_sqrtf(y);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  const float rdn = __builtin_fmaf(-dn, s, y);
  const float rup = __builtin_fmaf(-up, s, y);
  float r = rdn <= 0.0f ? dn : s;
  r = rup > 0.0f ? up : r;
  return r;
}

// Two Philox words -> two N(0,1) floats.
__device__ __forceinline__ void normal_pair(uint32_t wa, uint32_t wb, float& z0, float& z1) {
#pragma clang fp contract(off)
  const float u1 = __builtin_fmaf((float)wa, 0x1p-32f, 0x1p-33f);  // (0, 1]
  // -2 ln(u1) is +-0 or in [5.9e-8, 46.1]
  const float r = sqrt_cr(-2.0f * ln_poly(u1));
  const uint32_t q = wb >> 30;
  const int32_t gi = (int32_t)(wb & 0x3FFFFFFFu) - (int32_t)0x20000000;
  const float g = (float)gi * 0x1p-30f;  // [-1/2, 1/2)
  const float x = g * g;
  float sp = 0x1.d53624p-14f;
  sp = __builtin_fmaf(sp, x, -0x1.b1de6p-9f);
  sp = __builtin_fmaf(sp, x, 0x1.cda0f6p-5f);
  sp = __builtin_fmaf(sp, x, -0x1.d3ba5cp-2f);
  sp = __builtin_fmaf(sp, x, 0x1.1c5832p+0f);
  float cp = -0x1.26d016p-16f;
  cp = __builtin_fmaf(cp, x, 0x1.54c234p-11f);
  cp = __builtin_fmaf(cp, x, -0x1.e36a9ep-7f);
  cp = __builtin_fmaf(cp, x, 0x1.6f5a4ap-3f);
  cp = __builtin_fmaf(cp, x, -0x1.bea5b6p-1f);
  cp = __builtin_fmaf(cp, x, 0x1.6a09e6p-1f);
  const float s = g * sp;   // sin(pi g/2)/sqrt2
  const float cq = cp - s;  // cos(pi/4 + pi g/2)
  const float sq = cp + s;  // sin(pi/4 + pi g/2)
  const bool odd = (q & 1u) != 0u;
  const float a = odd ? sq : cq;
  const float b = odd ? cq : sq;
  const uint32_t cs = ((q ^ (q >> 1)) & 1u) << 31;
  const uint32_t ss = (q >> 1) << 31;
  z0 = __uint_as_float(__float_as_uint(r * a) ^ cs);
  z1 = __uint_as_float(__float_as_uint(r * b) ^ ss);
}

// Four normals of Philox call q of walk `seed`.
__device__ __forceinline__ void normals4(uint64_t q, uint32_t seed, uint32_t rng_stream, float out[4]) {
  const u32x4 w = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), seed, 0u, rng_stream, kKey1);
  normal_pair(w.x, w.y, out[0], out[1]);
  normal_pair(w.z, w.w, out[2], out[3]);
}

}  // namespace jne
