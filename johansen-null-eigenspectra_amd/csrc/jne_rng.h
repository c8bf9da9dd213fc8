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
// the same operations.  Three steps below are written in forms that reach the
// same values with fewer instructions (the exponent split, the logarithm at
// twice the scale, the quadrant turn); tests/test_transform_forms.py compares
// them with the plain forms over every radius input the stream can produce
// and the angle words near the quadrant edges, where a sign of zero could
// differ.
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

// 2 ln(u) for u in [2^-33, 1].  The logarithm: u = m * 2^e with m in (C, 2C],
// C = 0x1.6a09e6p-1 (sqrt(1/2) rounded down), and ln(u) = L = e ln2_hi +
// (e ln2_lo + lnm), lnm = f + f^2 Q(f), f = m - 1 (exact; Horner in float32
// with fixed coefficients).  Positive normal floats order as their bit
// patterns and adding k 2^23 to a pattern scales the value by 2^k, so u in
// (2C 2^(e-1), 2C 2^e] gives -e = floor((bits(2C) - bits(u)) / 2^23) (one
// subtraction, one shift) and m = u 2^-e (v_ldexp_f32): the (m, e) of the
// split by frexp (mant in [1/2, 1); m = 2 mant, e = exp - 1 when mant <= C)
// for every u of the range.  The return value is 2L computed at twice the
// scale: h = 2f = 2m - 2, Horner in h with coefficients c_k 2^(-k-1) (= Q/2,
// every partial sum exactly 2^(-k-1) times the one in f), 2 lnm = h + h^2 Q/2
// and the ln 2 terms doubled, -e entering with their signs turned.  Scaling
// by a power of two commutes with rounding (no partial result is subnormal),
// so this is exactly 2L, +0 at u = 1, and -(2L) is bit for bit -2 * L.
__device__ __forceinline__ float two_ln(float u) {
#pragma clang fp contract(off)
  const int32_t ub = (int32_t)__float_as_uint(u);
  const int32_t ne = ((int32_t)0x3FB504F3 - ub) >> 23;  // 0x3FB504F3 = bits(2C); ne = -e
  const float m = __builtin_amdgcn_ldexpf(u, ne);  // exact: m is a normal float
  const float h = __builtin_fmaf(m, 2.0f, -2.0f);  // 2f, exact
  float q = 0x1.15425cp-14f;
  q = __builtin_fmaf(q, h, -0x1.d962f2p-13f);
  q = __builtin_fmaf(q, h, 0x1.e3d33ap-12f);
  q = __builtin_fmaf(q, h, -0x1.fcc1aep-11f);
  q = __builtin_fmaf(q, h, 0x1.237a56p-9f);
  q = __builtin_fmaf(q, h, -0x1.5559e6p-8f);
  q = __builtin_fmaf(q, h, 0x1.99a13p-7f);
  q = __builtin_fmaf(q, h, -0x1.000008p-5f);
  q = __builtin_fmaf(q, h, 0x1.55554ep-4f);
  q = __builtin_fmaf(q, h, -0x1p-2f);
  const float h2 = h * h;
  const float lnm2 = __builtin_fmaf(h2, q, h);
  const float nef = (float)ne;
  return __builtin_fmaf(nef, -0x1.62e43p+0f, __builtin_fmaf(nef, 0x1.05c61p-28f, lnm2));
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
  // -2 ln(u1) is -0 (u1 = 1) or in [5.9e-8, 46.1]; the negation is a source
  // modifier of the square root's instructions
  const float r = sqrt_cr(-two_ln(u1));
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
  const float s = g * sp;  // sin(pi g/2)/sqrt2
  // (cp - s, cp + s) = (cos, sin)(pi/4 + pi g/2), turned by q = wb >> 30
  // quarter turns: q = 0: (cp - s, cp + s), 1: (-(cp + s), cp - s),
  // 2: (-(cp - s), -(cp + s)), 3: (cp + s, -(cp - s)).  The factor pair
  // (cp + s', cp - s'), s' = s negated when bit 30 of wb is clear, picks the
  // entries; the signs go on r as sign flips of a product: bit 31 flips both
  // outputs, bit 30 flips z0 once more.  Sign flips and x + (-y) = x - y are
  // exact, so this is bit for bit the select-and-flip of the rotation.
  // v_bitop3_b32 truth tables over (x, y, 0x80000000): 0x78 = x ^ (y & sign),
  // 0xD2 = x ^ (~y & sign)
  const uint32_t t = wb << 1;  // bit 30 of wb at the sign position
  const float sf = __uint_as_float(__builtin_amdgcn_bitop3_b32(__float_as_uint(s), t, 0x80000000u, 0xD2));
  const float a = cp + sf;
  const float b = cp - sf;
  const float r1 = __uint_as_float(__builtin_amdgcn_bitop3_b32(__float_as_uint(r), wb, 0x80000000u, 0x78));
  const float r0 = __uint_as_float(__builtin_amdgcn_bitop3_b32(__float_as_uint(r1), t, 0x80000000u, 0x78));
  z0 = r0 * a;
  z1 = r1 * b;
}

// Four normals of Philox call q of walk `seed`.
__device__ __forceinline__ void normals4(uint64_t q, uint32_t seed, uint32_t rng_stream, float out[4]) {
  const u32x4 w = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), seed, 0u, rng_stream, kKey1);
  normal_pair(w.x, w.y, out[0], out[1]);
  normal_pair(w.z, w.w, out[2], out[3]);
}

}  // namespace jne
