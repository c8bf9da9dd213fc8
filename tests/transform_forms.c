/* transform_forms.c -- the device normal transform's operation forms against
 * the oracle's, over whole input ranges (CPU; built and run by
 * test_transform_forms.py).
 *
 * johansen-null-eigenspectra_amd/csrc/jne_rng.h computes the Box-Muller
 * radius from 2 ln(u) formed at twice the scale with an exponent split by
 * bit arithmetic, and applies the quadrant turn as sign flips of r and s
 * instead of a select.  Each form is restated here in the device's operation
 * order and compared bit for bit with the oracle's jo_normal_pair and
 * ln_poly (oracle/jne_oracle.c, included below):
 *   radius: every float u1 in [2^-33, 1] (all u1 the stream can produce);
 *   turn:   every quadrant, with the angle word's low 30 bits covering the
 *           2^21 values at each end of the range (where cos/sin of the
 *           turned angle pass through zero) and a stride-1021 sample of the
 *           rest, for three radius words (r = -0 among them).
 * The square root is sqrtf (correctly rounded, as the device's sqrt_cr).
 * Negations are written as sign-bit flips, as the device applies them
 * (source modifiers), so no compiler rewrite of -fma(a, b, c) is involved.
 * Prints one line of counts and exits 1 on any difference. */
#include <stdio.h>

#include "jne_oracle.c"

static float neg(float x) { return u2f(f2u(x) ^ 0x80000000u); }

static float two_ln_device(float u) {
    const int32_t ub = (int32_t)f2u(u);
    const int32_t ne = ((int32_t)0x3FB504F3 - ub) >> 23;
    const float m = ldexpf(u, ne);
    const float h = fmaf(m, 2.0f, -2.0f);
    float q = 0x1.15425cp-14f;
    q = fmaf(q, h, -0x1.d962f2p-13f);
    q = fmaf(q, h, 0x1.e3d33ap-12f);
    q = fmaf(q, h, -0x1.fcc1aep-11f);
    q = fmaf(q, h, 0x1.237a56p-9f);
    q = fmaf(q, h, -0x1.5559e6p-8f);
    q = fmaf(q, h, 0x1.99a13p-7f);
    q = fmaf(q, h, -0x1.000008p-5f);
    q = fmaf(q, h, 0x1.55554ep-4f);
    q = fmaf(q, h, -0x1p-2f);
    const float h2 = h * h;
    const float lnm2 = fmaf(h2, q, h);
    const float nef = (float)ne;
    return fmaf(nef, -0x1.62e43p+0f, fmaf(nef, 0x1.05c61p-28f, lnm2));
}

static void normal_pair_device(uint32_t wa, uint32_t wb, float out[2]) {
    const float u1 = fmaf((float)wa, 0x1p-32f, 0x1p-33f);
    const float r = sqrtf(neg(two_ln_device(u1)));
    const int32_t gi = (int32_t)(wb & 0x3FFFFFFFu) - (int32_t)0x20000000;
    const float g = (float)gi * 0x1p-30f;
    const float x = g * g;
    float sp = 0x1.d53624p-14f;
    sp = fmaf(sp, x, -0x1.b1de6p-9f);
    sp = fmaf(sp, x, 0x1.cda0f6p-5f);
    sp = fmaf(sp, x, -0x1.d3ba5cp-2f);
    sp = fmaf(sp, x, 0x1.1c5832p+0f);
    float cp = -0x1.26d016p-16f;
    cp = fmaf(cp, x, 0x1.54c234p-11f);
    cp = fmaf(cp, x, -0x1.e36a9ep-7f);
    cp = fmaf(cp, x, 0x1.6f5a4ap-3f);
    cp = fmaf(cp, x, -0x1.bea5b6p-1f);
    cp = fmaf(cp, x, 0x1.6a09e6p-1f);
    const float s = g * sp;
    const uint32_t t = wb << 1;
    const float sf = u2f(f2u(s) ^ (~t & 0x80000000u));
    const float a = cp + sf;
    const float b = cp - sf;
    const float r1 = u2f(f2u(r) ^ (wb & 0x80000000u));
    const float r0 = u2f(f2u(r1) ^ (t & 0x80000000u));
    out[0] = r0 * a;
    out[1] = r1 * b;
}

int main(void) {
    long radius_bad = 0, radius_n = 0;
    for (uint32_t ub = f2u(0x1p-33f); ub <= f2u(1.0f); ++ub) {
        const float u = u2f(ub);
        if (f2u(neg(two_ln_device(u))) != f2u(-2.0f * ln_poly(u))) ++radius_bad;
        ++radius_n;
    }
    const uint32_t was[3] = {0xFFFFFFFFu /* u1 = 1, r = -0 */, 0u, 0x9E3779B9u};
    long turn_bad = 0, turn_n = 0;
    for (uint32_t low = 0; low < (1u << 30); low += (low < (1u << 21) || low >= (1u << 30) - (1u << 21)) ? 1u : 1021u) {
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t wb = (q << 30) | low;
            for (int k = 0; k < 3; ++k) {
                float o[2], d[2];
                jo_normal_pair(was[k], wb, o);
                normal_pair_device(was[k], wb, d);
                if (f2u(o[0]) != f2u(d[0]) || f2u(o[1]) != f2u(d[1])) ++turn_bad;
                ++turn_n;
            }
        }
    }
    printf("radius %ld/%ld differ, turn %ld/%ld differ\n", radius_bad, radius_n, turn_bad, turn_n);
    return (radius_bad || turn_bad) ? 1 : 0;
}
