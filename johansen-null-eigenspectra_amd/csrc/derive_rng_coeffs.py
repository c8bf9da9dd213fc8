"""Derive the float32 polynomial coefficients of the normal generator.

The device normal generator turns two Philox4x32-10 words into two N(0,1)
floats by the Box-Muller transform.  To make the device stream bit-identical
to a CPU implementation, the transform uses only IEEE-754 operations that are
correctly rounded on both sides (add, mul, fma, sqrt, int<->float conversion)
and no library transcendental.  The three transcendental pieces are replaced
by polynomials whose coefficients this script derives:

  ln(1+f) = f + f^2 * Q(f)            f in [sqrt(1/2)-1, sqrt(2)-1]
  sin(pi*g/2) / sqrt(2) = g * S(g^2)  g in [-1/2, 1/2]
  cos(pi*g/2) / sqrt(2) = C(g^2)      g in [-1/2, 1/2]

Each polynomial is a Chebyshev interpolant of the target function (near
minimax), converted to the monomial basis and rounded to float32.  The script
prints the coefficients as C hex-float literals and the worst error of the
float32 Horner evaluation against numpy's float64 reference on a dense grid.

Run:  python derive_rng_coeffs.py
"""
import numpy as np
from numpy.polynomial import chebyshev as C
from numpy.polynomial import polynomial as P


def cheb_monomial(fn, lo, hi, deg):
    """Chebyshev interpolant of fn on [lo, hi], returned as monomial coeffs."""
    # interpolate on the canonical interval then map back to x
    cheb = C.Chebyshev.interpolate(fn, deg, domain=[lo, hi])
    mono = cheb.convert(kind=P.Polynomial, domain=[lo, hi], window=[lo, hi])
    return mono.coef.astype(np.float64)


def horner_f32(coef32, x32):
    """Float32 Horner with a fused multiply-add per step (emulated in f64)."""
    acc = np.full_like(x32, coef32[-1], dtype=np.float32)
    for c in coef32[-2::-1]:
        acc = (acc.astype(np.float64) * x32.astype(np.float64) + np.float64(c)).astype(np.float32)
    return acc


def hexf(v):
    mant, exp = float(np.float32(v)).hex().split("p")
    if "." in mant:
        mant = mant.rstrip("0").rstrip(".")
    return f"{mant}p{exp}f"


def main():
    # ---- Q(f) = (ln(1+f) - f) / f^2 ----
    lo, hi = np.sqrt(0.5) - 1.0, np.sqrt(2.0) - 1.0

    def q_fn(f):
        f = np.asarray(f, dtype=np.float64)
        out = np.empty_like(f)
        small = np.abs(f) < 1e-4
        fs = f[small]
        out[small] = -0.5 + fs / 3 - fs**2 / 4 + fs**3 / 5
        fb = f[~small]
        out[~small] = (np.log1p(fb) - fb) / fb**2
        return out

    qdeg = 9
    q = cheb_monomial(q_fn, lo, hi, qdeg).astype(np.float32)
    f = np.linspace(lo, hi, 2_000_001).astype(np.float32)
    f = f[(f >= lo) & (f <= hi)]
    f2 = (f.astype(np.float64) * f).astype(np.float32)
    qv = horner_f32(q, f)
    ln32 = (f2.astype(np.float64) * qv + f).astype(np.float32)
    ref = np.log1p(f.astype(np.float64))
    nz = ref != 0
    relerr = np.max(np.abs(ln32[nz] - ref[nz]) / np.abs(ref[nz]))
    print(f"// ln(1+f) = f + f^2*Q(f), deg {qdeg}, max rel err {relerr:.3e}")
    print("Q =", ", ".join(hexf(c) for c in q))

    # ---- S(x) = sin(pi*sqrt(x)/2)/(sqrt(2)*sqrt(x)), x = g^2 in [0, 1/4] ----
    def s_fn(x):
        x = np.asarray(x, dtype=np.float64)
        g = np.sqrt(np.maximum(x, 1e-300))
        out = np.where(x < 1e-12, np.pi / 2 / np.sqrt(2.0), np.sin(np.pi * g / 2) / g / np.sqrt(2.0))
        return out

    def c_fn(x):
        x = np.asarray(x, dtype=np.float64)
        return np.cos(np.pi * np.sqrt(x) / 2) / np.sqrt(2.0)

    sdeg, cdeg = 4, 5
    s = cheb_monomial(s_fn, 0.0, 0.25, sdeg).astype(np.float32)
    c = cheb_monomial(c_fn, 0.0, 0.25, cdeg).astype(np.float32)
    g = np.linspace(-0.5, 0.5, 2_000_001).astype(np.float32)
    x = (g.astype(np.float64) * g).astype(np.float32)
    sv = (g.astype(np.float64) * horner_f32(s, x)).astype(np.float32)
    cv = horner_f32(c, x)
    sref = np.sin(np.pi * g.astype(np.float64) / 2) / np.sqrt(2.0)
    cref = np.cos(np.pi * g.astype(np.float64) / 2) / np.sqrt(2.0)
    print(f"// sin(pi g/2)/sqrt2 = g*S(g^2), deg {sdeg}, max abs err {np.max(np.abs(sv - sref)):.3e}")
    print("S =", ", ".join(hexf(v) for v in s))
    print(f"// cos(pi g/2)/sqrt2 = C(g^2), deg {cdeg}, max abs err {np.max(np.abs(cv - cref)):.3e}")
    print("C =", ", ".join(hexf(v) for v in c))
    ln2 = np.float32(np.log(2.0))
    ln2lo = np.float32(np.log(2.0) - np.float64(ln2))
    print("LN2_HI =", hexf(ln2), " LN2_LO =", hexf(ln2lo))


if __name__ == "__main__":
    main()
