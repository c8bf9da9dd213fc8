// jne_tridiag.h -- eigenvalues of small symmetric matrices.
//
// tridiag_eigvals: a matrix held in one thread's registers, by Householder
// reduction to tridiagonal form, then the implicit QL iteration with the shift
// from the leading 2x2 block (the tql1 scheme of the EISPACK/Numerical Recipes
// family), eigenvalues only.  sturm_*: the Gershgorin interval and the Sturm
// count of a tridiagonal matrix, from which the lanes of a wave find its
// eigenvalues one each (or a few lanes per eigenvalue); sturm_eigval is the
// plain bisection on them.
//
// Every array index is a compile-time constant once the loops are unrolled,
// so the matrix stays in registers: the QL sweep runs over the whole trailing
// range with a per-thread predicate instead of a data-dependent loop bound,
// and the split point m is selected, not indexed.  The iteration count per
// eigenvalue is bounded (kQlMaxIter), so every thread leaves the loop.
//
// Cost for N = 12: about 1.5e3 FMAs for the reduction and 25 operations per
// rotation, 2-3 sweeps per eigenvalue, against about 4e3 operations per sweep
// of cyclic Jacobi and 6-8 sweeps.  Accuracy: both steps are orthogonal
// similarity transforms, so the eigenvalues carry an absolute error of a few
// ulp of the matrix norm.
#pragma once

#if defined(__HIPCC__)
#define JNE_TRI_HD __device__ __forceinline__
#else
#define JNE_TRI_HD inline
#endif

namespace jne {

constexpr int kQlMaxIter = 30;

// Packed upper triangle of an N x N symmetric matrix: a[tri_idx<N>(i, j)].
template <int N>
JNE_TRI_HD constexpr int tri_idx(int i, int j) {
  return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j);
}

// Math supplies sqrt(x) for x >= 0, 1/x (x != 0) and 1/sqrt(x) (x > 0), each
// to a few ulp.
template <int N, class Math>
JNE_TRI_HD void tridiag_eigvals(double* a, double* d) {
  double e[N];  // e[i] couples d[i] and d[i + 1]; e[N - 1] = 0
  // ---- Householder: column k below the subdiagonal to zero, k = 0 .. N-3
#pragma unroll
  for (int k = 0; k < N - 2; ++k) {
    const double x0 = a[tri_idx<N>(k + 1, k)];
    double sigma = 0.0;
#pragma unroll
    for (int i = k + 2; i < N; ++i) sigma = fma(a[tri_idx<N>(i, k)], a[tri_idx<N>(i, k)], sigma);
    const double nrm = Math::sqrt(fma(x0, x0, sigma));
    const double alpha = x0 > 0.0 ? -nrm : nrm;
    // v = x - alpha e_1, |v|^2 = sigma + v0^2; tau = 2 / |v|^2 (0 when x = 0:
    // the reflection is the identity and the subdiagonal entry is 0)
    const double v0 = x0 - alpha;
    const double vn = fma(v0, v0, sigma);
    const double tau = vn > 0.0 ? 2.0 * Math::rcp(vn) : 0.0;
    double v[N], p[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = i == k + 1 ? v0 : (i > k + 1 ? a[tri_idx<N>(i, k)] : 0.0);
    // p = tau A v over the trailing block, K = tau/2 p.v, w = p - K v
    double pv = 0.0;
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = k + 1; j < N; ++j) s = fma(a[tri_idx<N>(i, j)], v[j], s);
      p[i] = tau * s;
      pv = fma(p[i], v[i], pv);
    }
    const double K = 0.5 * tau * pv;
#pragma unroll
    for (int i = k + 1; i < N; ++i) p[i] = fma(-K, v[i], p[i]);  // w
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
#pragma unroll
      for (int j = i; j < N; ++j) {
        double& aij = a[tri_idx<N>(i, j)];
        aij = fma(-v[i], p[j], fma(-p[i], v[j], aij));
      }
    }
    e[k] = tau > 0.0 ? alpha : x0;
  }
  e[N - 2] = a[tri_idx<N>(N - 1, N - 2)];
  e[N - 1] = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) d[i] = a[tri_idx<N>(i, i)];

  // ---- implicit QL, eigenvalue l converging at the top of block [l, m]
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    for (int iter = 0; iter < kQlMaxIter; ++iter) {
      // m: the first j >= l whose e[j] is negligible against its diagonal
      // neighbours (N - 1 if none)
      int m = N - 1;
#pragma unroll
      for (int j = N - 2; j >= l; --j) {
        const double dd = fabs(d[j]) + fabs(d[j + 1]);
        if (fabs(e[j]) <= 0x1p-53 * dd) m = j;
      }
      if (m == l) break;
      double dm = d[N - 1];
#pragma unroll
      for (int j = l + 1; j < N - 1; ++j) dm = m == j ? d[j] : dm;
      double g = (d[l + 1] - d[l]) * Math::rcp(2.0 * e[l]);
      double r = Math::sqrt(fma(g, g, 1.0));
      g = dm - d[l] + e[l] * Math::rcp(g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      // the sweep from i = m - 1 down to l, run over i = N - 2 .. l with the
      // steps above m - 1 predicated off.  A rotation of zero length (f = g =
      // 0: the block has split below l) ends the sweep: d[i + 1] takes its
      // shift back and the iteration restarts on the split blocks.
      bool split = false;
#pragma unroll
      for (int i = N - 2; i >= l; --i) {
        const bool act = i < m && !split;
        const double f = s * e[i], b = c * e[i];
        const double h = fma(f, f, g * g);
        const double ri = h > 0.0 ? Math::rsqrt(h) : 0.0;  // 1 / |(f, g)|
        const double rr = h * ri;                           // |(f, g)|
        const double sn = f * ri, cn = g * ri;
        const double g2 = d[i + 1] - p;
        const double r2 = fma(d[i] - g2, sn, 2.0 * cn * b);
        const double pn = sn * r2;
        if (act) {
          e[i + 1] = rr;
          split = !(rr > 0.0);
          s = sn;
          c = cn;
          p = split ? 0.0 : pn;
          d[i + 1] = split ? g2 : g2 + pn;
          g = fma(cn, r2, -b);
        }
      }
      if (!split) {
        d[l] -= p;
        e[l] = g;
      }
#pragma unroll
      for (int j = l + 1; j < N; ++j)
        if (m == j) e[j] = 0.0;
    }
  }
}

// ---- bisection on Sturm counts (the scheme of LAPACK's dstebz/dlaebz)
//
// The tridiagonal matrix T of order n <= NMAX has diagonal d[0..n-1] and
// squared off-diagonal e2[0..n-2] (e2[i] couples i and i + 1).  The pivots of
// T - xI = L D L^T, q_0 = d_0 - x and q_i = d_i - x - e2_{i-1} / q_{i-1}, have
// as many negative entries as T has eigenvalues below x.  A pivot q <= pivmin
// counts as negative and is replaced by min(q, -pivmin), which keeps every
// quotient finite; pivmin = (smallest normal) * max(1, max e2).  Bisection
// halves [lo, hi] (Gershgorin bounds, which hold every eigenvalue) on the
// count, so eigenvalue k (ascending) is found independently of the others.

constexpr double kSafeMin = 0x1p-1022;

// Gershgorin interval of T, widened by a few ulp of its norm, and pivmin.
template <int NMAX>
JNE_TRI_HD void sturm_bounds(const double* d, const double* e2, int n, double& lo, double& hi, double& pivmin) {
  double emax = 0.0, l = d[0], h = d[0];
  double ep = 0.0;  // |e_{i-1}|
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if (i < n) {
      const double en = i + 1 < n ? sqrt(e2[i]) : 0.0;
      const double r = ep + en;
      l = fmin(l, d[i] - r);
      h = fmax(h, d[i] + r);
      if (i + 1 < n) emax = fmax(emax, e2[i]);
      ep = en;
    }
  }
  const double tnorm = fmax(fabs(l), fabs(h));
  const double pad = 0x1p-50 * tnorm * (double)n;
  pivmin = kSafeMin * fmax(1.0, emax);
  lo = l - pad - 2.0 * pivmin;
  hi = h + pad + 2.0 * pivmin;
}

// Number of eigenvalues of T below x.  Math::rcp(x) is 1/x to a few ulp.
template <int NMAX, class Math>
JNE_TRI_HD int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  bool neg = q <= pivmin;
  int c = neg ? 1 : 0;
  q = neg ? fmin(q, -pivmin) : q;
#pragma unroll
  for (int i = 1; i < NMAX; ++i) {
    if (i < n) {
      q = fma(-e2[i - 1], Math::rcp(q), d[i] - x);
      neg = q <= pivmin;
      c += neg ? 1 : 0;
      q = neg ? fmin(q, -pivmin) : q;
    }
  }
  return c;
}

// Eigenvalue k (0-based, ascending) of T by `iters` halvings of [lo, hi] (the
// one-lane form; the wave-parallel solve runs the same counts at several
// points per round).
// From the Gershgorin interval (width <= 2 |T|), 54 halvings leave a width of
// at most 2^-53 |T|.
template <int NMAX, class Math>
JNE_TRI_HD double sturm_eigval(const double* d, const double* e2, int n, int k, double lo, double hi, double pivmin,
                               int iters) {
  for (int it = 0; it < iters; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_count<NMAX, Math>(d, e2, n, mid, pivmin) > k)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

constexpr int kBisectIters = 54;

}  // namespace jne
