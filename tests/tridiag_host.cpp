// tridiag_host.cpp -- the eigenvalue solves of jne_tridiag.h built for the
// host, for test_tridiag.py.
//   eig(n, a, out): Householder + implicit QL (the device's solve for dims
//     7-12); a is the packed upper triangle of an n x n symmetric matrix
//     (row-major, i <= j), out its n eigenvalues.
//   sturm_eigs(n, d, e, out): bisection on Sturm counts (the device's solve of
//     the tridiagonal form for dims 13-32, one eigenvalue per lane); d, e the
//     diagonal and off-diagonal of an n x n tridiagonal matrix, n <= 32, out
//     its eigenvalues in ascending order.
#include <cmath>
#include <cstring>

using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::sqrt;

#include "jne_tridiag.h"

namespace {
struct HostMath {
  static double sqrt(double x) { return std::sqrt(x); }
  static double rcp(double x) { return 1.0 / x; }
  static double rsqrt(double x) { return 1.0 / std::sqrt(x); }
};
template <int N>
void run(const double* a_in, double* out) {
  double a[N * (N + 1) / 2];
  std::memcpy(a, a_in, sizeof a);
  jne::tridiag_eigvals<N, HostMath>(a, out);
}
}  // namespace

extern "C" int eig(int n, const double* a, double* out) {
  switch (n) {
    case 1: run<1>(a, out); return 0;
    case 2: run<2>(a, out); return 0;
    case 3: run<3>(a, out); return 0;
    case 4: run<4>(a, out); return 0;
    case 5: run<5>(a, out); return 0;
    case 6: run<6>(a, out); return 0;
    case 7: run<7>(a, out); return 0;
    case 8: run<8>(a, out); return 0;
    case 9: run<9>(a, out); return 0;
    case 10: run<10>(a, out); return 0;
    case 11: run<11>(a, out); return 0;
    case 12: run<12>(a, out); return 0;
  }
  return 1;
}

extern "C" int sturm_eigs(int n, const double* d, const double* e, double* out) {
  constexpr int NMAX = 32;
  if (n < 1 || n > NMAX) return 1;
  double dd[NMAX] = {}, e2[NMAX] = {};
  for (int i = 0; i < n; ++i) dd[i] = d[i];
  for (int i = 0; i + 1 < n; ++i) e2[i] = e[i] * e[i];
  double lo, hi, pivmin;
  jne::sturm_bounds<NMAX>(dd, e2, n, lo, hi, pivmin);
  for (int k = 0; k < n; ++k)
    out[k] = jne::sturm_eigval<NMAX, HostMath>(dd, e2, n, k, lo, hi, pivmin, jne::kBisectIters);
  return 0;
}
