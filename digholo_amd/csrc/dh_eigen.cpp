// FP64 spectrum of a Hermitian matrix, for the singular values behind the
// mode-dependent-loss metric (the reference calls LAPACK cgesvd,
// digHolo.cpp:3021-3059; here the Gram matrix comes from the GPU).
//
// Householder reduction to real symmetric tridiagonal form (unitary
// similarity, eigenvalues unchanged), then Sturm-sequence bisection for the
// requested eigenvalues.
#include "dh_eigen.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

namespace dh {

using cd = std::complex<double>;

void hermitian_tridiagonal(int n, std::vector<cd>& A, std::vector<double>& d, std::vector<double>& e) {
  d.assign(n, 0.0);
  e.assign(n > 1 ? n - 1 : 0, 0.0);
  std::vector<cd> v(n), p(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;            // entries below the diagonal in column k
    double alpha2 = 0;
    for (int i = 0; i < L; ++i) alpha2 += std::norm(A[(size_t)(k + 1 + i) * n + k]);
    const double alpha = std::sqrt(alpha2);
    d[k] = A[(size_t)k * n + k].real();
    if (alpha == 0.0) {
      e[k] = 0.0;
      continue;
    }
    const cd x0 = A[(size_t)(k + 1) * n + k];
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1, 0);
    for (int i = 0; i < L; ++i) v[i] = A[(size_t)(k + 1 + i) * n + k];
    v[0] += ph * alpha;                 // H x = -ph * alpha * e1
    double vn = 0;
    for (int i = 0; i < L; ++i) vn += std::norm(v[i]);
    const double tau = 2.0 / vn;
    // p = tau * B v, B the trailing (L x L) block
    for (int i = 0; i < L; ++i) {
      cd s = 0;
      const cd* row = &A[(size_t)(k + 1 + i) * n + k + 1];
      for (int j = 0; j < L; ++j) s += row[j] * v[j];
      p[i] = tau * s;
    }
    cd vp = 0;
    for (int i = 0; i < L; ++i) vp += std::conj(v[i]) * p[i];
    const cd K = 0.5 * tau * vp;
    for (int i = 0; i < L; ++i) p[i] -= K * v[i];          // w
    // B <- B - v w^H - w v^H
    for (int i = 0; i < L; ++i) {
      cd* row = &A[(size_t)(k + 1 + i) * n + k + 1];
      const cd vi = v[i], wi = p[i];
      for (int j = 0; j < L; ++j) row[j] -= vi * std::conj(p[j]) + wi * std::conj(v[j]);
    }
    e[k] = alpha;                        // |subdiagonal|; its phase is a diagonal similarity
  }
  if (n >= 2) {
    d[n - 2] = A[(size_t)(n - 2) * n + n - 2].real();
    d[n - 1] = A[(size_t)(n - 1) * n + n - 1].real();
    e[n - 2] = std::abs(A[(size_t)(n - 1) * n + n - 2]);
  } else if (n == 1) {
    d[0] = A[0].real();
  }
}

// number of eigenvalues of the tridiagonal (d, e) that are < x
static int sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x) {
  const int n = (int)d.size();
  int cnt = 0;
  double q = 1.0;
  const double tiny = std::numeric_limits<double>::min() * 4;
  for (int i = 0; i < n; ++i) {
    const double ee = i ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i ? ee / q : 0.0);
    if (q == 0.0) q = -tiny;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// k-th smallest eigenvalue (k = 0 .. n-1) by bisection
double tridiagonal_eigenvalue(const std::vector<double>& d, const std::vector<double>& e, int k) {
  const int n = (int)d.size();
  double lo = std::numeric_limits<double>::max(), hi = -lo;
  for (int i = 0; i < n; ++i) {        // Gershgorin
    const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double scale = std::max(std::fabs(lo), std::fabs(hi));
  lo -= 1e-12 * scale + std::numeric_limits<double>::min();
  hi += 1e-12 * scale + std::numeric_limits<double>::min();
  for (int it = 0; it < 200 && hi - lo > 2e-16 * scale; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (sturm_count(d, e, mid) > k) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

HermitianSpectrum hermitian_extremes(int n, std::vector<cd>& A) {
  HermitianSpectrum s{0, 0, 0};
  if (n <= 0) return s;
  for (int i = 0; i < n; ++i) s.trace += A[(size_t)i * n + i].real();
  std::vector<double> d, e;
  hermitian_tridiagonal(n, A, d, e);
  s.min = tridiagonal_eigenvalue(d, e, 0);
  s.max = tridiagonal_eigenvalue(d, e, n - 1);
  return s;
}

}  // namespace dh
