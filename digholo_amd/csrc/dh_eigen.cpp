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

// trailing blocks at least this tall are reduced with the rows spread over
// the host's threads (OpenMP; OMP_NUM_THREADS sets how many)
constexpr int kParallelRows = 96;

void hermitian_tridiagonal(int n, std::vector<cd>& A, std::vector<double>& d, std::vector<double>& e) {
  hermitian_tridiagonal_q(n, A, d, e, nullptr, nullptr);
}

void hermitian_tridiagonal_q(int n, std::vector<cd>& A, std::vector<double>& d, std::vector<double>& e,
                             std::vector<Reflector>* refl, std::vector<cd>* sub) {
  if (refl) refl->clear();
  if (sub) sub->assign(n > 1 ? n - 1 : 0, cd(0, 0));
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
    if (refl) refl->push_back(Reflector{k, tau, std::vector<cd>(v.begin(), v.begin() + L)});
    if (sub) (*sub)[k] = -ph * alpha;
    // p = tau * B v, B the trailing (L x L) block (rows in parallel on the
    // host cores once the block is large enough to pay for the threads)
#pragma omp parallel for schedule(static) if (L >= kParallelRows)
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
#pragma omp parallel for schedule(static) if (L >= kParallelRows)
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
    if (sub) (*sub)[n - 2] = A[(size_t)(n - 1) * n + n - 2];
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

// Solution of (T - lambda I) y = b for the real symmetric tridiagonal T
// (diagonal d, off-diagonal e) by Gaussian elimination with partial pivoting;
// a zero pivot is replaced by a tiny one (lambda is an eigenvalue).
static void tridiagonal_shifted_solve(const std::vector<double>& d, const std::vector<double>& e, double lambda,
                                      double tiny, std::vector<double>& b) {
  const int n = (int)d.size();
  std::vector<double> dd(n), dl(n > 1 ? n - 1 : 0), du(n > 1 ? n - 1 : 0), du2(n > 2 ? n - 2 : 0, 0.0);
  std::vector<char> swp(n > 1 ? n - 1 : 0, 0);
  for (int i = 0; i < n; ++i) dd[i] = d[i] - lambda;
  for (int i = 0; i + 1 < n; ++i) dl[i] = du[i] = e[i];
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      if (dd[i] == 0.0) dd[i] = tiny;
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      swp[i] = 1;
    }
  }
  if (n > 0 && dd[n - 1] == 0.0) dd[n - 1] = tiny;
  for (int i = 0; i + 1 < n; ++i) {
    if (swp[i]) std::swap(b[i], b[i + 1]);
    b[i + 1] -= dl[i] * b[i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    if (i + 1 < n) s -= du[i] * b[i + 1];
    if (i + 2 < n) s -= du2[i] * b[i + 2];
    b[i] = s / dd[i];
  }
}

HermitianExtremePairs hermitian_extreme_pairs(int n, std::vector<cd>& A) {
  HermitianExtremePairs r;
  if (n <= 0) return r;
  std::vector<double> d, e;
  std::vector<Reflector> refl;
  std::vector<cd> sub;
  hermitian_tridiagonal_q(n, A, d, e, &refl, &sub);
  r.lmin = tridiagonal_eigenvalue(d, e, 0);
  r.lmax = tridiagonal_eigenvalue(d, e, n - 1);
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = std::max(scale, std::fabs(d[i]) + (i ? e[i - 1] : 0.0) + (i + 1 < n ? e[i] : 0.0));
  const double tiny = std::max(scale, std::numeric_limits<double>::min()) * 1e-15;
  // phases taking the real tridiagonal (d, |t|) back to the complex one
  std::vector<cd> ph(n, cd(1, 0));
  for (int k = 0; k + 1 < n; ++k) {
    const double a = std::abs(sub[k]);
    ph[k + 1] = a > 0 ? ph[k] * (sub[k] / a) : ph[k];
  }
  auto vec = [&](double lambda) {
    std::vector<double> y(n, 1.0 / std::sqrt((double)n));
    for (int it = 0; it < 3; ++it) {
      tridiagonal_shifted_solve(d, e, lambda, tiny, y);
      double nrm = 0;
      for (double v : y) nrm += v * v;
      nrm = std::sqrt(nrm);
      if (!(nrm > 0) || !std::isfinite(nrm)) break;
      for (double& v : y) v /= nrm;
    }
    std::vector<cd> x(n);
    for (int i = 0; i < n; ++i) x[i] = ph[i] * y[i];
    for (int k = (int)refl.size() - 1; k >= 0; --k) {   // x = H_0 H_1 ... H_{n-3} x
      const Reflector& R = refl[k];
      const int L = (int)R.v.size();
      cd s = 0;
      for (int i = 0; i < L; ++i) s += std::conj(R.v[i]) * x[R.k + 1 + i];
      s *= R.tau;
      for (int i = 0; i < L; ++i) x[R.k + 1 + i] -= R.v[i] * s;
    }
    return x;
  };
  r.vmin = vec(r.lmin);
  r.vmax = vec(r.lmax);
  return r;
}

}  // namespace dh
