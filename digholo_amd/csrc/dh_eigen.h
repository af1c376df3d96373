// FP64 Hermitian spectrum helpers (host).
#pragma once
#include <complex>
#include <vector>

namespace dh {

struct HermitianSpectrum {
  double min, max, trace;
};

// A: n x n row-major Hermitian (overwritten).  Returns the smallest and largest
// eigenvalues and the trace.
HermitianSpectrum hermitian_extremes(int n, std::vector<std::complex<double>>& A);

// Householder reflector H = I - tau v v^H acting on entries k+1 .. k+len(v).
struct Reflector {
  int k;
  double tau;
  std::vector<std::complex<double>> v;
};

// Smallest and largest eigenvalues of an n x n row-major Hermitian matrix
// (overwritten) with unit eigenvectors: tridiagonal reduction keeping the
// reflectors, inverse iteration on the real tridiagonal, back transformation.
struct HermitianExtremePairs {
  double lmin = 0, lmax = 0;
  std::vector<std::complex<double>> vmin, vmax;
};
HermitianExtremePairs hermitian_extreme_pairs(int n, std::vector<std::complex<double>>& A);

// Reduction to real symmetric tridiagonal (d, |e|) with the same eigenvalues;
// the _q form also returns the reflectors and the complex subdiagonal.
void hermitian_tridiagonal_q(int n, std::vector<std::complex<double>>& A, std::vector<double>& d,
                             std::vector<double>& e, std::vector<Reflector>* refl,
                             std::vector<std::complex<double>>* sub);
void hermitian_tridiagonal(int n, std::vector<std::complex<double>>& A, std::vector<double>& d,
                           std::vector<double>& e);
double tridiagonal_eigenvalue(const std::vector<double>& d, const std::vector<double>& e, int k);

}  // namespace dh
