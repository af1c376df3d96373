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

// Reduction to real symmetric tridiagonal (d, |e|) with the same eigenvalues.
void hermitian_tridiagonal(int n, std::vector<std::complex<double>>& A, std::vector<double>& d,
                           std::vector<double>& e);
double tridiagonal_eigenvalue(const std::vector<double>& d, const std::vector<double>& e, int k);

}  // namespace dh
