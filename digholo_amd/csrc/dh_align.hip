// Kernels of the AutoAlign parameter estimate (digHolo.cpp:11517-11930): the
// tilt search of the batch-summed Fourier-plane intensity, a DCT-domain
// Gaussian low-pass (pattern match against the Fourier window), followed by the
// masked peak search.
//
// The 2-D transforms are the unnormalised real-to-real transforms the
// reference plans with FFTW (:7393-7420):
//   type 10 (DCT-II, REDFT10):  Y_k = 2 sum_n X_n cos(pi (2n+1) k / 2N)
//   type 01 (DCT-III, REDFT01): Y_k = X_0 + 2 sum_{n>=1} X_n cos(pi n (2k+1) / 2N)
// evaluated as direct sums in FP64 (one map per polarisation, at most a few
// hundred thousand points, once per AutoAlign call), with the cosines read
// from a table of cos(pi m / 2N), m < 4N, so every angle is reduced exactly.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "dh_internal.h"

namespace dh {
namespace {

constexpr int kSThreads = 256;

__global__ void k_cos_table(int N, double* __restrict__ tab) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < 4 * N) tab[m] = cospi((double)m / (2.0 * N));
}

// One 1-D transform along x (axis 0) or y (axis 1) of an Ny x Nx map.
template <typename TIn>
__global__ __launch_bounds__(kSThreads) void k_r2r_axis(const TIn* __restrict__ in, double* __restrict__ out, int Ny,
                                                        int Nx, int axis, int type, const double* __restrict__ ctab) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= Ny * Nx) return;
  const int j = idx / Nx, i = idx - j * Nx;
  const int N = axis ? Ny : Nx;
  const int k = axis ? j : i;
  const size_t base = axis ? (size_t)i : (size_t)j * Nx;
  const size_t stride = axis ? (size_t)Nx : 1;
  const int N4 = 4 * N;
  double s = 0.0;
  if (type == 10) {
    int m = k % N4;              // (2n + 1) k mod 4N, stepped by 2k
    const int step = (2 * k) % N4;
    for (int n = 0; n < N; ++n) {
      s += (double)in[base + n * stride] * ctab[m];
      m += step;
      if (m >= N4) m -= N4;
    }
    s *= 2.0;
  } else {
    const int step = (2 * k + 1) % N4;   // n (2k + 1) mod 4N
    int m = step;
    for (int n = 1; n < N; ++n) {
      s += (double)in[base + n * stride] * ctab[m];
      m += step;
      if (m >= N4) m -= N4;
    }
    s = (double)in[base] + 2.0 * s;
  }
  out[idx] = s;
}

// Gaussian filter of the DCT-domain map.  The reference walks the map in
// blocks of 8 floats with the loop counters of a (8 floor(Nx/8)) x Ny array
// (:11636-11660) while the map's rows are Nx long, so element f of the flat
// map (f < Ny * NxR, NxR = 8 floor(Nx/8)) is weighted with the counters
// j = f / NxR, c = f % NxR, and the last Ny * (Nx - NxR) elements keep their
// value.  The exponent is formed in float as there; exp in FP64.
__global__ void k_snap_gauss(double* __restrict__ m, int Ny, int Nx, float px, float wx2, float wy2) {
  const int NxR = 8 * (Nx / 8);
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= Ny * NxR) return;
  const int j = f / NxR, c = f - j * NxR;
  const float px2 = __fmul_rn(px, px);
  const float y2 = __fmul_rn((float)(j * j), px2);
  const float yw2 = __fmul_rn(y2, wy2);
  const float x = __fmul_rn((float)c, px);
  const float xw2 = __fmul_rn(__fmul_rn(x, x), wx2);
  m[f] *= exp((double)__fadd_rn(xw2, yw2));
}

// Largest value of the filtered map outside the zero-order disc
// kx^2 + ky^2 > ad (float, as :11672-11695); ties go to the first pixel in
// row-major order (the reference's strict comparison).  One workgroup.
__global__ __launch_bounds__(1024) void k_snap_argmax(const double* __restrict__ m, int Ny, int Nx,
                                                      const float* __restrict__ kx, const float* __restrict__ ky,
                                                      float ad, int* __restrict__ out) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  double best = -DBL_MAX;
  int bi = -1;
  for (int f = threadIdx.x; f < Ny * Nx; f += blockDim.x) {
    const int j = f / Nx, i = f - j * Nx;
    const float kr = __fadd_rn(__fmul_rn(kx[i], kx[i]), __fmul_rn(ky[j], ky[j]));
    if (kr > ad) {
      const double v = m[f];
      if (v > best) {   // f increases within a thread: the first maximum stays
        best = v;
        bi = f;
      }
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double v2 = sv[threadIdx.x + s];
      const int i2 = si[threadIdx.x + s];
      const double v1 = sv[threadIdx.x];
      const int i1 = si[threadIdx.x];
      if (i2 >= 0 && (i1 < 0 || v2 > v1 || (v2 == v1 && i2 < i1))) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = si[0];
}

inline int blocks(size_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

size_t snap_work_elems(int Ny, int Nx) { return 2 * (size_t)Ny * Nx + 4 * (size_t)(Ny + Nx); }

hipError_t launch_snap_tilt_search(const float* map, int Ny, int Nx, float px, float wx2, float wy2, const float* kx,
                                   const float* ky, float ad, double* work, int* outIdx, hipStream_t s) {
  if (Ny <= 0 || Nx <= 0) return hipErrorInvalidValue;
  const size_t N = (size_t)Ny * Nx;
  double* a = work;
  double* b = work + N;
  double* tx = work + 2 * N;
  double* ty = tx + 4 * (size_t)Nx;
  k_cos_table<<<blocks(4 * (size_t)Nx, kSThreads), kSThreads, 0, s>>>(Nx, tx);
  k_cos_table<<<blocks(4 * (size_t)Ny, kSThreads), kSThreads, 0, s>>>(Ny, ty);
  const int nb = blocks(N, kSThreads);
  k_r2r_axis<float><<<nb, kSThreads, 0, s>>>(map, a, Ny, Nx, 0, 10, tx);
  k_r2r_axis<double><<<nb, kSThreads, 0, s>>>(a, b, Ny, Nx, 1, 10, ty);
  k_snap_gauss<<<nb, kSThreads, 0, s>>>(b, Ny, Nx, px, wx2, wy2);
  k_r2r_axis<double><<<nb, kSThreads, 0, s>>>(b, a, Ny, Nx, 0, 1, tx);
  k_r2r_axis<double><<<nb, kSThreads, 0, s>>>(a, b, Ny, Nx, 1, 1, ty);
  k_snap_argmax<<<1, 1024, 0, s>>>(b, Ny, Nx, kx, ky, ad, outIdx);
  return hipGetLastError();
}

}  // namespace dh
