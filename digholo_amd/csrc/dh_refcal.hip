// Reference-wave calibration plane (digHolo.cpp:6729-6850, :8253-8292 and the
// complex-to-complex branch of the window copy, :1808-2122 with r2c = 0).
//
// The stored calibration frame of each (wavelength, pol) is cropped to the
// pol's FFT window, transformed (forward, unnormalised), its Fourier window
// centred on zero frequency is weighted with mask * fill-factor envelope *
// norm, placed in the field plane and inverse transformed.  The transforms are
// evaluated once per configuration as separable direct DFTs restricted to the
// window's rows and columns (FP64 accumulation, twiddles from a table of
// exp(-2 pi i m / N), m < N, so every angle is reduced exactly).
#include <hip/hip_runtime.h>

#include "dh_internal.h"

namespace dh {
namespace {

constexpr int kCThreads = 256;

__global__ void k_twiddle_f64(int N, double2* __restrict__ tw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  double s, c;
  sincospi(-2.0 * (double)m / (double)N, &s, &c);
  tw[m] = make_double2(c, s);
}

// out(o, k) = sum_{n < N} in(o, n) exp(sign 2 pi i f_k n / N)
__global__ __launch_bounds__(kCThreads) void k_cdft_axis(const float2* __restrict__ in, size_t inO, size_t inN,
                                                         float2* __restrict__ out, size_t outO, size_t outK,
                                                         int nOuter, int N, int K, const int* __restrict__ fk,
                                                         int sign, const double2* __restrict__ tw) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nOuter * K) return;
  const int o = (int)(idx / K), k = (int)(idx - (size_t)o * K);
  int f = fk ? fk[k] % N : k;
  if (f < 0) f += N;
  const float2* src = in + (size_t)o * inO;
  double ar = 0.0, ai = 0.0;
  int m = 0;
  for (int n = 0; n < N; ++n) {
    const double2 t = tw[m];
    const double ti = sign > 0 ? -t.y : t.y;
    const float2 v = src[(size_t)n * inN];
    ar += (double)v.x * t.x - (double)v.y * ti;
    ai += (double)v.x * ti + (double)v.y * t.x;
    m += f;
    if (m >= N) m -= N;
  }
  out[(size_t)o * outO + (size_t)k * outK] = make_float2((float)ar, (float)ai);
}

// Field-plane input of the inverse transform: window element (j, i) of Y times
// its weight lands at row (j + h/2) mod h, column (i + w/2) mod w (low
// resolution) or at its own frequency, row (j + SY + H/2) mod H, column
// (i + SX + W/2) mod W (full resolution); everything else is zero.
__global__ void k_cal_place(const float2* __restrict__ Y, const float* __restrict__ wt, int w, int h, int fw, int fh,
                            int lowRes, int SX, int SY, int W, int H, float2* __restrict__ plane) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= fw * fh) return;
  const int r = e / fw, col = e - r * fw;
  int j, i;
  if (lowRes) {
    j = (r + h - h / 2) % h;
    i = (col + w - w / 2) % w;
  } else {
    j = ((r - SY - H / 2) % H + H) % H;
    i = ((col - SX - W / 2) % W + W) % W;
  }
  float2 v = make_float2(0.f, 0.f);
  if (j < h && i < w) {
    const float a = wt[j * w + i];
    const float2 y = Y[j * w + i];
    v = make_float2(y.x * a, y.y * a);
  }
  plane[e] = v;
}

// a[q][e] *= b[q % nb][e] over nq planes of n elements
__global__ void k_cmul_planes(float2* __restrict__ a, const float2* __restrict__ b, size_t n, int nq, int nb) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nq) return;
  const size_t q = e / n;
  const float2 x = a[e], y = b[(q % nb) * n + (e - q * n)];
  a[e] = make_float2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}

__global__ void k_zero2(float2* __restrict__ a, size_t n) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = make_float2(0.f, 0.f);
}

inline unsigned blocks(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_zero_complex(float2* a, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_zero2<<<blocks(n, kCThreads), kCThreads, 0, s>>>(a, n);
  return hipGetLastError();
}

hipError_t launch_twiddle_f64(int N, double2* tw, hipStream_t s) {
  k_twiddle_f64<<<blocks(N, kCThreads), kCThreads, 0, s>>>(N, tw);
  return hipGetLastError();
}

hipError_t launch_cdft_axis(const float2* in, size_t inO, size_t inN, float2* out, size_t outO, size_t outK,
                            int nOuter, int N, int K, const int* fk, int sign, const double2* tw, hipStream_t s) {
  if (nOuter <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
  k_cdft_axis<<<blocks((size_t)nOuter * K, kCThreads), kCThreads, 0, s>>>(in, inO, inN, out, outO, outK, nOuter, N,
                                                                           K, fk, sign, tw);
  return hipGetLastError();
}

hipError_t launch_cal_place(const float2* Y, const float* wt, int w, int h, int fw, int fh, int lowRes, int SX, int SY,
                            int W, int H, float2* plane, hipStream_t s) {
  k_cal_place<<<blocks((size_t)fw * fh, kCThreads), kCThreads, 0, s>>>(Y, wt, w, h, fw, fh, lowRes, SX, SY, W, H,
                                                                       plane);
  return hipGetLastError();
}

hipError_t launch_cmul_planes(float2* a, const float2* b, size_t n, int nq, int nb, hipStream_t s) {
  if (n == 0 || nq <= 0 || nb <= 0) return hipSuccess;
  k_cmul_planes<<<blocks(n * nq, kCThreads), kCThreads, 0, s>>>(a, b, n, nq, nb);
  return hipGetLastError();
}

}  // namespace dh
