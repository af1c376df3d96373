// Power-meter reductions over the coefficient matrix (digHolo.cpp:10714-11129).
//
// The coefficients of one batch form S matrices A_s (S = 1, or 2 with
// independent polarisations) of N = batch rows by M columns:
//   A_s[i][j] = coefs[i*rowStride + s*M + j].
// k_metric_rows: per row i (frame i exciting mode i), FP64 sums of the power on
//   the diagonal (j == i, or the same mode in the other polarisation), off the
//   diagonal, inside / outside the frame's HG mode group (weighted by
//   1/group as the reference does), and the row's total power.
// k_gram_f64: the Gram matrix whose eigenvalues are the squared singular values
//   of A_s, in FP64: A_s A_s^H (N x N) when N <= M, else A_s^H A_s (M x M).
#include <hip/hip_runtime.h>

#include "dh_internal.h"

namespace dh {

__device__ __forceinline__ int hg_group(int j) {
  // mode j belongs to group g with g(g+1)/2 <= j < (g+1)(g+2)/2
  int g = (int)((sqrtf(8.f * j + 1.f) - 1.f) * 0.5f);
  while ((g + 1) * (g + 2) / 2 <= j) ++g;
  while (g * (g + 1) / 2 > j) --g;
  return g;
}

// grid (rows, S), 256 threads; out[(s*N + i)*5 + {signal, noise, mgSignal, mgNoise, power}]
__global__ __launch_bounds__(256) void k_metric_rows(const float2* __restrict__ coefs, int N, int M, int Mpol,
                                                     int rowStride, double* __restrict__ out) {
  const int i = blockIdx.x, s = blockIdx.y;
  const float2* row = coefs + (size_t)i * rowStride + (size_t)s * M;
  const int gin = hg_group(i % Mpol) + 1;
  double sig = 0, noi = 0, mgs = 0, mgn = 0, pw = 0;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const float2 v = row[j];
    const float vp = v.x * v.x + v.y * v.y;          // float, like the reference's accumulands
    pw += vp;
    if (i == j || i == j - Mpol) sig += vp;
    else noi += vp;
    const int gout = hg_group(j % Mpol) + 1;
    if (gin == gout) mgs += (double)(vp / gin);
    else mgn += (double)(vp / gin);
  }
  __shared__ double red[5][256];
  red[0][threadIdx.x] = sig;
  red[1][threadIdx.x] = noi;
  red[2][threadIdx.x] = mgs;
  red[3][threadIdx.x] = mgn;
  red[4][threadIdx.x] = pw;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 5) out[((size_t)s * N + i) * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// Gram in FP64, 16x16 output tile per block.  rowsAreVectors: G[a][c] =
// sum_k A[a][k] conj(A[c][k]) over the rows of A (dimension N); otherwise over
// its columns (dimension M): G[a][c] = sum_i conj(A[i][a]) A[i][c].
__global__ __launch_bounds__(256) void k_gram_f64(const float2* __restrict__ coefs, int N, int M, int rowStride,
                                                  int s, int rowsAreVectors, double2* __restrict__ G) {
  const int dim = rowsAreVectors ? N : M;
  const int K = rowsAreVectors ? M : N;
  const int a0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  const int ta = threadIdx.x / 16, tc = threadIdx.x % 16;
  __shared__ double2 sa[16][17], sc[16][17];
  const float2* A = coefs + (size_t)s * M;
  double re = 0, im = 0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    // element (vector v, component k)
    {
      const int v = a0 + ta, k = k0 + tc;
      float2 x = make_float2(0.f, 0.f);
      if (v < dim && k < K) x = rowsAreVectors ? A[(size_t)v * rowStride + k] : A[(size_t)k * rowStride + v];
      sa[ta][tc] = make_double2(x.x, x.y);
      const int v2 = c0 + ta;
      float2 y = make_float2(0.f, 0.f);
      if (v2 < dim && k < K) y = rowsAreVectors ? A[(size_t)v2 * rowStride + k] : A[(size_t)k * rowStride + v2];
      sc[ta][tc] = make_double2(y.x, y.y);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const double2 x = sa[ta][k], y = sc[tc][k];
      if (rowsAreVectors) {   // x * conj(y)
        re += x.x * y.x + x.y * y.y;
        im += x.y * y.x - x.x * y.y;
      } else {                // conj(x) * y
        re += x.x * y.x + x.y * y.y;
        im += x.x * y.y - x.y * y.x;
      }
    }
    __syncthreads();
  }
  const int a = a0 + ta, c = c0 + tc;
  if (a < dim && c < dim) G[(size_t)a * dim + c] = make_double2(re, im);
}

// MDL term of the alignment-loss gradient: for the extreme singular triplet
// (sigma, u, v) of A_s, d sigma^2 / d conj(A) = sigma u v^H, so the seed of
// mu (ln s_max^2 - ln s_min^2) is coef conj(u[b]) v[j] with coef = +-mu/sigma
// (the conjugate because the adjoint takes dL/dc as conj(...)).  One of u, v is
// the Gram eigenvector x; the other is its image under A^H (rowsAreVectors, x
// of length N) or A (x of length M), over sigma.
// k_mdl_image: y[j] = sum_b conj(A[b][j]) x[b] / sig (rows), one thread per j;
// or y[b] = sum_j A[b][j] x[j] / sig (columns), one 256-thread block per b.
__global__ __launch_bounds__(256) void k_mdl_image(const float2* __restrict__ coefs, int N, int M, int rowStride,
                                                   int s, int rowsAreVectors, const double2* __restrict__ x,
                                                   double invSig, double2* __restrict__ y) {
  const float2* A = coefs + (size_t)s * M;
  if (rowsAreVectors) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    double re = 0, im = 0;
    for (int b = 0; b < N; ++b) {
      const float2 a = A[(size_t)b * rowStride + j];
      const double2 u = x[b];
      re += a.x * u.x + a.y * u.y;   // conj(a) u
      im += a.x * u.y - a.y * u.x;
    }
    y[j] = make_double2(re * invSig, im * invSig);
    return;
  }
  const int b = blockIdx.x;
  double re = 0, im = 0;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const float2 a = A[(size_t)b * rowStride + j];
    const double2 v = x[j];
    re += a.x * v.x - a.y * v.y;
    im += a.x * v.y + a.y * v.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = re;
  red[1][threadIdx.x] = im;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) y[b] = make_double2(red[0][0] * invSig, red[1][0] * invSig);
}

// g[b][s M + j] += coef conj(u[b]) v[j] for b < N, j < M
__global__ __launch_bounds__(256) void k_mdl_outer(float2* __restrict__ g, int N, int M, int rowStride, int s,
                                                   const double2* __restrict__ u, const double2* __restrict__ v,
                                                   double coef) {
  const size_t total = (size_t)N * M;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / M), j = (int)(e - (size_t)b * M);
    const double2 ub = u[b], vj = v[j];
    const double re = coef * (ub.x * vj.x + ub.y * vj.y), im = coef * (ub.x * vj.y - ub.y * vj.x);
    float2& o = g[(size_t)b * rowStride + (size_t)s * M + j];
    o.x += (float)re;
    o.y += (float)im;
  }
}

hipError_t launch_metric_rows(const float2* coefs, int N, int M, int Mpol, int rowStride, int S, double* out,
                              hipStream_t st) {
  dim3 grid(N, S);
  k_metric_rows<<<grid, 256, 0, st>>>(coefs, N, M, Mpol, rowStride, out);
  return hipGetLastError();
}

hipError_t launch_gram_f64(const float2* coefs, int N, int M, int rowStride, int s, int rowsAreVectors,
                           double2* G, hipStream_t st) {
  const int dim = rowsAreVectors ? N : M;
  dim3 grid((dim + 15) / 16, (dim + 15) / 16);
  k_gram_f64<<<grid, 256, 0, st>>>(coefs, N, M, rowStride, s, rowsAreVectors, G);
  return hipGetLastError();
}

hipError_t launch_mdl_seed_pair(const float2* coefs, int N, int M, int rowStride, int s, int rowsAreVectors,
                                const double2* x, double sig, double coef, double2* y, float2* g, hipStream_t st) {
  if (rowsAreVectors) k_mdl_image<<<(M + 255) / 256, 256, 0, st>>>(coefs, N, M, rowStride, s, 1, x, 1.0 / sig, y);
  else k_mdl_image<<<N, 256, 0, st>>>(coefs, N, M, rowStride, s, 0, x, 1.0 / sig, y);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const double2* u = rowsAreVectors ? x : y;
  const double2* v = rowsAreVectors ? y : x;
  const size_t n = (size_t)N * M;
  k_mdl_outer<<<stride_blocks(n, 256), 256, 0, st>>>(g, N, M, rowStride, s, u, v, coef);
  return hipGetLastError();
}

}  // namespace dh
