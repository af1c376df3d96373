// Basis transforms of the HG coefficients and the basis fields
// (digHolo.cpp:9469-9605 applyBasisTransform / HGtoLG, :5233-5296 GetFieldsBasis).
//
// LG: block-diagonal per HG group g, out[g0 + k] = sum_c U_g[k][c] in[g0 + c]
// with g0 = g(g+1)/2 and U_g stored row-major at lgOff(g) = sum_{h<g} (h+1)^2.
// Custom: a dense product out[r][j] = sum_i in[r][i] T[i * ldT + j] over the
// nIn x nOut block the batch's mode count allows.
// Basis fields: out[mode][p][y][x] = sum_i W[mode][i] HGx_{m_i}(x) HGy_{n_i}(y),
// separable in two stages per (mode, pol) workgroup.
#include <hip/hip_runtime.h>

#include "dh_internal.h"

namespace dh {
namespace {

__device__ __forceinline__ float2 cmla(float2 acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
  acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
  return acc;
}

__device__ __forceinline__ int group_of(int k) {
  int g = (int)((sqrtf(8.f * k + 1.f) - 1.f) * 0.5f);
  while ((g + 1) * (g + 2) / 2 <= k) ++g;
  while (g * (g + 1) / 2 > k) --g;
  return g;
}

// one thread per output coefficient; rows = B * P
__global__ void k_lg_apply(const float2* __restrict__ in, float2* __restrict__ out, size_t rows, int M,
                           const float2* __restrict__ tab) {
  const size_t total = rows * M;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t r = e / M;
    const int k = (int)(e - r * M);
    const int g = group_of(k);
    const int g0 = g * (g + 1) / 2, kk = k - g0;
    // sum_{h<g} (h+1)^2 = g (g+1) (2g+1) / 6
    const float2* U = tab + (size_t)g * (g + 1) * (2 * g + 1) / 6 + (size_t)kk * (g + 1);
    const float2* x = in + r * M + g0;
    float2 acc = make_float2(0.f, 0.f);
    for (int c = 0; c <= g; ++c) acc = cmla(acc, U[c], x[c]);
    out[e] = acc;
  }
}

// 64 x 64 output tile per 256-thread workgroup, 4 x 4 per thread, k-steps of 16
constexpr int kCT = 64, kCK = 16;
__global__ __launch_bounds__(256) void k_custom_apply(const float2* __restrict__ in, int ldIn, int rows, int nIn,
                                                      const float2* __restrict__ T, int ldT, int nOut,
                                                      float2* __restrict__ out) {
  __shared__ float2 sA[kCK][kCT + 1];   // [k][row]
  __shared__ float2 sB[kCK][kCT + 1];   // [k][col]
  const int r0 = blockIdx.y * kCT, c0 = blockIdx.x * kCT;
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  float2 acc[4][4];
  for (auto& a : acc)
    for (auto& v : a) v = make_float2(0.f, 0.f);
  for (int k0 = 0; k0 < nIn; k0 += kCK) {
    for (int e = threadIdx.x; e < kCK * kCT; e += 256) {
      const int kk = e % kCK, rr = e / kCK;       // A: consecutive threads along k
      const int r = r0 + rr, k = k0 + kk;
      sA[kk][rr] = (r < rows && k < nIn) ? in[(size_t)r * ldIn + k] : make_float2(0.f, 0.f);
      const int cc = e % kCT, kb = e / kCT;       // B: consecutive threads along columns
      const int c = c0 + cc, k2 = k0 + kb;
      sB[kb][cc] = (c < nOut && k2 < nIn) ? T[(size_t)k2 * ldT + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kCK; ++kk) {
      float2 a[4], b[4];
      for (int i = 0; i < 4; ++i) a[i] = sA[kk][tr + 16 * i];
      for (int j = 0; j < 4; ++j) b[j] = sB[kk][tc + 16 * j];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = cmla(acc[i][j], a[i], b[j]);
    }
    __syncthreads();
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + tr + 16 * i, c = c0 + tc + 16 * j;
      if (r < rows && c < nOut) out[(size_t)r * nOut + c] = acc[i][j];
    }
}

constexpr int kSXC = 32;   // field columns per chunk
__global__ __launch_bounds__(256) void k_basis_synth(Geometry g, const float2* __restrict__ W, int ldW,
                                                     const float* __restrict__ hgx, const float* __restrict__ hgy,
                                                     float2* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char bsm[];
  const int p = blockIdx.x, mode = blockIdx.y;
  const int G = g.G, M = g.M, fw = g.fw, fh = g.fh;
  float2* ws = reinterpret_cast<float2*>(bsm);   // [M]
  float2* T1 = ws + M;                           // [G][kSXC]
  for (int e = threadIdx.x; e < M; e += 256) ws[e] = W[(size_t)mode * ldW + e];
  const float* hx = hgx + (size_t)p * G * fw;
  const float* hy = hgy + (size_t)p * G * fh;
  float2* o = out + ((size_t)mode * g.P + p) * fh * fw;
  for (int x0 = 0; x0 < fw; x0 += kSXC) {
    __syncthreads();
    for (int e = threadIdx.x; e < G * kSXC; e += 256) {
      const int n = e / kSXC, xc = e - n * kSXC, x = x0 + xc;
      float2 s = make_float2(0.f, 0.f);
      if (x < fw)
        for (int m = 0; m < G - n; ++m) {
          const int gg = m + n;
          const float2 w = ws[gg * (gg + 1) / 2 + n];
          const float h = hx[(size_t)m * fw + x];
          s.x = fmaf(w.x, h, s.x);
          s.y = fmaf(w.y, h, s.y);
        }
      T1[e] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < fh * kSXC; e += 256) {
      const int y = e / kSXC, xc = e - y * kSXC, x = x0 + xc;
      if (x >= fw) continue;
      float2 a = make_float2(0.f, 0.f);
      for (int n = 0; n < G; ++n) {
        const float h = hy[(size_t)n * fh + y];
        const float2 t = T1[n * kSXC + xc];
        a.x = fmaf(h, t.x, a.x);
        a.y = fmaf(h, t.y, a.y);
      }
      o[(size_t)y * fw + x] = a;
    }
  }
}

}  // namespace

hipError_t launch_lg_apply(const float2* in, float2* out, size_t rows, int M, const float2* tab, hipStream_t s) {
  const size_t n = rows * M;
  if (n == 0) return hipSuccess;
  k_lg_apply<<<stride_blocks(n, 256), 256, 0, s>>>(in, out, rows, M, tab);
  return hipGetLastError();
}

hipError_t launch_custom_apply(const float2* in, int ldIn, int rows, int nIn, const float2* T, int ldT, int nOut,
                               float2* out, hipStream_t s) {
  if (rows <= 0 || nOut <= 0) return hipSuccess;
  // grid y covers at most 65535 row tiles: larger row counts run in chunks
  constexpr int kRowsPerLaunch = 65535 * kCT;
  for (int r0 = 0; r0 < rows; r0 += kRowsPerLaunch) {
    const int nr = rows - r0 < kRowsPerLaunch ? rows - r0 : kRowsPerLaunch;
    dim3 grid((nOut + kCT - 1) / kCT, (nr + kCT - 1) / kCT);
    k_custom_apply<<<grid, 256, 0, s>>>(in + (size_t)r0 * ldIn, ldIn, nr, nIn, T, ldT, nOut, out + (size_t)r0 * nOut);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_basis_synth(const Geometry& g, const float2* W, int modes, int ldW, const float* hgx,
                              const float* hgy, float2* out, hipStream_t s) {
  if (modes <= 0 || g.G <= 0) return hipSuccess;
  const size_t lds = ((size_t)g.M + (size_t)g.G * kSXC) * sizeof(float2);
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_basis_synth),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  k_basis_synth<<<dim3(g.P, modes), 256, lds, s>>>(g, W, ldW, hgx, hgy, out);
  return hipGetLastError();
}

}  // namespace dh
