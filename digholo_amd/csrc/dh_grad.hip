// Adjoint of the batch path: the gradient of the alignment loss
//   L = -(1/B) sum_{b,p,mode} |c|^2  [+ mu (ln s_max^2 - ln s_min^2)]
// with respect to each pol's waist, defocus, tilt and beam centre, with the
// window placement (FFT crops, Fourier-window position, size and mask) held.
// No counterpart in the reference, whose AutoAlign searches without
// derivatives (digHolo.cpp:12315-12885).
//
// With g = dL/dc = -conj(c)/B (Wirtinger; plus the mode-dependent-loss
// term's seed gExtra, formed on the host from the extreme singular triplets
// of the coefficient matrix) and the adjoint field
// A(x, y) = sum_mn g_mn HGx_m(x) HGy_n(y) (the transposed overlap), the
// field V = -conj(E R) responds to
//   tilt:    dV/dtheta_x = i k0 (x + bcX) V          (reference-wave slope and piston)
//   defocus: dV/dD = i (pi / lambda) (x^2 + y^2) V
//   centre:  the window ramp exp(i kx bcx) shifts E: dE/dbcx = IFFT(i kx W),
//            evaluated in the Fourier plane as sum_k kx Im(X1 conj X2) / N with
//            X1 = IFFT(-V R) = N conj(W), X2 = IFFT(conj(A) R)
//   waist:   through the basis, 2 Re sum (dHGx HGy + HGx dHGy) g V.
// k_grad_backproject forms A and the waist adjoint per (frame, pol) in LDS,
// reduces the real-space sums and writes -V R and conj(A) R; the two inverse
// transforms run on the staged IFFT kernels; k_grad_kspace reduces the
// Fourier-plane sums; k_grad_reduce sums the partials over the batch.
#include <hip/hip_runtime.h>

#include "dh_internal.h"

namespace dh {
namespace {

constexpr int kGThreads = 256;
constexpr int kGXC = 32;          // field columns per back-projection chunk

// d/dw of the HG tables of k_hg_basis: with u = sqrt2 t / w,
// psi_n'(u) = sqrt(n/2) psi_{n-1} - sqrt((n+1)/2) psi_{n+1},
// d/dw [psi_n(u) sqrt(dt sqrt2 / w)] = nrm (psi_n'(u) (-u/w) - psi_n / (2w)).
// The recurrence runs to n = G with the mantissa renormalised every step and
// the binary exponent carried separately, as for the basis itself.
__global__ void k_hg_dwaist(TableArgs a, float* __restrict__ dhgx, float* __restrict__ dhgy) {
  const Geometry& g = a.g;
  const int per = g.fw + g.fh;
  const int total = g.P * per;
  const double rt2 = 1.41421356237309504880;
  const double pim4 = 0.75112554446494248286;
  const double ln2 = 0.69314718055994530942;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int p = idx / per;
    const int k = idx - p * per;
    const PolParams& pp = a.pp[p];
    const bool isx = k < g.fw;
    const int i = isx ? k : k - g.fw;
    const int N = isx ? g.fw : g.fh;
    const double dt = isx ? pp.pxx : pp.pxy;
    const double t = (i - N / 2) * dt;
    const double w = pp.waist;
    const double u = rt2 * t / w;
    const double nrm = sqrt(dt * rt2 / w);
    const double gexp = -0.5 * u * u;
    float* out = isx ? (dhgx + (size_t)p * g.G * g.fw + i) : (dhgy + (size_t)p * g.G * g.fh + i);
    // psi values: prev = psi_{n-1}, cur = psi_n, computed one order ahead
    double m0 = pim4, e0 = 0.0;             // psi_0
    double m1 = rt2 * pim4 * u, e1 = 0.0;   // psi_1
    double prev = 0.0;
    double cur = m0 * exp(gexp);
    double next = m1 * exp(gexp);
    for (int n = 0; n < g.G; ++n) {
      // next = psi_{n+1}; advance the recurrence for the following order
      const double dpsi = (n > 0 ? sqrt(0.5 * n) * prev : 0.0) - sqrt(0.5 * (n + 1)) * next;
      out[(size_t)n * N] = (float)(nrm * (dpsi * (-u / w) - cur / (2.0 * w)));
      const int n2 = n + 2;   // psi_{n+2} from psi_{n+1} (m1) and psi_n (m0)
      const double c1 = sqrt(2.0 / n2), c2 = sqrt((n2 - 1.0) / n2);
      double mn = c1 * u * m1 - c2 * m0 * exp2(e0 - e1);
      int fe;
      mn = frexp(mn, &fe);
      const double en = e1 + fe;
      m0 = m1;
      e0 = e1;
      m1 = mn;
      e1 = en;
      prev = cur;
      cur = next;
      next = mn * exp(gexp + en * ln2);
    }
  }
}

__device__ __forceinline__ int mode_index(int m, int n) {
  const int gg = m + n;
  return gg * (gg + 1) / 2 + n;
}

template <int NS>
__device__ __forceinline__ void block_sum(double (&v)[NS], double* sh, double* out) {
  for (int k = 0; k < NS; ++k) sh[k * kGThreads + threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kGThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < NS; ++k) sh[k * kGThreads + threadIdx.x] += sh[k * kGThreads + threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < NS) out[threadIdx.x] = sh[threadIdx.x * kGThreads];
}

// One workgroup per (pol, frame).  partial[bp][8] = {sum Im(AV), sum x Im(AV),
// sum y Im(AV), sum r^2 Im(AV), sum Re(Aw V), sum |c|^2, -, -}.
__global__ __launch_bounds__(kGThreads) void k_grad_backproject(
    Geometry g, double invB, const float2* __restrict__ coefs, const float2* __restrict__ gExtra,
    const float* __restrict__ hgx,
    const float* __restrict__ hgy, const float* __restrict__ dhgx, const float* __restrict__ dhgy,
    const float2* __restrict__ fields, const float2* __restrict__ ref, const double* __restrict__ pxy2,
    float2* __restrict__ q1, float2* __restrict__ q2, double* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char gsm[];
  const int p = blockIdx.x, b = blockIdx.y + g.b0;
  const int G = g.G, M = g.M, fw = g.fw, fh = g.fh;
  float2* gs = reinterpret_cast<float2*>(gsm);                 // [M]
  float2* T1 = gs + M;                                          // [G][kGXC]
  float2* T1w = T1 + G * kGXC;                                  // [G][kGXC]
  const size_t bp = (size_t)b * g.P + p;
  const float2* c = coefs + bp * M;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int e = threadIdx.x; e < M; e += kGThreads) {
    const float2 v = c[e];
    acc[5] += (double)v.x * v.x + (double)v.y * v.y;
    float2 gv = make_float2((float)(-v.x * invB), (float)(v.y * invB));   // -conj(c) / B
    if (gExtra) {
      const float2 x = gExtra[bp * M + e];
      gv.x += x.x;
      gv.y += x.y;
    }
    gs[e] = gv;
  }
  const float* hx = hgx + (size_t)p * G * fw;
  const float* hy = hgy + (size_t)p * G * fh;
  const float* dx = dhgx + (size_t)p * G * fw;
  const float* dy = dhgy + (size_t)p * G * fh;
  const float2* V = fields + bp * (size_t)fh * fw;
  const float2* R = ref + (size_t)p * fh * fw;
  float2* o1 = q1 + bp * (size_t)fh * fw;
  float2* o2 = q2 + bp * (size_t)fh * fw;
  const double pxx = pxy2[2 * p], pyy = pxy2[2 * p + 1];
  for (int x0 = 0; x0 < fw; x0 += kGXC) {
    __syncthreads();   // gs written / previous chunk's T1 read
    for (int e = threadIdx.x; e < G * kGXC; e += kGThreads) {
      const int n = e / kGXC, xc = e - n * kGXC, x = x0 + xc;
      float2 s = make_float2(0.f, 0.f), sw = make_float2(0.f, 0.f);
      if (x < fw) {
        for (int m = 0; m < G - n; ++m) {
          const float2 gv = gs[mode_index(m, n)];
          const float h = hx[(size_t)m * fw + x], dh = dx[(size_t)m * fw + x];
          s.x = fmaf(gv.x, h, s.x);
          s.y = fmaf(gv.y, h, s.y);
          sw.x = fmaf(gv.x, dh, sw.x);
          sw.y = fmaf(gv.y, dh, sw.y);
        }
      }
      T1[e] = s;
      T1w[e] = sw;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < fh * kGXC; e += kGThreads) {
      const int y = e / kGXC, xc = e - y * kGXC, x = x0 + xc;
      if (x >= fw) continue;
      float2 A = make_float2(0.f, 0.f), Aw = make_float2(0.f, 0.f);
      for (int n = 0; n < G; ++n) {
        const float h = hy[(size_t)n * fh + y], dh = dy[(size_t)n * fh + y];
        const float2 t = T1[n * kGXC + xc], tw = T1w[n * kGXC + xc];
        A.x = fmaf(h, t.x, A.x);
        A.y = fmaf(h, t.y, A.y);
        Aw.x = fmaf(h, tw.x, fmaf(dh, t.x, Aw.x));
        Aw.y = fmaf(h, tw.y, fmaf(dh, t.y, Aw.y));
      }
      const size_t off = (size_t)y * fw + x;
      const float2 v = V[off], r = R[off];
      const double im = (double)A.x * v.y + (double)A.y * v.x;
      const double xc_ = (x - fw / 2) * pxx, yc_ = (y - fh / 2) * pyy;
      acc[0] += im;
      acc[1] += xc_ * im;
      acc[2] += yc_ * im;
      acc[3] += (xc_ * xc_ + yc_ * yc_) * im;
      acc[4] += (double)Aw.x * v.x - (double)Aw.y * v.y;
      o1[off] = make_float2(-(v.x * r.x - v.y * r.y), -(v.x * r.y + v.y * r.x));   // -V R
      o2[off] = make_float2(A.x * r.x + A.y * r.y, A.x * r.y - A.y * r.x);        // conj(A) R
    }
  }
  __shared__ double sh[6 * kGThreads];
  __syncthreads();
  block_sum<6>(acc, sh, partial + bp * 8);
}

// partial[bp][6..7] = sum kx Im(X1 conj X2), sum ky Im(X1 conj X2) (unnormalised)
__global__ __launch_bounds__(kGThreads) void k_grad_kspace(Geometry g, const float2* __restrict__ x1,
                                                           const float2* __restrict__ x2,
                                                           const double* __restrict__ kx,
                                                           const double* __restrict__ ky,
                                                           double* __restrict__ partial) {
  const int p = blockIdx.x, b = blockIdx.y + g.b0;
  const int fw = g.fw, fh = g.fh, N = fw * fh;
  const size_t bp = (size_t)b * g.P + p;
  const float2* a = x1 + bp * N;
  const float2* c = x2 + bp * N;
  double acc[2] = {0, 0};
  for (int e = threadIdx.x; e < N; e += kGThreads) {
    const int j = e / fw, i = e - j * fw;
    const float2 u = a[e], v = c[e];
    const double im = (double)u.y * v.x - (double)u.x * v.y;
    acc[0] += kx[(size_t)p * fw + i] * im;
    acc[1] += ky[(size_t)p * fh + j] * im;
  }
  __shared__ double sh[2 * kGThreads];
  block_sum<2>(acc, sh, partial + bp * 8 + 6);
}

// out[p][k] = sum_b partial[b][p][k]
__global__ __launch_bounds__(kGThreads) void k_grad_reduce(int B, int P, const double* __restrict__ partial,
                                                           double* __restrict__ out) {
  const int p = blockIdx.x;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < B; b += kGThreads)
    for (int k = 0; k < 8; ++k) acc[k] += partial[((size_t)b * P + p) * 8 + k];
  __shared__ double sh[8 * kGThreads];
  block_sum<8>(acc, sh, out + (size_t)p * 8);
}

}  // namespace

hipError_t launch_hg_dwaist(const TableArgs& a, float* dhgx, float* dhgy, hipStream_t s) {
  if (a.g.G <= 0) return hipSuccess;
  const int t2 = a.g.P * (a.g.fw + a.g.fh);
  k_hg_dwaist<<<(t2 + 63) / 64, 64, 0, s>>>(a, dhgx, dhgy);
  return hipGetLastError();
}

size_t grad_backproject_lds(const Geometry& g) {
  return (size_t)g.M * sizeof(float2) + 2 * (size_t)g.G * kGXC * sizeof(float2);
}

hipError_t launch_grad_backproject(const Geometry& g, double invB, const float2* coefs, const float2* gExtra,
                                   const float* hgx,
                                   const float* hgy, const float* dhgx, const float* dhgy, const float2* fields,
                                   const float2* ref, const double* pxy2, float2* q1, float2* q2, double* partial,
                                   hipStream_t s) {
  if (g.G <= 0 || g.B <= 0) return hipErrorInvalidValue;
  const size_t lds = grad_backproject_lds(g);
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_grad_backproject),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_grad_backproject<<<dim3(g.P, nb), kGThreads, lds, s>>>(gc, invB, coefs, gExtra, hgx, hgy, dhgx, dhgy, fields,
                                                             ref, pxy2, q1, q2, partial);
    return hipGetLastError();
  });
}

hipError_t launch_grad_kspace(const Geometry& g, const float2* x1, const float2* x2, const double* kx,
                              const double* ky, double* partial, hipStream_t s) {
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_grad_kspace<<<dim3(g.P, nb), kGThreads, 0, s>>>(gc, x1, x2, kx, ky, partial);
    return hipGetLastError();
  });
}

hipError_t launch_grad_reduce(int B, int P, const double* partial, double* out, hipStream_t s) {
  k_grad_reduce<<<P, kGThreads, 0, s>>>(B, P, partial, out);
  return hipGetLastError();
}

}  // namespace dh
