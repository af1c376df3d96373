// Plane analysis on the GPU: the statistics of digHoloFieldAnalysis
// (digHolo.cpp:2354-2654, finished on the host as in :9112-9140) for every
// (frame, pol) plane, the batch-summed intensity map, the same statistics of
// that map, and the int16 field export of GetFields16 (:4880-4906, :2820-2832).
//
// Per visited pixel the intensity is I = norm * |V|^2 (|V|^2 as
// fma(re, re, im * im), then one float multiply, as the reference forms it);
// the sums are carried in FP64, the peak in float with its index
// 2 * j * Nx + i (the reference's index convention, :2453-2456).
#include <hip/hip_runtime.h>

#include "dh_internal.h"

namespace dh {
namespace {

constexpr int kAThreads = 256;

struct Acc {
  double s0, sx, sy, syw, s4;
  float mx;
  int idx;
};

__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b) {
  a.s0 += b.s0;
  a.sx += b.sx;
  a.sy += b.sy;
  a.syw += b.syw;
  a.s4 += b.s4;
  if (b.mx > a.mx || (b.mx == a.mx && b.idx > a.idx)) {   // ties: the later pixel
    a.mx = b.mx;
    a.idx = b.idx;
  }
}

// One workgroup per plane.  CPLX: complex input V (the peak is of |V|^2);
// otherwise an intensity map (the peak is of I itself, as for the batch sum).
template <bool CPLX>
__global__ __launch_bounds__(kAThreads) void k_plane_stats(PlaneArgs a, const void* __restrict__ in,
                                                           size_t planeStride, float* __restrict__ pw,
                                                           double* __restrict__ out) {
  const int bp = blockIdx.x;
  const int N = a.Ny * a.Nx;
  Acc acc{0, 0, 0, 0, 0, -1.f, -1};
  for (int e = threadIdx.x; e < N; e += kAThreads) {
    const int j = e / a.Nx, i = e - j * a.Nx;
    const bool visited = i >= a.x0 && i < a.x1 && j != a.skip0 && j != a.skip1;
    float I = 0.f;
    if (visited) {
      const size_t off = (size_t)bp * planeStride + (size_t)j * a.inStride + i;
      float peak;
      if (CPLX) {
        const bool inc = a.include ? a.include[e] != 0 : true;
        const float2 v = static_cast<const float2*>(in)[off];
        const float vr = inc ? v.x : 0.f, vi = inc ? v.y : 0.f;
        const float v2 = __fmaf_rn(vr, vr, __fmul_rn(vi, vi));
        I = __fmul_rn(a.norm, v2);
        peak = v2;
      } else {
        I = static_cast<const float*>(in)[off];
        peak = I;
      }
      acc.s0 += I;
      acc.sx += (double)a.x[i] * I;
      acc.sy += (double)a.y[j] * I;
      acc.syw += (double)a.yw[j] * I;
      acc.s4 += (double)I * I;
      if (peak >= acc.mx) {
        acc.mx = peak;
        acc.idx = 2 * j * a.Nx + i;
      }
    }
    if (pw) pw[(size_t)bp * N + e] = I;
  }
  __shared__ Acc sh[kAThreads];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kAThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      Acc t = sh[threadIdx.x];
      acc_merge(t, sh[threadIdx.x + s]);
      sh[threadIdx.x] = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = out + (size_t)bp * 8;
    const Acc r = sh[0];
    o[0] = r.s0;
    o[1] = r.sx;
    o[2] = r.sy;
    o[3] = r.syw;
    o[4] = r.s4;
    o[5] = r.mx;
    o[6] = r.idx;
    o[7] = 0;
  }
}

// sum[p][e] = sum over frames of pw[b][p][e], added in frame order in float
// like the reference's running map (:2478-2500).  grid (ceil(N/256), P)
__global__ void k_batch_sum(const float* __restrict__ pw, int B, int P, int N, float* __restrict__ sum,
                            size_t polStride) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (e >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s = __fadd_rn(s, pw[((size_t)b * P + p) * N + e]);
  sum[(size_t)p * polStride + e] = s;
}

// int16 planes R, I = round(scale * V) with saturation; scale[p][B + 1]
__global__ void k_fields16(const float2* __restrict__ fields, int B, int P, int hw, const float* __restrict__ scale,
                           short* __restrict__ re, short* __restrict__ im) {
  const size_t total = (size_t)B * P * hw;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int bp = (int)(e / hw);
    const int b = bp / P, p = bp - b * P;
    const float s = scale[(size_t)p * (B + 1) + b];
    const float2 v = fields[e];
    const float r = fminf(fmaxf(rintf(__fmul_rn(s, v.x)), -32768.f), 32767.f);
    const float i = fminf(fmaxf(rintf(__fmul_rn(s, v.y)), -32768.f), 32767.f);
    re[e] = (short)r;
    im[e] = (short)i;
  }
}

}  // namespace

hipError_t launch_plane_stats(const PlaneArgs& a, const float2* inC, const float* inR, int planes, size_t planeStride,
                              float* pw, double* out, hipStream_t s) {
  if (planes <= 0 || a.Nx <= 0 || a.Ny <= 0) return hipErrorInvalidValue;
  if (inC)
    k_plane_stats<true><<<planes, kAThreads, 0, s>>>(a, inC, planeStride, pw, out);
  else
    k_plane_stats<false><<<planes, kAThreads, 0, s>>>(a, inR, planeStride, pw, out);
  return hipGetLastError();
}

hipError_t launch_batch_sum(const float* pw, int B, int P, int N, float* sum, size_t polStride, hipStream_t s) {
  if (B <= 0 || P <= 0 || N <= 0) return hipErrorInvalidValue;
  dim3 grid((N + 255) / 256, P);
  k_batch_sum<<<grid, 256, 0, s>>>(pw, B, P, N, sum, polStride);
  return hipGetLastError();
}

hipError_t launch_fields16(const float2* fields, int B, int P, int hw, const float* scale, short* re, short* im,
                           hipStream_t s) {
  const size_t total = (size_t)B * P * hw;
  if (total == 0) return hipErrorInvalidValue;
  k_fields16<<<stride_blocks(total, 256), 256, 0, s>>>(fields, B, P, hw, scale, re, im);
  return hipGetLastError();
}

}  // namespace dh
