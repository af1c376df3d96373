// In-place batched Stockham FFT staged in LDS (radix 2/3/4/5/8), shared by
// the fused window kernels and the staged kernels.
#pragma once
#include <algorithm>

#include "dh_fft.h"
#include "dh_internal.h"

namespace dh {

constexpr int kMaxV = 8;                    // complex values per thread in one stage

__device__ __forceinline__ int fdiv(int q, unsigned mag) {
  return mag ? (int)__umulhi((unsigned)q, mag) : q;
}

// Padded position of point k within a transform: one spare float2 after every
// 32, so the strided writes of the early Stockham stages (stride R) and the
// reads (stride 1) spread over the LDS banks.
__host__ __device__ constexpr int pad32(int k) { return k + (k >> 5); }
// stride of padded transforms of length n (odd, so that transforms side by
// side start on different banks)
__host__ __device__ constexpr int pad_stride(int n) { return (pad32(n - 1) + 1) | 1; }

// In-place radix-R Stockham stage over nb transforms of length n; element
// (t, k) lives at buf[t*ts + k*es] (PAD: at buf[t*ts + pad32(k)], es unused).
// tmajor: consecutive threads take consecutive transforms (use when ts == 1),
// else consecutive butterflies.
template <int R, bool INV, int NT, bool PAD = false>
__device__ __forceinline__ void stage_ip(float2* buf, int n, int nb, int ts, int es, int Ns, unsigned magNbf,
                                         unsigned magNs, unsigned magNb, bool tmajor, const float2* tw) {
  constexpr int BMAX = kMaxV / R > 0 ? kMaxV / R : 1;
  const int nbf = n / R;
  const int total = nb * nbf;
  const int step = n / (Ns * R);
  float2 v[BMAX][R];
  int wb[BMAX], wk[BMAX];
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    const int q = threadIdx.x + bb * NT;
    wb[bb] = -1;
    wk[bb] = 0;
    if (q < total) {
      int t, j;
      if (tmajor) {
        j = fdiv(q, magNb);
        t = q - j * nb;
      } else {
        t = fdiv(q, magNbf);
        j = q - t * nbf;
      }
      const int jm = j - fdiv(j, magNs) * Ns;
      const int tb = t * ts;
#pragma unroll
      for (int r = 0; r < R; ++r) v[bb][r] = PAD ? buf[tb + pad32(j + r * nbf)] : buf[tb + (j + r * nbf) * es];
      if (jm) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[bb][r] = cmul(v[bb][r], twiddle<INV>(tw, jm * r * step));
      }
      dftR<R, INV>(v[bb]);
      if (PAD) {
        wb[bb] = tb;
        wk[bb] = (j - jm) * R + jm;
      } else {
        wb[bb] = tb + ((j - jm) * R + jm) * es;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    if (wb[bb] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (PAD)
          buf[wb[bb] + pad32(wk[bb] + r * Ns)] = v[bb][r];
        else
          buf[wb[bb] + r * Ns * es] = v[bb][r];
      }
    }
  }
  __syncthreads();
}

// nb transforms, at most `chunk` per in-place pass.
template <bool INV, int NT, bool PAD = false>
__device__ void lds_fft_ip(float2* buf, int nb, int ts, int es, bool tmajor, const FFTPlan& pl, const float2* tw,
                           int chunk) {
  for (int t0 = 0; t0 < nb; t0 += chunk) {
    const int cnt = min(chunk, nb - t0);
    const unsigned magNb = tmajor ? magic_div((unsigned)cnt) : 0u;
    float2* b = buf + t0 * ts;
    int Ns = 1;
    for (int s = 0; s < pl.nstages; ++s) {
      const int R = pl.radix[s];
      switch (R) {
        case 8:
          stage_ip<8, INV, NT, PAD>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 4:
          stage_ip<4, INV, NT, PAD>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 2:
          stage_ip<2, INV, NT, PAD>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 3:
          stage_ip<3, INV, NT, PAD>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        default:
          stage_ip<5, INV, NT, PAD>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
      }
      Ns *= R;
    }
  }
}

// Transforms one in-place pass can take with nt threads (each thread holds at
// most kMaxV values per stage); 0 when a radix has no in-place butterfly.
inline int ip_chunk(const FFTPlan& pl, int nt) {
  int best = 1 << 30;
  for (int s = 0; s < pl.nstages; ++s) {
    const int R = pl.radix[s];
    if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) return 0;
    const int bmax = kMaxV / R > 0 ? kMaxV / R : 1;
    best = std::min(best, bmax * nt * R / pl.n);
  }
  return best;
}

}  // namespace dh
