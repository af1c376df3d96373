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
template <int R, bool INV, int NT, bool PAD = false, int VMAX = kMaxV>
__device__ __forceinline__ void stage_ip(float2* buf, int n, int nb, int ts, int es, int Ns, unsigned magNbf,
                                         unsigned magNs, unsigned magNb, bool tmajor, const float2* tw) {
  constexpr int BMAX = VMAX / R > 0 ? VMAX / R : 1;
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
template <bool INV, int NT, bool PAD = false, int VMAX = kMaxV>
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
          stage_ip<8, INV, NT, PAD, VMAX>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 4:
          stage_ip<4, INV, NT, PAD, VMAX>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 2:
          stage_ip<2, INV, NT, PAD, VMAX>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        case 3:
          stage_ip<3, INV, NT, PAD, VMAX>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
        default:
          stage_ip<5, INV, NT, PAD, VMAX>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw);
          break;
      }
      Ns *= R;
    }
  }
}

// ---------------------------------------------------------------------------
// The same transform for a compile-time length N (power of two: radix 8
// stages, then one of radix 4 or 2; else radices 8/4/2/3/5 as the run-time
// plans), padded, transforms at stride ts: divisions by constants (shifts for
// powers of two), and the LDS offsets of a butterfly's points are constants
// added to one computed base wherever the stride is a multiple of 32.
// ---------------------------------------------------------------------------
template <int N, int R, int NS, bool INV, int NT, int VMAX>
__device__ __forceinline__ void stage_pow2(float2* buf, int nb, int ts, const float2* tw) {
  constexpr int NBF = N / R;
  constexpr int STEP = N / (NS * R);
  constexpr int BMAX = VMAX / R > 0 ? VMAX / R : 1;
  const int total = nb * NBF;
  float2 v[BMAX][R];
  int wb[BMAX], wk[BMAX];
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    const int q = threadIdx.x + bb * NT;
    wb[bb] = -1;
    wk[bb] = 0;
    if (q < total) {
      const int t = q / NBF, j = q % NBF, jm = j % NS;
      const int tb = t * ts;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (NBF % 32 == 0)
          v[bb][r] = buf[tb + pad32(j) + r * (NBF + NBF / 32)];
        else
          v[bb][r] = buf[tb + pad32(j + r * NBF)];
      }
      if (NS > 1 && jm) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[bb][r] = cmul(v[bb][r], twiddle<INV>(tw, jm * r * STEP));
      }
      dftR<R, INV>(v[bb]);
      wb[bb] = tb;
      wk[bb] = (j - jm) * R + jm;
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    if (wb[bb] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (NS % 32 == 0)
          buf[wb[bb] + pad32(wk[bb]) + r * (NS + NS / 32)] = v[bb][r];
        else
          buf[wb[bb] + pad32(wk[bb] + r * NS)] = v[bb][r];
      }
    }
  }
  __syncthreads();
}

// radix of the next stage for the m points still to factor: 8, 4, 2, 3, 5 in
// that order of preference (the run-time plans use the same order); 0 when m
// has another prime factor
__host__ __device__ constexpr int ct_radix(int m) {
  return m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : m % 5 == 0 ? 5 : 0;
}

// true when N factors into 2, 3 and 5
__host__ __device__ constexpr bool ct_ok(int N) {
  for (int m = N; m > 1;) {
    const int r = ct_radix(m);
    if (!r) return false;
    m /= r;
  }
  return N > 1;
}

template <int N, int NS, bool INV, int NT, int VMAX>
__device__ __forceinline__ void stages_pow2(float2* buf, int nb, int ts, const float2* tw) {
  if constexpr (NS < N) {
    constexpr int R = ct_radix(N / NS);
    static_assert(R > 0, "length with a prime factor above 5");
    stage_pow2<N, R, NS, INV, NT, VMAX>(buf, nb, ts, tw);
    stages_pow2<N, NS * R, INV, NT, VMAX>(buf, nb, ts, tw);
  }
}

// Most transforms of length N one pass of NT threads takes (each thread holds
// at most VMAX values, at least one butterfly, per stage).
template <int N, int NT, int VMAX>
__host__ __device__ constexpr int pow2_chunk() {
  int best = 1 << 30;
  for (int m = N; m > 1;) {
    const int r = ct_radix(m);
    const int bmax = VMAX / r > 0 ? VMAX / r : 1;
    const int c = bmax * NT * r / N;
    best = c < best ? c : best;
    m /= r;
  }
  return best;
}

// nb transforms of length N at buf + t * ts (padded points), in passes of
// pow2_chunk transforms.  Called by the whole block; ends with a barrier.
template <int N, bool INV, int NT, int VMAX = kMaxV>
__device__ void lds_fft_pow2(float2* buf, int nb, int ts, const float2* tw) {
  constexpr int CH = pow2_chunk<N, NT, VMAX>();
  static_assert(CH >= 1, "transform too long for one pass of the block");
  for (int t0 = 0; t0 < nb; t0 += CH) stages_pow2<N, 1, INV, NT, VMAX>(buf + t0 * ts, min(CH, nb - t0), ts, tw);
}

// nb padded transforms of plan pl at buf + t * ts: the compile-time form for
// the lengths listed (inverse-window sizes in multiples of 16), else the
// run-time one (chunk: ip_chunk of pl).  Called by the whole block.
template <bool INV, int NT, int VMAX = kMaxV>
__device__ void lds_fft_any(float2* buf, int nb, int ts, const FFTPlan& pl, const float2* tw, int chunk) {
  switch (pl.n) {
    case 32: lds_fft_pow2<32, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 48: lds_fft_pow2<48, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 64: lds_fft_pow2<64, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 80: lds_fft_pow2<80, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 96: lds_fft_pow2<96, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 128: lds_fft_pow2<128, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 144: lds_fft_pow2<144, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 160: lds_fft_pow2<160, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 192: lds_fft_pow2<192, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 240: lds_fft_pow2<240, INV, NT, VMAX>(buf, nb, ts, tw); break;
    case 256: lds_fft_pow2<256, INV, NT, VMAX>(buf, nb, ts, tw); break;
    default: lds_fft_ip<INV, NT, true, VMAX>(buf, nb, ts, 1, false, pl, tw, chunk); break;
  }
}

// Transforms one in-place pass can take with nt threads (each thread holds at
// most vmax values per stage); 0 when a radix has no in-place butterfly.
inline int ip_chunk(const FFTPlan& pl, int nt, int vmax = kMaxV) {
  int best = 1 << 30;
  for (int s = 0; s < pl.nstages; ++s) {
    const int R = pl.radix[s];
    if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) return 0;
    const int bmax = vmax / R > 0 ? vmax / R : 1;
    best = std::min(best, bmax * nt * R / pl.n);
  }
  return best;
}

}  // namespace dh
