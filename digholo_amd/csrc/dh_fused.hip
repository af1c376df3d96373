// Fused window path: one workgroup per (frame, pol) turns the camera window
// into the reconstructed field without any intermediate leaving LDS.
//
//   1. rows of the window (uint16/float) -> packed r2c FFT of W/2 points,
//      keeping only the w columns [c0, c0+w) the Fourier window reads;
//   2. the column transform is split by decimation in time over R row
//      residues:  F[k][.] = sum_r W_H^(k r) G_r[k mod Hc][.],  G_r the
//      Hc = H/R point FFT of rows y = R t + r.  Each thread accumulates its
//      own window pixels (k = the h window rows) in registers;
//   3. window factor (mask, fill-factor envelope, centring ramp, norm);
//   4. 2-D inverse FFT (w along x, h along y) in place;
//   5. V = -conj(E * ref) -> fields.
//
// Every transform is an in-place Stockham FFT: in each stage a thread reads
// all of its butterflies into registers, the block synchronises, then writes;
// batches of transforms larger than the register budget run in chunks.  The
// next row batch is fetched from HBM into registers while the current one is
// transformed.  Supported radices 2/3/4/5/8; other shapes use the staged
// kernels of dh_kernels.hip.
#include <algorithm>

#include "dh_fft.h"
#include "dh_internal.h"

namespace dh {

constexpr int kFThreads = 512;
constexpr int kMaxV = 8;                    // complex values per thread in one stage
constexpr int kPre = 8;                     // prefetched row elements per thread
constexpr size_t kFusedLdsMax = 80 * 1024;  // two workgroups per CU
constexpr int kMaxFPT = 16;                 // window pixels accumulated per thread

__device__ __forceinline__ int fdiv(int q, unsigned mag) {
  return mag ? (int)__umulhi((unsigned)q, mag) : q;
}

// In-place radix-R Stockham stage over nb transforms of length n; element
// (t, k) lives at buf[t*ts + k*es].  tmajor: consecutive threads take
// consecutive transforms (use when ts == 1), else consecutive butterflies.
template <int R, bool INV>
__device__ __forceinline__ void stage_ip(float2* buf, int n, int nb, int ts, int es, int Ns, unsigned magNbf,
                                         unsigned magNs, unsigned magNb, bool tmajor, const float2* tw) {
  constexpr int BMAX = kMaxV / R > 0 ? kMaxV / R : 1;
  const int nbf = n / R;
  const int total = nb * nbf;
  const int step = n / (Ns * R);
  float2 v[BMAX][R];
  int wb[BMAX];
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    const int q = threadIdx.x + bb * kFThreads;
    wb[bb] = -1;
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
      for (int r = 0; r < R; ++r) v[bb][r] = buf[tb + (j + r * nbf) * es];
      if (jm) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[bb][r] = cmul(v[bb][r], twiddle<INV>(tw, jm * r * step));
      }
      dftR<R, INV>(v[bb]);
      wb[bb] = tb + ((j - jm) * R + jm) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < BMAX; ++bb) {
    if (wb[bb] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wb[bb] + r * Ns * es] = v[bb][r];
    }
  }
  __syncthreads();
}

// nb transforms, at most `chunk` per in-place pass.
template <bool INV>
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
        case 8: stage_ip<8, INV>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw); break;
        case 4: stage_ip<4, INV>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw); break;
        case 2: stage_ip<2, INV>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw); break;
        case 3: stage_ip<3, INV>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw); break;
        default: stage_ip<5, INV>(b, pl.n, cnt, ts, es, Ns, pl.magNbf[s], pl.magNs[s], magNb, tmajor, tw); break;
      }
      Ns *= R;
    }
  }
}

struct FusedLayout {
  int regionX;   // float2 elements shared by (row buffer + band) and the field
  int tables;    // float2 elements of twiddle tables
  size_t bytes;
};

__host__ __device__ inline FusedLayout fused_layout(const FusedArgs& a) {
  FusedLayout L;
  const int fwd = a.RB * a.row.n + a.Hc * a.g.w;
  const int inv = a.g.h * a.g.w;
  L.regionX = fwd > inv ? fwd : inv;
  L.tables = a.row.n + a.Hc + a.g.w + a.g.H + a.g.w + a.g.h;
  L.bytes = sizeof(float2) * (size_t)(L.regionX + L.tables) + sizeof(int) * (size_t)a.g.h;
  return L;
}

struct FusedMagic {
  unsigned n, w, Hc, H;
};

// fetch row batch (r, t0) of the window into registers: element e = tid + s*NT
// of the [rows][n] packed-complex batch
__device__ __forceinline__ void fetch_rows(float2 (&pre)[kPre], const void* frames, int u16, size_t fbase,
                                           int frameW, int R, int r, int t0, int rows, int n, unsigned magN) {
#pragma unroll
  for (int s = 0; s < kPre; ++s) {
    const int e = threadIdx.x + s * kFThreads;
    if (e < rows * n) {
      const int tt = fdiv(e, magN), k = e - tt * n;
      const size_t off = fbase + (size_t)(R * (t0 + tt) + r) * frameW + 2 * k;
      if (u16) {
        const ushort2 v = *reinterpret_cast<const ushort2*>(static_cast<const unsigned short*>(frames) + off);
        pre[s] = make_float2((float)v.x, (float)v.y);
      } else {
        pre[s] = *reinterpret_cast<const float2*>(static_cast<const float*>(frames) + off);
      }
    }
  }
}

template <int FPT>
__global__ __launch_bounds__(kFThreads, 4) void k_fused_window(FusedArgs a, FusedMagic mg,
                                                               const float2* __restrict__ twW,
                                                               const float2* __restrict__ twH,
                                                               const void* __restrict__ frames, int u16,
                                                               const float2* __restrict__ factor,
                                                               const float2* __restrict__ ref,
                                                               float2* __restrict__ fields) {
  extern __shared__ float2 fsm[];
  const Geometry& g = a.g;
  const int p = blockIdx.x, b = blockIdx.y;
  const int n = a.row.n, Hc = a.Hc, R = a.R, RB = a.RB;
  const int H = g.H, w = g.w, h = g.h, hw = h * w;
  const int tid = threadIdx.x;
  const FusedLayout L = fused_layout(a);
  float2* rowbuf = fsm;                   // [RB][n]
  float2* band = fsm + RB * n;            // [Hc][w]
  float2* fbuf = fsm;                     // [h][w], after the forward pass
  float2* twn = fsm + L.regionX;          // n
  float2* twc = twn + n;                  // Hc
  float2* twr = twc + Hc;                 // w: exp(-2 pi i (c0 + i) / W)
  float2* twy = twr + w;                  // H: exp(-2 pi i k / H)
  float2* twix = twy + H;                 // w
  float2* twiy = twix + w;                // h
  int* kyTab = reinterpret_cast<int*>(twiy + h);   // h: r2c row of window row j
  const int c0 = g.c0[p];
  const size_t fbase = (size_t)b * g.frameH * g.frameW + (size_t)g.cy[p] * g.frameW + g.cx[p];

  // first row batch in flight while the tables are staged
  float2 pre[kPre];
  const int nbat = (Hc + RB - 1) / RB;     // row batches per residue
  fetch_rows(pre, frames, u16, fbase, g.frameW, R, 0, 0, min(RB, Hc), n, mg.n);

  for (int e = tid; e < n; e += kFThreads) twn[e] = a.row.tw[e];
  for (int e = tid; e < Hc; e += kFThreads) twc[e] = a.col.tw[e];
  for (int e = tid; e < w; e += kFThreads) {
    twr[e] = twW[c0 + e];
    twix[e] = a.ix.tw[e];
  }
  for (int e = tid; e < H; e += kFThreads) twy[e] = twH[e];
  for (int e = tid; e < h; e += kFThreads) {
    twiy[e] = a.iy.tw[e];
    int row = (e + g.SY[p] + H / 2) % H;
    kyTab[e] = row < 0 ? row + H : row;
  }

  float2 F[FPT];
#pragma unroll
  for (int s = 0; s < FPT; ++s) F[s] = make_float2(0.f, 0.f);

  for (int r = 0; r < R; ++r) {
    for (int bi = 0; bi < nbat; ++bi) {
      const int t0 = bi * RB;
      const int rows = min(RB, Hc - t0);
      // the batch fetched earlier goes to LDS; start fetching the next one
#pragma unroll
      for (int s = 0; s < kPre; ++s) {
        const int e = tid + s * kFThreads;
        if (e < rows * n) rowbuf[e] = pre[s];
      }
      {
        int nr = r, nb = bi + 1;
        if (nb == nbat) {
          nb = 0;
          ++nr;
        }
        if (nr < R) fetch_rows(pre, frames, u16, fbase, g.frameW, R, nr, nb * RB, min(RB, Hc - nb * RB), n, mg.n);
      }
      __syncthreads();
      lds_fft_ip<false>(rowbuf, rows, n, 1, false, a.row, twn, RB);
      // unpack the packed real transform at the window columns
      for (int e = tid; e < rows * w; e += kFThreads) {
        const int tt = fdiv(e, mg.w), i = e - tt * w;
        const int k = c0 + i;
        const float2 zk = rowbuf[tt * n + k];
        float2 zc = rowbuf[tt * n + (k == 0 ? 0 : n - k)];
        zc.y = -zc.y;
        const float2 ev = cscale(cadd(zk, zc), 0.5f);
        const float2 od = mul_mi<false>(cscale(csub(zk, zc), 0.5f));
        band[(t0 + tt) * w + i] = cadd(ev, cmul(twr[i], od));
      }
      __syncthreads();
    }
    // Hc-point FFT down each of the w band columns
    lds_fft_ip<false>(band, w, 1, w, true, a.col, twc, a.chunkCol);
    // accumulate this residue into the window pixels each thread owns
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      const int e = tid + s * kFThreads;
      if (e < hw) {
        const int j = fdiv(e, mg.w), i = e - j * w;
        const int ky = kyTab[j];
        const int kk = ky - fdiv(ky, mg.Hc) * Hc;
        const int pr = ky * r;
        const int idx = pr - fdiv(pr, mg.H) * H;
        F[s] = cadd(F[s], cmul(twy[idx], band[kk * w + i]));
      }
    }
    __syncthreads();
  }
  // window factor, then the field plane in LDS
  const float2* fac = factor + (size_t)p * hw;
#pragma unroll
  for (int s = 0; s < FPT; ++s) {
    const int e = tid + s * kFThreads;
    if (e < hw) fbuf[e] = cmul(F[s], fac[e]);
  }
  __syncthreads();
  lds_fft_ip<true>(fbuf, h, w, 1, false, a.ix, twix, a.chunkIx);   // along x
  lds_fft_ip<true>(fbuf, w, 1, w, true, a.iy, twiy, a.chunkIy);    // along y
  const float2* rp = ref + (size_t)p * hw;
  float2* out = fields + (size_t)(b * g.P + p) * hw;
  for (int e = tid; e < hw; e += kFThreads) {
    const float2 t = cmul(fbuf[e], rp[e]);
    out[e] = make_float2(-t.x, t.y);
  }
}

// Most transforms of plan pl one in-place pass can hold (0: unsupported radix).
static int ip_chunk(const FFTPlan& pl) {
  int best = 1 << 30;
  for (int s = 0; s < pl.nstages; ++s) {
    const int R = pl.radix[s];
    if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) return 0;
    const int bmax = kMaxV / R > 0 ? kMaxV / R : 1;
    best = std::min(best, bmax * kFThreads * R / pl.n);
  }
  return best;
}

static int fpt_of(const FusedArgs& a) { return (a.g.h * a.g.w + kFThreads - 1) / kFThreads; }

bool fused_prepare(FusedArgs& a) {
  if (a.R < 1 || a.Hc * a.R != a.g.H) return false;
  if (fpt_of(a) > kMaxFPT) return false;
  const int cr = ip_chunk(a.row), cc = ip_chunk(a.col), cx = ip_chunk(a.ix), cy = ip_chunk(a.iy);
  if (cr < 1 || cc < 1 || cx < 1 || cy < 1) return false;
  a.RB = std::min(std::min(a.Hc, cr), kPre * kFThreads / a.row.n);
  if (a.RB < 1) return false;
  a.chunkCol = cc;
  a.chunkIx = cx;
  a.chunkIy = cy;
  return fused_layout(a).bytes <= kFusedLdsMax;
}

template <int FPT>
static hipError_t fused_launch(const FusedArgs& a, const FusedMagic& mg, const float2* twW, const float2* twH,
                               const void* frames, int u16, const float2* factor, const float2* ref,
                               float2* fields, hipStream_t s) {
  const size_t lds = fused_layout(a).bytes;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_fused_window<FPT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  dim3 grid(a.g.P, a.g.B);
  k_fused_window<FPT><<<grid, kFThreads, lds, s>>>(a, mg, twW, twH, frames, u16, factor, ref, fields);
  return hipGetLastError();
}

// ===========================================================================
// MFMA column variant (uint16 frames).  Rows are transformed as above, RB rows
// at a time in natural order, and the column transform restricted to the h
// window rows is a complex GEMM on the matrix cores:
//     F[j][c] = sum_y D[j][y] X[y][c],   D[j][y] = exp(-2 pi i ky(j) y / H),
// X the band of row spectra, ky(j) the spectrum row of window row j.  Real form
// over k = 2y + {re, im} (K = 2 RB per row batch):
//     Fr = Ar . B,  Fi = Ai . B,  B[2y][c] = Xr, B[2y+1][c] = Xi,
//     Ar[j][2y] = Dr, Ar[j][2y+1] = -Di,  Ai[j][2y] = Di, Ai[j][2y+1] = Dr.
// Every operand is split x = hi + lo (fp16) and hi*hi + hi*lo + lo*hi run on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation.  X is scaled by a fixed power
// of two chosen from the bound |X| <= W * 65535 of uint16 frames, which keeps
// every value inside the fp16 range; D needs no scale.  A (= D, per pol, built
// once with the tables) is read from global memory straight into registers;
// B goes through LDS.  Combos (j tile, re/im part) are dealt to the waves; a
// wave keeps its combos' accumulators for every column tile.
// ===========================================================================

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int mfma_bpitch(int RB) { return 2 * RB + 8; }   // halfs per B row (one column c)

struct MfmaLayout {
  int region;    // bytes of (row buffer + B planes) / field plane
  size_t bytes;
};

__host__ __device__ inline MfmaLayout mfma_layout(const FusedArgs& a) {
  MfmaLayout L;
  const int fwd = a.RB * a.row.n * 8 + 2 * a.g.w * mfma_bpitch(a.RB) * 2;
  const int inv = a.g.h * a.g.w * 8;
  L.region = ((fwd > inv ? fwd : inv) + 15) & ~15;
  L.bytes = (size_t)L.region + sizeof(float2) * (size_t)(a.row.n + a.g.w + a.g.w + a.g.h);
  return L;
}

template <int CPW, int NU>
__global__ __launch_bounds__(kFThreads, 4) void k_window_mfma(FusedArgs a, FusedMagic mg, float xscale,
                                                              const float2* __restrict__ twW,
                                                              const unsigned short* __restrict__ frames,
                                                              const _Float16* __restrict__ dtab,
                                                              const float2* __restrict__ factor,
                                                              const float2* __restrict__ ref,
                                                              float2* __restrict__ fields) {
  extern __shared__ __align__(16) unsigned char wm_smem[];
  const Geometry& g = a.g;
  const int p = blockIdx.x, b = blockIdx.y;
  const int n = a.row.n, RB = a.RB, H = g.H, w = g.w, h = g.h, hw = h * w;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, q = lane >> 4;
  const MfmaLayout L = mfma_layout(a);
  float2* rowbuf = reinterpret_cast<float2*>(wm_smem);                                 // [RB][n]
  const int pk = mfma_bpitch(RB);
  _Float16* Bl = reinterpret_cast<_Float16*>(wm_smem + (size_t)RB * n * sizeof(float2));   // [2][w][pk]
  float2* fbuf = reinterpret_cast<float2*>(wm_smem);                                   // [h][w] at the end
  float2* twn = reinterpret_cast<float2*>(wm_smem + L.region);                         // n
  float2* twr = twn + n;                                                               // w
  float2* twix = twr + w;                                                              // w
  float2* twiy = twix + w;                                                             // h
  const int c0 = g.c0[p];
  const size_t fbase = (size_t)b * g.frameH * g.frameW + (size_t)g.cy[p] * g.frameW + g.cx[p];
  const int K = 2 * H;
  const int nCombo = 2 * (h / 16);
  const _Float16* tabp = dtab + (size_t)p * 4 * h * K;

  float2 pre[kPre];
  fetch_rows(pre, frames, 1, fbase, g.frameW, 1, 0, 0, RB, n, mg.n);
  for (int e = tid; e < n; e += kFThreads) twn[e] = a.row.tw[e];
  for (int e = tid; e < w; e += kFThreads) {
    twr[e] = twW[c0 + e];
    twix[e] = a.ix.tw[e];
  }
  for (int e = tid; e < h; e += kFThreads) twiy[e] = a.iy.tw[e];

  floatx4 acc[CPW][NU];
#pragma unroll
  for (int ci = 0; ci < CPW; ++ci)
#pragma unroll
    for (int u = 0; u < NU; ++u) acc[ci][u] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int nch = H / RB;
  for (int ch = 0; ch < nch; ++ch) {
#pragma unroll
    for (int s = 0; s < kPre; ++s) {
      const int e = tid + s * kFThreads;
      if (e < RB * n) rowbuf[e] = pre[s];
    }
    if (ch + 1 < nch) fetch_rows(pre, frames, 1, fbase, g.frameW, 1, 0, (ch + 1) * RB, RB, n, mg.n);
    __syncthreads();
    lds_fft_ip<false>(rowbuf, RB, n, 1, false, a.row, twn, RB);
    // unpack the window columns of the packed real transform, scale, split into B
    for (int e = tid; e < RB * w; e += kFThreads) {
      const int tt = fdiv(e, mg.w), i = e - tt * w;
      const int k = c0 + i;
      const float2 zk = rowbuf[tt * n + k];
      float2 zc = rowbuf[tt * n + (k == 0 ? 0 : n - k)];
      zc.y = -zc.y;
      const float2 ev = cscale(cadd(zk, zc), 0.5f);
      const float2 od = mul_mi<false>(cscale(csub(zk, zc), 0.5f));
      const float2 X = cscale(cadd(ev, cmul(twr[i], od)), xscale);
      const _Float16 rh = (_Float16)X.x, ih = (_Float16)X.y;
      const _Float16 rl = (_Float16)(X.x - (float)rh), il = (_Float16)(X.y - (float)ih);
      _Float16* bp = Bl + i * pk + 2 * tt;
      bp[0] = rh;
      bp[1] = ih;
      bp[w * pk] = rl;
      bp[w * pk + 1] = il;
    }
    __syncthreads();
    // F += A[:, chunk] . B_chunk for this wave's combos
    const int kb = ch * 2 * RB;
    for (int s = 0; s < 2 * RB; s += 32) {
      half8 ah[CPW], al[CPW];
#pragma unroll
      for (int ci = 0; ci < CPW; ++ci) {
        const int cb = wave + 8 * ci;
        if (cb < nCombo) {
          const int t = cb >> 1, part = cb & 1;
          const _Float16* ap = tabp + ((size_t)(part * 2) * h + 16 * t + lr) * K + kb + s + 8 * q;
          ah[ci] = *reinterpret_cast<const half8*>(ap);
          al[ci] = *reinterpret_cast<const half8*>(ap + (size_t)h * K);
        }
      }
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const _Float16* bq = Bl + (16 * u + lr) * pk + s + 8 * q;
        const half8 bh = *reinterpret_cast<const half8*>(bq);
        const half8 bl = *reinterpret_cast<const half8*>(bq + w * pk);
#pragma unroll
        for (int ci = 0; ci < CPW; ++ci) {
          if (wave + 8 * ci < nCombo) {
            acc[ci][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ci], bh, acc[ci][u], 0, 0, 0);
            acc[ci][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ci], bl, acc[ci][u], 0, 0, 0);
            acc[ci][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[ci], bh, acc[ci][u], 0, 0, 0);
          }
        }
      }
    }
  }
  __syncthreads();   // B and the row buffer are free: the field plane takes their place
  const float inv = 1.f / xscale;
  float* fl = reinterpret_cast<float*>(fbuf);
#pragma unroll
  for (int ci = 0; ci < CPW; ++ci) {
    const int cb = wave + 8 * ci;
    if (cb < nCombo) {
      const int t = cb >> 1, part = cb & 1;
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int i = 0; i < 4; ++i) fl[2 * ((16 * t + 4 * q + i) * w + 16 * u + lr) + part] = acc[ci][u][i] * inv;
    }
  }
  __syncthreads();
  const float2* fac = factor + (size_t)p * hw;
  for (int e = tid; e < hw; e += kFThreads) fbuf[e] = cmul(fbuf[e], fac[e]);
  __syncthreads();
  lds_fft_ip<true>(fbuf, h, w, 1, false, a.ix, twix, a.chunkIx);   // along x
  lds_fft_ip<true>(fbuf, w, 1, w, true, a.iy, twiy, a.chunkIy);    // along y
  const float2* rp = ref + (size_t)p * hw;
  float2* out = fields + (size_t)(b * g.P + p) * hw;
  for (int e = tid; e < hw; e += kFThreads) {
    const float2 t = cmul(fbuf[e], rp[e]);
    out[e] = make_float2(-t.x, t.y);
  }
}

// D as the A operands: tab[p][part][hi/lo][j][k], k = 2y + {0, 1}, FP64 angles
__global__ void k_coldft_table(Geometry g, _Float16* __restrict__ tab) {
  const size_t per = (size_t)g.h * g.H;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)g.P * per) return;
  const int p = (int)(e / per);
  const int rem = (int)(e - (size_t)p * per);
  const int j = rem / g.H, y = rem - j * g.H;
  int ky = (j + g.SY[p] + g.H / 2) % g.H;
  if (ky < 0) ky += g.H;
  const long long m = ((long long)ky * y) % g.H;
  const double ang = -2.0 * 3.14159265358979323846 * (double)m / g.H;
  double sn, cs;
  sincos(ang, &sn, &cs);
  const int K = 2 * g.H;
  auto put = [&](int part, int k, double v) {
    const _Float16 hi = (_Float16)v;
    const _Float16 lo = (_Float16)(v - (double)hi);
    _Float16* base = tab + ((size_t)(p * 2 + part) * 2 * g.h) * K;
    base[(size_t)j * K + k] = hi;
    base[(size_t)(g.h + j) * K + k] = lo;
  };
  put(0, 2 * y, cs);
  put(0, 2 * y + 1, -sn);
  put(1, 2 * y, sn);
  put(1, 2 * y + 1, cs);
}

size_t coldft_table_elems(const Geometry& g) { return (size_t)g.P * 4 * g.h * 2 * g.H; }

hipError_t launch_coldft_table(const Geometry& g, void* tab, hipStream_t s) {
  const size_t total = (size_t)g.P * g.h * g.H;
  if (total == 0) return hipErrorInvalidValue;
  k_coldft_table<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(g, static_cast<_Float16*>(tab));
  return hipGetLastError();
}

static int mfma_cpw(const Geometry& g) { return (2 * (g.h / 16) + 7) / 8; }

bool window_mfma_prepare(FusedArgs& a) {
  const Geometry& g = a.g;
  if (g.w % 16 || g.h % 16 || g.w < 16 || g.h < 16) return false;
  const int nu = g.w / 16, cpw = mfma_cpw(g);
  const bool inst = (cpw == 1 && (nu == 3 || nu == 4 || nu == 5 || nu == 6 || nu == 8)) ||
                    (cpw == 2 && (nu == 5 || nu == 6 || nu == 8));
  if (!inst) return false;
  const int cr = ip_chunk(a.row), cx = ip_chunk(a.ix), cy = ip_chunk(a.iy);
  if (cr < 1 || cx < 1 || cy < 1) return false;
  int RB = std::min(32, std::min(cr, kPre * kFThreads / a.row.n));
  RB -= RB % 16;
  if (RB < 16) return false;
  if (g.H % RB) RB = 16;
  if (g.H % RB) return false;
  a.R = 1;
  a.Hc = g.H;
  a.RB = RB;
  a.chunkCol = 0;
  a.chunkIx = cx;
  a.chunkIy = cy;
  return mfma_layout(a).bytes <= kFusedLdsMax;
}

template <int CPW, int NU>
static hipError_t wm_launch(const FusedArgs& a, const FusedMagic& mg, float xscale, const float2* twW,
                            const void* frames, const void* dtab, const float2* factor, const float2* ref,
                            float2* fields, hipStream_t s) {
  const size_t lds = mfma_layout(a).bytes;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_window_mfma<CPW, NU>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  dim3 grid(a.g.P, a.g.B);
  k_window_mfma<CPW, NU><<<grid, kFThreads, lds, s>>>(a, mg, xscale, twW,
                                                       static_cast<const unsigned short*>(frames),
                                                       static_cast<const _Float16*>(dtab), factor, ref, fields);
  return hipGetLastError();
}

hipError_t launch_window_mfma(const FusedArgs& a0, const float2* twW, const void* frames_u16, const void* dtab,
                              const float2* factor, const float2* ref, float2* fields, hipStream_t s) {
  FusedArgs a = a0;
  if (!window_mfma_prepare(a)) return hipErrorNotSupported;
  FusedMagic mg{magic_div(a.row.n), magic_div(a.g.w), 0u, 0u};
  // |X| <= W * 65535 < 2^(16 + log2 W): scale into [2^14, 2^15) at most
  int lw = 0;
  while ((1 << lw) < a.g.W) ++lw;
  const float xscale = ldexpf(1.f, 15 - (16 + lw));
  const int nu = a.g.w / 16, cpw = mfma_cpw(a.g);
#define WM_CASE(C, N) \
  if (cpw == C && nu == N) return wm_launch<C, N>(a, mg, xscale, twW, frames_u16, dtab, factor, ref, fields, s);
  WM_CASE(1, 3) WM_CASE(1, 4) WM_CASE(1, 5) WM_CASE(1, 6) WM_CASE(1, 8)
  WM_CASE(2, 5) WM_CASE(2, 6) WM_CASE(2, 8)
#undef WM_CASE
  return hipErrorNotSupported;
}

hipError_t launch_fused_window(const FusedArgs& a0, const float2* twW, const float2* twH, const void* frames,
                               int frame_is_u16, const float2* factor, const float2* ref, float2* fields,
                               hipStream_t s) {
  FusedArgs a = a0;
  if (!fused_prepare(a)) return hipErrorNotSupported;
  FusedMagic mg{magic_div(a.row.n), magic_div(a.g.w), magic_div(a.Hc), magic_div(a.g.H)};
  const int fpt = fpt_of(a);
  if (fpt <= 4) return fused_launch<4>(a, mg, twW, twH, frames, frame_is_u16, factor, ref, fields, s);
  if (fpt <= 8) return fused_launch<8>(a, mg, twW, twH, frames, frame_is_u16, factor, ref, fields, s);
  return fused_launch<16>(a, mg, twW, twH, frames, frame_is_u16, factor, ref, fields, s);
}

}  // namespace dh
