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
#include <type_traits>

#include "dh_fft_ip.h"

namespace dh {

constexpr int kFThreads = 512;
constexpr int kPre = 8;                     // prefetched row elements per thread
constexpr size_t kFusedLdsMax = 80 * 1024;  // two workgroups per CU
constexpr int kMaxFPT = 16;                 // window pixels accumulated per thread

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
  const int p = blockIdx.x, b = blockIdx.y + g.b0;
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
      lds_fft_ip<false, kFThreads>(rowbuf, rows, n, 1, false, a.row, twn, RB);
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
    lds_fft_ip<false, kFThreads>(band, w, 1, w, true, a.col, twc, a.chunkCol);
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
  const float2* fac = factor + (size_t)(lam_in(g, b) * g.P + p) * hw;
#pragma unroll
  for (int s = 0; s < FPT; ++s) {
    const int e = tid + s * kFThreads;
    if (e < hw) fbuf[e] = cmul(F[s], fac[e]);
  }
  __syncthreads();
  lds_fft_ip<true, kFThreads>(fbuf, h, w, 1, false, a.ix, twix, a.chunkIx);   // along x
  lds_fft_ip<true, kFThreads>(fbuf, w, 1, w, true, a.iy, twiy, a.chunkIy);    // along y
  const float2* rp = ref + (size_t)(lam_in(g, b) * g.P + p) * hw;
  float2* out = fields + (size_t)(out_pos(g, b) * g.P + p) * hw;
  for (int e = tid; e < hw; e += kFThreads) {
    const float2 t = cmul(fbuf[e], rp[e]);
    out[e] = make_float2(-t.x, t.y);
  }
}

static int fpt_of(const FusedArgs& a) { return (a.g.h * a.g.w + kFThreads - 1) / kFThreads; }

bool fused_prepare(FusedArgs& a) {
  if (a.g.fw != a.g.w || a.g.fh != a.g.h) return false;
  if (a.R < 1 || a.Hc * a.R != a.g.H) return false;
  if (fpt_of(a) > kMaxFPT) return false;
  const int cr = ip_chunk(a.row, kFThreads), cc = ip_chunk(a.col, kFThreads);
  const int cx = ip_chunk(a.ix, kFThreads), cy = ip_chunk(a.iy, kFThreads);
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
  return for_frame_chunks(a.g.B, [&](int b0, int nb) {
    FusedArgs ac = a;
    ac.g.b0 = b0;
    k_fused_window<FPT><<<dim3(a.g.P, nb), kFThreads, lds, s>>>(ac, mg, twW, twH, frames, u16, factor, ref, fields);
    return hipGetLastError();
  });
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
  const int p = blockIdx.x, b = blockIdx.y + g.b0;
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
    lds_fft_ip<false, kFThreads>(rowbuf, RB, n, 1, false, a.row, twn, RB);
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
  const float2* fac = factor + (size_t)(lam_in(g, b) * g.P + p) * hw;
  for (int e = tid; e < hw; e += kFThreads) fbuf[e] = cmul(fbuf[e], fac[e]);
  __syncthreads();
  lds_fft_ip<true, kFThreads>(fbuf, h, w, 1, false, a.ix, twix, a.chunkIx);   // along x
  lds_fft_ip<true, kFThreads>(fbuf, w, 1, w, true, a.iy, twiy, a.chunkIy);    // along y
  const float2* rp = ref + (size_t)(lam_in(g, b) * g.P + p) * hw;
  float2* out = fields + (size_t)(out_pos(g, b) * g.P + p) * hw;
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
  if (g.fw != g.w || g.fh != g.h) return false;
  if (g.w % 16 || g.h % 16 || g.w < 16 || g.h < 16) return false;
  const int nu = g.w / 16, cpw = mfma_cpw(g);
  const bool inst = (cpw == 1 && (nu == 3 || nu == 4 || nu == 5 || nu == 6 || nu == 8)) ||
                    (cpw == 2 && (nu == 5 || nu == 6 || nu == 8));
  if (!inst) return false;
  const int cr = ip_chunk(a.row, kFThreads), cx = ip_chunk(a.ix, kFThreads), cy = ip_chunk(a.iy, kFThreads);
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
  return for_frame_chunks(a.g.B, [&](int b0, int nb) {
    FusedArgs ac = a;
    ac.g.b0 = b0;
    k_window_mfma<CPW, NU><<<dim3(a.g.P, nb), kFThreads, lds, s>>>(ac, mg, xscale, twW,
                                                                  static_cast<const unsigned short*>(frames),
                                                                  static_cast<const _Float16*>(dtab), factor, ref,
                                                                  fields);
    return hipGetLastError();
  });
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

// ===========================================================================
// Size-specialised window kernel (uint16 frames), 256 threads per (frame, pol).
//
// Row transform, in registers.  A row of W = 2 NR reals is packed into NR
// complex points z[n] = I[2n] + i I[2n+1] and transformed by L lanes holding M
// points each (NR = L M, n = l + L m):
//     pass A (per lane):   v[a] = w_NR^(l a) sum_m z[l + L m] w_M^(m a)
//     pass B (over lanes): Z[a + M b] = sum_l w_L^(l b) v_l[a]
// The transpose between the passes goes through a wave-private LDS region in
// two halves, so no block barrier is involved.  Pass B deals the a values to
// lanes in pairs (a, M - a) (and (0, M/2)), so that Z[k] and its r2c partner
// Z[NR - k] sit in the same lane and the real-to-complex unpack
//     X[k] = (Z[k] + Z*[NR-k])/2 - i w_W^k (Z[k] - Z*[NR-k])/2
// needs no further exchange.  Only columns c0 <= k < c0 + w are kept.
//
// Column transform restricted to the h window rows, a complex GEMM on MFMA:
//     F[j][c] = sum_y D[j][y] X[y][c],  D[j][y] = exp(-2 pi i ky(j) y / H),
// split into fp16 hi + lo with hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16
// and FP32 accumulation (the operand layout of k_window_mfma).  The D operands
// are read per k-step from an LDS table holding, for each m, cos and sin of
// 2 pi m / H already split into fp16 hi and lo, so nothing but the frame
// streams from memory.  Row-spectrum values outside the window are written to
// a spare B column instead of being branched around.  A round is 4 RPW window rows (each wave
// transforms RPW of them); units (j-tile, re/im part) are dealt to the waves.
//
// With hgx16 set, the epilogue also computes the HG coefficients from the
// field plane left in LDS: the split-fp16 separable overlap of k_overlap_mfma
// (field scaled by a power of two to a peak in [128, 256), stage 1 per 16 field
// rows T = V HGx^T for the real and imaginary parts, stage 2 C += T^T HGy^T),
// with the basis operands read from L2 and the stage-1 output double-buffered
// in LDS, so coefficients leave the kernel with the fields.
//
// Epilogue: the window plane goes to LDS (row pitch w + 4), and each inverse
// transform of N = 16 U points is done by 4 lanes holding N/4 points (the
// same two-pass split, in place in the plane), 16 transforms per wave pass:
// rows with the window factor applied on load, then columns with the
// reference-wave tilt applied on the way to HBM.
// ===========================================================================

template <int B0, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B0 < E) {
    f(std::integral_constant<int, B0>{});
    sfor<B0 + 1, E>(f);
  }
}

constexpr double kPiD = 3.14159265358979323846;

// sin(x) by its Taylor series, |x| <= 3 pi / 2 (compile time only)
constexpr double ct_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 26; ++n) {
    t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    s += t;
  }
  return s;
}

// v * exp(-/+ 2 pi i E / N) for compile-time E, N
template <int E, int N, bool INV>
__device__ __forceinline__ float2 tw_mul(float2 v) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (4 * e == N) {
    return mul_mi<INV>(v);
  } else if constexpr (2 * e == N) {
    return make_float2(-v.x, -v.y);
  } else if constexpr (4 * e == 3 * N) {
    return mul_mi<!INV>(v);
  } else {
    constexpr double th = 2.0 * kPiD * (double)(2 * e <= N ? e : e - N) / (double)N;   // (-pi, pi]
    constexpr float c = (float)ct_sin(kPiD / 2 - th);
    constexpr float sn = (float)ct_sin(th);
    return cmul(v, make_float2(c, INV ? sn : -sn));
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v);

// In-register DFT of P Q points, natural order in and out:
//   V[a1 + Q a2] = sum_p w_P^(p a2) w_PQ^(p a1) sum_q v[p + P q] w_Q^(q a1)
template <int P, int Q, bool INV>
__device__ __forceinline__ void fft_pq(float2* v) {
  constexpr int M = P * Q;
  float2 y[M];
  sfor<0, P>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    float2 t[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) t[q] = v[p + P * q];
    dft_any<Q, INV>(t);
    sfor<0, Q>([&](auto ac) {
      constexpr int a1 = decltype(ac)::value;
      y[a1 * P + p] = tw_mul<p * a1, M, INV>(t[a1]);
    });
  });
#pragma unroll
  for (int a1 = 0; a1 < Q; ++a1) {
    float2 t[P];
#pragma unroll
    for (int p = 0; p < P; ++p) t[p] = y[a1 * P + p];
    dft_any<P, INV>(t);
#pragma unroll
    for (int a2 = 0; a2 < P; ++a2) v[a1 + Q * a2] = t[a2];
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v) {
  if constexpr (R == 16) fft_pq<4, 4, INV>(v);
  else if constexpr (R == 24) fft_pq<3, 8, INV>(v);
  else if constexpr (R == 12) fft_pq<3, 4, INV>(v);
  else if constexpr (R == 20) fft_pq<4, 5, INV>(v);
  else if constexpr (R == 32) fft_pq<4, 8, INV>(v);
  else if constexpr (R == 10) fft_pq<2, 5, INV>(v);
  else if constexpr (R == 1) (void)v;
  else dftR<R, INV>(v);
}

// LDS order point within a wave: every lane's earlier LDS accesses have
// completed before any later one is issued.
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));

// (xr, xi) split into fp16 pairs hi = (fp16(xr), fp16(xi)) and
// lo = (fp16(xr - hi.x), fp16(xi - hi.y)), both rounded to nearest even: the
// residuals are exact in FP32, and v_fma_mix forms them from the fp16 halves
// of hi and rounds them to fp16 in one instruction each (the same values as
// converting, subtracting and converting again, in three instructions
// instead of eight).  hi is an ordinary conversion, so the compiler sees it
// read xr and xi and keeps the wait states an MFMA result needs before a VALU
// reads it; the inline v_fma_mix depend on hi and so come after it (the
// compiler does not check the operands of inline assembly for that hazard).
__device__ __forceinline__ void split_f16x2(float xr, float xi, unsigned& hi, unsigned& lo) {
  const half2v hv = {(_Float16)xr, (_Float16)xi};
  hi = __builtin_bit_cast(unsigned, hv);
  asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %0, %1, -1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=&v"(lo)
      : "v"(hi), "v"(xr), "v"(xi));
}

constexpr int kW3Threads = 256;

// R: the MFMA row transform where it exists (128-point rows); false keeps the
// register transform (the autotuner times both)
template <int L, int M, int U, int J, bool R = true>
struct W3Shape {
  static constexpr int NR = L * M;            // complex points per row transform
  // MR: the row transform on MFMA (128-point rows: a 16-point DFT, then per
  // residue pair one matrix for the 8-point DFT, its twiddles and the r2c
  // unpack), in rounds of 16 rows; the exchange holds whole rows, grouped by
  // residue pair, and it and the B buffer are double-buffered
  static constexpr bool MR = R && L == 8 && M == 16;
  static constexpr int RPW = 64 / L;          // rows per wave per round (register row transform)
  static constexpr int RR = MR ? 16 : 4 * RPW;   // rows per round
  static constexpr int KS = RR / 16;          // MFMA k-steps per round
  static constexpr int S = M / L / 2;         // (a, M - a) pairs per lane
  static constexpr int w = 16 * U, h = 16 * J;
  static constexpr int NU = (2 * J + 3) / 4;  // (j-tile, part) units per wave
  static constexpr int LP = M / 2 + 2;        // exchange: complex per lane (half)
  static constexpr int RPc = L * LP + (8 - (L * LP) % 16 + 16) % 16;   // per row, = 8 mod 16
  static constexpr int PK = 2 * RR + 8;       // halfs per B column
  static constexpr int EPP = 16 * 8 + 16;     // MR exchange: bytes per (row, pair), 16 complex + pad
  static constexpr int ERP = 8 * EPP + 16;    // MR exchange: bytes per row
  static constexpr int EBYTES = MR ? 2 * RR * ERP : 4 * RPW * RPc * 8;
  static constexpr int BBYTES = 2 * 2 * (w + 1) * PK * 2;   // two buffers; column w takes discarded values
  static constexpr int WP = w + 4;            // field-plane row pitch (complex)
};

constexpr int kMrMats = 10;
constexpr int kMrMatHalfs = 2 * 2 * 64 * 8;
constexpr double kMrScaleA = 1.0 / 32.0;

// the column k of output o of residue pair v
__host__ __device__ inline int kmr(int v, int o) {
  if (v == 0 || v == 8) {
    if (o == 0) return v == 8 ? 128 : 0;
    return o < 8 ? 16 * o : 8 + 16 * (o - 8);
  }
  return o < 8 ? v + 16 * o : 16 - v + 16 * (o - 8);
}

// offset (complex elements, 16-byte aligned) of the MR fragments in the table
__host__ __device__ inline int w3_mr_off(int H, int L, int M) { return (H + L * (M + 1) + M + 1) & ~1; }
// MFMA row path: the column product's A operands after the pair matrices
// (float2 units; 16-byte aligned), [P][H / 16][2 J][hi, lo][64 lanes] uint4
__host__ __device__ inline int w3_atab_off(int H, int L, int M) { return w3_mr_off(H, L, M) + kMrMats * kMrMatHalfs / 4; }
__host__ __device__ inline size_t w3_atab_elems(int P, int H, int J) { return (size_t)P * (H / 16) * (2 * J) * 256; }

// LDS bytes of the aliased region: the row phase or the field plane
__host__ __device__ inline int w3_region(int ebytes, int bbytes, int H, int L, int M, int w, int h) {
  const int row = ebytes + bbytes + 8 * H + 8 * L * (M + 1) + 8 * H;
  const int pl = 8 * h * (w + 4);
  return ((row > pl ? row : pl) + 15) & ~15;
}

struct W3Args {
  Geometry g;
  FFTPlan ix, iy;                     // twiddle tables of the w- and h-point transforms
  // fused split-fp16 HG overlap (null hgx16: coefficients computed elsewhere)
  const _Float16* hgx16;              // [P][hi/lo][Gp][w], rows m >= G zero
  const _Float16* hgy16;              // [P][hi/lo][Gp][h]
  // fused FP32 HG overlap instead (null: none): the basis [P][G][w], [P][G][h]
  const float* hgx32;
  const float* hgy32;
  float2* coefs;                      // [B][P][M]
  BatchCal cal;                       // per-frame window calibration
  int region;         // bytes of the aliased row-phase / field-plane region
  int ovlOff;         // byte offset of the overlap's stage-1 buffers (after the plane)
  unsigned magH;      // q / H == __umulhi(q, magH) for q < 2^20
  float xscale;       // power of two keeping the row spectra inside fp16
};

constexpr int kOvMaxTiles = 4;   // stage-2 tiles per wave (G up to ~50)

// Split-fp16 separable overlap of the h x w field in the LDS plane (pitch
// w + 4), 256 threads.  Writes coefs[b][p][mode], modes ordered by group g,
// k = 0..g (m = g - k, n = k).
//
// Stage 1 per 16-row chunk of the field: T_part[m][yy] = sum_x V_part[yy][x]
// HGx[m][x] on v_mfma_f32_16x16x32_f16, one unit (16 modes m, real or
// imaginary part) per output tile; lane (m, q) receives T[m][4q .. 4q + 3],
// which is exactly the A operand of v_mfma_f32_16x16x16_f16 (row m, k-slots
// 4q .. 4q + 3).  Stage 2, C_part[m][n] += sum_yy T_part[m][yy] HGy[n][yy],
// therefore runs on the same wave from registers with K = 16 = the chunk's
// rows: no exchange through LDS and no barrier per chunk.
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
template <int U, int J>
__device__ __forceinline__ void w3_overlap(const W3Args& a, const float2* plane, float vmax, float* red, int p,
                                           int b, int wave, int lane) {
  constexpr int w = 16 * U, h = 16 * J, WP = w + 4;
  const Geometry& g = a.g;
  const int G = g.G, Gp = (G + 15) & ~15, ct1 = Gp / 16;
  const int lr = lane & 15, q = lane >> 4;
  const int part = wave & 1;   // the wave's units u1 = wave + 4 i all have this part
  float scale = 1.f;   // per-field power of two, set below once the basis loads are in flight
  const _Float16* bx = a.hgx16 + (size_t)p * 2 * Gp * w;
  const _Float16* by = a.hgy16 + (size_t)p * 2 * Gp * h;
  auto split8 = [&](const float* v, half8& hi, half8& lo) {
    uint4 h4, l4;
    unsigned* hw = reinterpret_cast<unsigned*>(&h4);
    unsigned* lw = reinterpret_cast<unsigned*>(&l4);
#pragma unroll
    for (int e = 0; e < 4; ++e) split_f16x2(v[2 * e] * scale, v[2 * e + 1] * scale, hw[e], lw[e]);
    hi = __builtin_bit_cast(half8, h4);
    lo = __builtin_bit_cast(half8, l4);
  };
  // stage-1 units (m-tile, part) of this wave and their basis operands, which
  // are the same for every chunk
  constexpr int KS1 = (w + 31) / 32;
  constexpr int NU1 = 2;                       // units per wave: 2 ct1 <= 8 (Gp <= 64)
  constexpr int NT2 = 4;                       // stage-2 n-tiles: Gp / 16 <= 4
  half8 b1h[NU1][KS1], b1l[NU1][KS1];
#pragma unroll
  for (int i = 0; i < NU1; ++i) {
    const int u1 = wave + 4 * i;
    if (u1 < 2 * ct1) {
      const int mt = u1 >> 1;
#pragma unroll
      for (int ks = 0; ks < KS1; ++ks) {
        const int kb = 32 * ks + 8 * q;
        if (kb < w) {
          const _Float16* bp = bx + (16 * mt + lr) * w + kb;
          b1h[i][ks] = *reinterpret_cast<const half8*>(bp);
          b1l[i][ks] = *reinterpret_cast<const half8*>(bp + Gp * w);
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) b1h[i][ks][e] = b1l[i][ks][e] = (_Float16)0.f;
        }
      }
    }
  }
  // stage-2 basis operands of chunk ch (lane (n, q): HGy[16 nt + n][16 ch + 4q ..
  // + 3], rows n >= G zero), loaded one chunk ahead of their use
  auto load_b2 = [&](int ch, half4 (&dh)[NT2], half4 (&dl)[NT2]) {
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
      if (nt < Gp / 16) {
        const _Float16* brow = by + (nt * 16 + lr) * h + 16 * ch + 4 * q;
        dh[nt] = *reinterpret_cast<const half4*>(brow);
        dl[nt] = *reinterpret_cast<const half4*>(brow + Gp * h);
      }
    }
  };
  half4 b2h[NT2], b2l[NT2];
  load_b2(0, b2h, b2l);
  floatx4 cacc[NU1][NT2];
#pragma unroll
  for (int i = 0; i < NU1; ++i)
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) cacc[i][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  // per-field power-of-two scale: peak magnitude into [128, 256)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o));
  if (lane == 0) red[wave] = vmax;
  __syncthreads();   // also: the whole field plane is written
  vmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  if (vmax > 0.f) {
    int ex;
    frexpf(vmax, &ex);
    scale = ldexpf(1.f, 8 - ex);
  }
  for (int ch = 0; ch < J; ++ch) {
    half4 b2nh[NT2], b2nl[NT2];
    const bool more = ch + 1 < J;
    if (more) load_b2(ch + 1, b2nh, b2nl);
    // stage 1: the field operand is split once per k-step for both units
    floatx4 d[NU1];
#pragma unroll
    for (int i = 0; i < NU1; ++i) d[i] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks) {
      const int kb = 32 * ks + 8 * q;
      half8 ah, al;
      if (kb < w) {   // a ragged last step (w % 32 == 16) zero-fills the upper lane groups
        const float4* vp = reinterpret_cast<const float4*>(plane + (16 * ch + lr) * WP + kb);
        float vv[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float4 t = vp[e];
          vv[2 * e] = part ? t.y : t.x;
          vv[2 * e + 1] = part ? t.w : t.z;
        }
        split8(vv, ah, al);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) ah[e] = al[e] = (_Float16)0.f;
      }
#pragma unroll
      for (int i = 0; i < NU1; ++i) {
        if (wave + 4 * i < 2 * ct1) {
          d[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, b1h[i][ks], d[i], 0, 0, 0);
          d[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, b1l[i][ks], d[i], 0, 0, 0);
          d[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, b1h[i][ks], d[i], 0, 0, 0);
        }
      }
    }
    // stage 2 from registers: T (split into fp16 hi + lo) times the chunk's HGy
    // rows, over the n-tiles that hold some m + n < G
#pragma unroll
    for (int i = 0; i < NU1; ++i) {
      const int u1 = wave + 4 * i;
      if (u1 < 2 * ct1) {
        const int mt = u1 >> 1;
        uint2 th, tl;
        split_f16x2(d[i][0], d[i][1], th.x, tl.x);
        split_f16x2(d[i][2], d[i][3], th.y, tl.y);
        const half4 ah = __builtin_bit_cast(half4, th), al = __builtin_bit_cast(half4, tl);
#pragma unroll
        for (int nt = 0; nt < NT2; ++nt) {
          if (nt < Gp / 16 && 16 * mt + 16 * nt < G) {
            cacc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, b2h[nt], cacc[i][nt], 0, 0, 0);
            cacc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, b2l[nt], cacc[i][nt], 0, 0, 0);
            cacc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, b2h[nt], cacc[i][nt], 0, 0, 0);
          }
        }
      }
    }
    if (more) {
#pragma unroll
      for (int nt = 0; nt < NT2; ++nt) {
        b2h[nt] = b2nh[nt];
        b2l[nt] = b2nl[nt];
      }
    }
  }
  // the triangle m + n < G; lane (n, q) holds m = 16 mt + 4q .. + 3 of this part
  const float inv = 1.f / scale;
  float* out = reinterpret_cast<float*>(a.coefs + (size_t)(b * g.P + p) * g.M) + part;
#pragma unroll
  for (int i = 0; i < NU1; ++i) {
    const int u1 = wave + 4 * i;
    if (u1 < 2 * ct1) {
      const int mt = u1 >> 1;
#pragma unroll
      for (int nt = 0; nt < NT2; ++nt) {
        const int n = nt * 16 + lr;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int m = 16 * mt + 4 * q + e;
          if (nt < Gp / 16 && m + n < G) {
            const int gs = m + n;
            out[2 * (gs * (gs + 1) / 2 + n)] = cacc[i][nt][e] * inv;
          }
        }
      }
    }
  }
}

// FP32 form of the fused overlap (v_mfma_f32_16x16x4_f32, FP32 operands and
// accumulation; no scale): the same two stages and tiles as above, lane group q
// of every MFMA taking the contiguous k-slots q K/4 .. (q + 1) K/4 - 1 of A and B.
// Stage-1 output buffers at st1: 2 x [2 Gp][kOvPK32] floats.
constexpr int kOvPK32 = 20;   // floats per stage-1 output row (16 field rows + pad)
template <int U, int J>
__device__ __forceinline__ void w3_overlap32(const W3Args& a, const float2* plane, unsigned char* st1, int p, int b,
                                             int wave, int lane) {
  constexpr int w = 16 * U, h = 16 * J, WP = w + 4, KQ = w / 4;
  const Geometry& g = a.g;
  const int G = g.G, Gp = (G + 15) & ~15, ct1 = Gp / 16;
  const int lr = lane & 15, q = lane >> 4;
  const float* bx = a.hgx32 + (size_t)p * G * w;
  const float* by = a.hgy32 + (size_t)p * G * h;
  float* A2b = reinterpret_cast<float*>(st1);
  const int bufF = 2 * Gp * kOvPK32;   // floats per buffer
  const int rt2 = (2 * Gp) / 16, ct2 = Gp / 16;
  int myTile[kOvMaxTiles];
  int nMine = 0;
  {
    int cnt = 0;
    for (int t = 0; t < rt2 * ct2; ++t) {
      const int mt = t / ct2, nt = t - mt * ct2;
      if (8 * mt + 16 * nt >= G) continue;   // no (m, n) of the tile has m + n < G
      if (cnt % 4 == wave && nMine < kOvMaxTiles) myTile[nMine++] = t;
      ++cnt;
    }
  }
  floatx4 cacc[kOvMaxTiles];
#pragma unroll
  for (int i = 0; i < kOvMaxTiles; ++i) cacc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  // stage-1 basis operands of the wave's units (m-tile, part), every chunk
  constexpr int NU1 = 2;
  const int part = wave & 1;   // a wave's units u1 = wave + 4 i share it
  float b1[NU1][KQ];
#pragma unroll
  for (int i = 0; i < NU1; ++i) {
    const int u1 = wave + 4 * i;
    const int m = 16 * (u1 >> 1) + lr;
#pragma unroll
    for (int s4 = 0; s4 < KQ; ++s4)
      b1[i][s4] = (u1 < 2 * ct1 && m < G) ? bx[(size_t)m * w + q * KQ + s4] : 0.f;
  }
  float4 b2[kOvMaxTiles];
  for (int ch = 0; ch < J; ++ch) {
    float* A2 = A2b + (ch & 1) * bufF;
#pragma unroll
    for (int i = 0; i < kOvMaxTiles; ++i) {
      if (i < nMine) {
        const int n = (myTile[i] % ct2) * 16 + lr;
        b2[i] = n < G ? *reinterpret_cast<const float4*>(by + (size_t)n * h + 16 * ch + 4 * q)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    // stage 1: T_part[yy][m] = sum_x V_part[16 ch + yy][x] HGx[m][x]
    floatx4 d[NU1];
#pragma unroll
    for (int i = 0; i < NU1; ++i) d[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float2* vrow = plane + (16 * ch + lr) * WP + q * KQ;
#pragma unroll
    for (int s4 = 0; s4 < KQ; s4 += 2) {
      const float4 t = *reinterpret_cast<const float4*>(vrow + s4);
      const float a0 = part ? t.y : t.x, a1 = part ? t.w : t.z;
#pragma unroll
      for (int i = 0; i < NU1; ++i) {
        if (wave + 4 * i < 2 * ct1) {
          d[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1[i][s4], d[i], 0, 0, 0);
          d[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1[i][s4 + 1], d[i], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NU1; ++i) {
      const int u1 = wave + 4 * i;
      if (u1 < 2 * ct1) {
        const int row = 2 * (16 * (u1 >> 1) + lr) + part;
#pragma unroll
        for (int e = 0; e < 4; ++e) A2[row * kOvPK32 + 4 * q + e] = d[i][e];
      }
    }
    __syncthreads();   // this chunk's A2 is complete (the other buffer is free, as above)
    // stage 2: C[(m, part)][n] += sum_yy A2[(m, part)][yy] HGy[n][16 ch + yy]
#pragma unroll
    for (int i = 0; i < kOvMaxTiles; ++i) {
      if (i < nMine) {
        const int mt = myTile[i] / ct2;
        const float4 av = *reinterpret_cast<const float4*>(A2 + (mt * 16 + lr) * kOvPK32 + 4 * q);
        cacc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, b2[i].x, cacc[i], 0, 0, 0);
        cacc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, b2[i].y, cacc[i], 0, 0, 0);
        cacc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, b2[i].z, cacc[i], 0, 0, 0);
        cacc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, b2[i].w, cacc[i], 0, 0, 0);
      }
    }
  }
  // the triangle m + n < G
  float2* out = a.coefs + (size_t)(b * g.P + p) * g.M;
#pragma unroll
  for (int i = 0; i < kOvMaxTiles; ++i) {
    if (i < nMine) {
      const int t = myTile[i];
      const int mt = t / ct2, nt = t - mt * ct2;
      const int n = nt * 16 + lr;
#pragma unroll
      for (int ii = 0; ii < 4; ii += 2) {
        const int m = (mt * 16 + 4 * q + ii) >> 1;   // rows ii, ii + 1 are (m, re), (m, im)
        if (n < G && m < G && m + n < G) {
          const int gs = m + n;
          out[gs * (gs + 1) / 2 + n] = make_float2(cacc[i][ii], cacc[i][ii + 1]);
        }
      }
    }
  }
}

// Split basis tables for the fused overlap: dst[p][hl][m][x] for m < Gp (0 for m >= G)
__global__ void k_hg_split(const float* __restrict__ hg, int P, int G, int n, _Float16* __restrict__ dst) {
  const int Gp = (G + 15) & ~15;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * Gp * n) return;
  const int pp = e / (Gp * n), rem = e - pp * Gp * n, m = rem / n, x = rem - m * n;
  const float v = m < G ? hg[((size_t)pp * G + m) * n + x] : 0.f;
  const _Float16 hi = (_Float16)v;
  dst[((size_t)pp * 2 + 0) * Gp * n + rem] = hi;
  dst[((size_t)pp * 2 + 1) * Gp * n + rem] = (_Float16)(v - (float)hi);
}

// (pol, frame) of this workgroup in a one-dimensional grid of P B.  With two
// pols, workgroups 16 q + r and 16 q + r + 8 are the two pols of frame
// 8 q + r: the dispatcher deals workgroups to the eight XCDs in turn, so both
// land on the same XCD and the camera rows their windows share are read from
// HBM once and then from that XCD's L2.  The frames past the last full group
// of eight are mapped plainly.
__device__ __forceinline__ void w3_item(int P, int B, int id, int& p, int& b) {
  if (P != 2) {
    p = id % P;
    b = id / P;
    return;
  }
  const int full = (B >> 3) << 4;
  if (id < full) {
    const int q = id >> 4, r = id & 15;
    p = r >> 3;
    b = 8 * q + (r & 7);
  } else {
    p = (id - full) & 1;
    b = ((B >> 3) << 3) + ((id - full) >> 1);
  }
}

// SQ: a square FFT window (H = W = 2 L M), so the column transform's length
// and its table offsets are compile-time constants.  item: the (pol, frame)
// work item of the batch this workgroup processes (w3_item).
template <int L, int M, int U, int J, bool SQ, bool R = true>
__device__ __forceinline__ void w3_body(const W3Args& a, const unsigned short* __restrict__ frames,
                                        const float2* __restrict__ tabs, const float2* __restrict__ factor,
                                        const float2* __restrict__ ref, float2* __restrict__ fields, int item) {
  using Sh = W3Shape<L, M, U, J, R>;
  constexpr int NT = kW3Threads, RPW = Sh::RPW, RR = Sh::RR, KS = Sh::KS, S = Sh::S, w = Sh::w, h = Sh::h;
  constexpr int NU = Sh::NU, LP = Sh::LP, RPc = Sh::RPc, PK = Sh::PK, NR = Sh::NR, WP = Sh::WP;
  extern __shared__ __align__(16) unsigned char w3_smem[];
  const Geometry& g = a.g;
  const int H = SQ ? 2 * NR : g.H;
  int p, b;
  w3_item(a.g.P, a.g.B, item, p, b);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wv = (wave + b) & 3;              // units rotate with the frame so every SIMD gets its share
  const int r = lane / L, l = lane % L;       // row of the wave's pass, lane within the row
  const int lr = lane & 15, q = lane >> 4;    // MFMA operand coordinates
  const int c0 = g.c0[p];

  // row-phase layout from LDS address 0 (so the column-DFT table reads need no
  // base add): twT [part][hi, lo][H] (part 0 {cos, sin}, part 1 {-sin, cos}, fp16
  // pairs), twA [L][M + 1], the exchange rows, the two B buffers
  unsigned* twT = reinterpret_cast<unsigned*>(w3_smem);
  float2* twA = reinterpret_cast<float2*>(w3_smem + 16 * H);
  unsigned char* rowBase = w3_smem + 16 * H + 8 * L * (M + 1);
  float2* E = reinterpret_cast<float2*>(rowBase) + (wave * RPW + r) * RPc;      // this lane's exchange row
  _Float16* Bs0 = reinterpret_cast<_Float16*>(rowBase + Sh::EBYTES);           // [2 buffers][2][w + 1][PK]
  constexpr int BSTRIDE = 2 * (w + 1) * PK;
  float2* fbuf = reinterpret_cast<float2*>(w3_smem);                            // [h][WP] after the rows
  float2* twix = reinterpret_cast<float2*>(w3_smem + a.region);                 // [w]
  float2* twiy = twix + w;                                                      // [h]

  // the first round of this lane's row in flight while the tables are staged
  const size_t fw2 = (size_t)(g.frameW >> 1);
  const unsigned* src = reinterpret_cast<const unsigned*>(frames + ((size_t)b * g.frameH + g.cy[p]) * g.frameW +
                                                          g.cx[p]) + l + (size_t)(wave * RPW + r) * fw2;
  unsigned pre[M];
  // MR: pass-A operand words of this wave's two tiles of the round; tile t
  // holds rows 4 wave + 2t + (0, 1) of the round, tile row lr = 8 r' + l being
  // sequence l of row r'.  Lane (lr, q) holds the words l + 32 q + 8 e (e =
  // 0..3) of its row, i.e. z[l + 8 m], m = 4q + e: the 8 lanes of a row read
  // 32 contiguous bytes.
  unsigned fr[2][4];
  const unsigned* srcMR = reinterpret_cast<const unsigned*>(frames + ((size_t)b * g.frameH + g.cy[p]) * g.frameW +
                                                            g.cx[p]) + 32 * q + (lr & 7) +
                          (size_t)(4 * wave + (lr >> 3)) * fw2;
  auto load_mr = [&](int rd) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const unsigned* s = srcMR + (size_t)(rd * RR + 2 * t) * fw2;
#pragma unroll
      for (int e = 0; e < 4; ++e) fr[t][e] = __builtin_nontemporal_load(s + 8 * e);
    }
  };
  if constexpr (Sh::MR) {
    load_mr(0);
  } else {
#pragma unroll
    for (int m = 0; m < M; ++m) pre[m] = __builtin_nontemporal_load(src + L * m);
  }

  // (the MFMA row path reads its column-product operands from a table of their own)
  for (int e = tid; e < (Sh::MR ? 0 : H); e += NT) {
    const uint2 t = reinterpret_cast<const uint2*>(tabs)[e];
    twT[e] = t.x;
    twT[H + e] = t.y;
    // the imaginary-part operand (-sin, cos): halves swapped, then the low one negated
    twT[2 * H + e] = __builtin_amdgcn_alignbit(t.x, t.x, 16) ^ 0x8000u;
    twT[3 * H + e] = __builtin_amdgcn_alignbit(t.y, t.y, 16) ^ 0x8000u;
  }
  // the inter-pass twiddles carry the power-of-two scale of the B operand, so
  // pass B and the unpack produce scaled values (v[0] is scaled on its own)
  const float sc = 0.5f * a.xscale;           // ev + w od = 2 X
  if constexpr (!Sh::MR)
    for (int e = tid; e < L * (M + 1); e += NT) twA[e] = cscale(tabs[H + e], sc);
  for (int e = tid; e < w; e += NT) twix[e] = a.ix.tw[e];
  for (int e = tid; e < h; e += NT) twiy[e] = a.iy.tw[e];

  // this lane's a values in pass B and exp(-2 pi i a / W) for them
  int alo[S], ahi[S];
  float2 wlo[S], whi[S];
  if constexpr (!Sh::MR) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int pi = l * S + s;
      alo[s] = pi;
      ahi[s] = pi == 0 ? M / 2 : M - pi;
      wlo[s] = tabs[H + L * (M + 1) + alo[s]];
      whi[s] = tabs[H + L * (M + 1) + ahi[s]];
    }
  }
  // window rows of this lane's MFMA units; the twiddle index ky * y mod H of
  // the lane's first k-slot, carried from k-step to k-step (y advances by 16)
  // With H a power of two the indices run unreduced in bytes and are masked at
  // use, the part's table selected by a disjoint bit: (t & (8H - 8)) | 8H.
  const bool hp2 = (H & (H - 1)) == 0;
  int ky[NU], mrun[NU], ky16[NU];
#pragma unroll
  for (int i = 0; i < NU; ++i) {
    const int jt = (wv + 4 * i) >> 1;
    int t = (16 * jt + lr + g.SY[p] + H / 2) % H;
    ky[i] = t < 0 ? t + H : t;
    if (hp2) {   // in bytes of the 4-byte table words
      mrun[i] = 16 * ky[i] * q;
      ky16[i] = 64 * ky[i];
      ky[i] *= 4;
    } else {
      mrun[i] = (ky[i] * 4 * q) % H;
      ky16[i] = (ky[i] * 16) % H;
    }
  }

  floatx4 acc[NU][U];
#pragma unroll
  for (int i = 0; i < NU; ++i)
#pragma unroll
    for (int u = 0; u < U; ++u) acc[i][u] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int rho = wave * RPW + r;             // this lane's row within a round
  auto unpack = [](float2 zk, float2 zp, float2 wk) {
    const float2 ev = make_float2(zk.x + zp.x, zk.y - zp.y);
    const float2 od = make_float2(zk.y + zp.y, zp.x - zk.x);     // -i (zk - conj zp)
    return cadd(ev, cmul(wk, od));
  };
  const int nrd = H / RR;

  // Row transforms of round rd (data in pre[]) into B buffer Bs; fetches the
  // next round's data into pre[].
  auto rows = [&](int rd, _Float16* Bs) {
    auto put = [&](int k, float2 X) {
      const unsigned c0k = (unsigned)(k - c0);
      const int c = c0k < (unsigned)w ? (int)c0k : w;
      unsigned hi, lo;
      split_f16x2(X.x, X.y, hi, lo);
      *reinterpret_cast<unsigned*>(Bs + c * PK + 2 * rho) = hi;
      *reinterpret_cast<unsigned*>(Bs + (w + 1 + c) * PK + 2 * rho) = lo;
    };
    float2 v[M];
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = make_float2((float)(pre[m] & 0xffffu), (float)(pre[m] >> 16));
    if (rd + 1 < nrd) {
      const unsigned* s2 = src + (size_t)(rd + 1) * RR * fw2;
#pragma unroll
      for (int m = 0; m < M; ++m) pre[m] = __builtin_nontemporal_load(s2 + L * m);
    }
    // pass A and the inter-pass twiddle w_NR^(l a)
    dft_any<M, false>(v);
    v[0] = cscale(v[0], sc);
#pragma unroll
    for (int aa = 1; aa < M; ++aa) v[aa] = cmul(v[aa], twA[l * (M + 1) + aa]);
    // transpose through the wave-private rows, low then high half of a
    float2 zl[S][L], zh[S][L];
#pragma unroll
    for (int aa = 0; aa < M / 2; aa += 2)
      *reinterpret_cast<float4*>(E + l * LP + aa) = make_float4(v[aa].x, v[aa].y, v[aa + 1].x, v[aa + 1].y);
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ls = 0; ls < L; ++ls) zl[s][ls] = E[ls * LP + alo[s]];
    wave_lds_sync();
#pragma unroll
    for (int aa = 0; aa < M / 2; aa += 2)
      *reinterpret_cast<float4*>(E + l * LP + aa) =
          make_float4(v[M / 2 + aa].x, v[M / 2 + aa].y, v[M / 2 + aa + 1].x, v[M / 2 + aa + 1].y);
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ls = 0; ls < L; ++ls) zh[s][ls] = E[ls * LP + ahi[s] - M / 2];
    wave_lds_sync();
    // pass B: zl[s][b] = Z[alo + M b], zh[s][b] = Z[ahi + M b]
#pragma unroll
    for (int s = 0; s < S; ++s) {
      dft_any<L, false>(zl[s]);
      dft_any<L, false>(zh[s]);
    }
    // unpack at the window columns into the B operand (fp16 hi / lo)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const bool selfpair = (l * S + s) == 0;   // a = 0 and a = M/2 pair within their own group
      sfor<0, L>([&](auto bc) {
        constexpr int bb = decltype(bc)::value;
        // bins M bb .. M bb + M - 1; a block wholly outside the window is skipped
        // (c0 is uniform, so this is a scalar branch)
        if (M * bb >= c0 + w || M * bb + M <= c0) return;
        // partners as values first: a select between two array elements would
        // become a dynamically indexed (scratch) load
        const float2 zlr = zl[s][(L - bb) % L], zhr = zh[s][L - 1 - bb], zlm = zl[s][L - 1 - bb];
        const float2 pl = make_float2(selfpair ? zlr.x : zhr.x, selfpair ? zlr.y : zhr.y);
        const float2 ph = make_float2(selfpair ? zhr.x : zlm.x, selfpair ? zhr.y : zlm.y);
        put(alo[s] + M * bb, unpack(zl[s][bb], pl, tw_mul<bb, 2 * L, false>(wlo[s])));
        put(ahi[s] + M * bb, unpack(zh[s][bb], ph, tw_mul<bb, 2 * L, false>(whi[s])));
      });
      // Nyquist column k = NR from Z[0] (other lanes discard)
      if (NR >= c0 && NR < c0 + w) {
        const float2 z0 = zl[s][0];
        put(selfpair ? NR : c0 - 1, make_float2(2.f * (z0.x - z0.y), 0.f));
      }
    }
  };

  // F += D . X over the rows of round rd (B buffer Bs); rounds run in order,
  // so mrun[i] is ky * (rd RR + 16 ks + 4 q) mod H on entry to each k-step
  const unsigned* twU = twT + ((wv & 1) ? 2 * H : 0);   // every unit of a wave has the wave's part
  // power-of-two H: LDS byte address of the hi word (t & 4(H - 1)) + the part's
  // table base, the lo word 4H further on
  using lds_u32 = __attribute__((address_space(3))) const unsigned;
  const unsigned m4 = 4u * (unsigned)(H - 1), lo4 = 4u * (unsigned)H;
  // the part's table: a bit above every masked index (8H), OR-ed in; the
  // tables start at LDS address 0 (ldsBase folds to that constant)
  const unsigned partB = (wv & 1) ? 8u * (unsigned)H : 0u;
  const unsigned ldsBase = (unsigned)(size_t)((__attribute__((address_space(3))) unsigned char*)w3_smem);
  const int nUnits = (2 * J - wv + 3) / 4;   // units wv, wv + 4, ... below 2 J: NU or NU - 1
  // the wave's unit count NI is a compile-time argument, so a round holds no
  // branch on it.  mma_units: k-step ks of the product for A operands ah / al
  // of the wave's units.
  auto mma_units = [&](auto ni, const half8* ah, const half8* al, const _Float16* Bs, int ks) {
    constexpr int NI = decltype(ni)::value;
    // one product (hi hi, hi lo, lo hi) at a time over the units and a pair of
    // n-tiles, so consecutive MFMAs write different accumulators
    constexpr int UC = 2;
#pragma unroll
    for (int u0 = 0; u0 < U; u0 += UC) {
      half8 bh[UC], bl[UC];
#pragma unroll
      for (int uu = 0; uu < UC; ++uu) {
        if (u0 + uu < U) {
          // MR: column c at position (c mod 16) U + c / 16, so the pair stage's
          // columns (16 apart) are written to consecutive positions
          const int bcol = Sh::MR ? lr * U + u0 + uu : 16 * (u0 + uu) + lr;
          const _Float16* bq = Bs + bcol * PK + 32 * ks + 8 * q;
          bh[uu] = *reinterpret_cast<const half8*>(bq);
          bl[uu] = *reinterpret_cast<const half8*>(bq + (w + 1) * PK);
        }
      }
#pragma unroll
      for (int uu = 0; uu < UC; ++uu)
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (u0 + uu < U)
            acc[i][u0 + uu] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[uu], acc[i][u0 + uu], 0, 0, 0);
#pragma unroll
      for (int uu = 0; uu < UC; ++uu)
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (u0 + uu < U)
            acc[i][u0 + uu] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[uu], acc[i][u0 + uu], 0, 0, 0);
#pragma unroll
      for (int uu = 0; uu < UC; ++uu)
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (u0 + uu < U)
            acc[i][u0 + uu] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[uu], acc[i][u0 + uu], 0, 0, 0);
    }
  };
  // A operands from the LDS table
  auto gemm_n = [&](auto ni, const _Float16* Bs) {
    constexpr int NI = decltype(ni)::value;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      half8 ah[NI > 0 ? NI : 1], al[NI > 0 ? NI : 1];
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        {
          uint4 dh, dl;   // k-slot pairs (Re, Im) of A: part 0 (cos, sin), part 1 (-sin, cos)
          unsigned* ph = reinterpret_cast<unsigned*>(&dh);
          unsigned* pl = reinterpret_cast<unsigned*>(&dl);
          if (hp2) {
            unsigned tb = (unsigned)mrun[i];
            mrun[i] += ky16[i];
#pragma unroll
            for (int e2 = 0; e2 < 4; ++e2) {
              const unsigned ad = ((tb & m4) | partB) + ldsBase;
              ph[e2] = *reinterpret_cast<lds_u32*>((size_t)ad);
              pl[e2] = *reinterpret_cast<lds_u32*>((size_t)(ad + lo4));
              tb += (unsigned)ky[i];
            }
          } else {
            int mm = mrun[i];
            mrun[i] += ky16[i];
            if (mrun[i] >= H) mrun[i] -= H;
#pragma unroll
            for (int e2 = 0; e2 < 4; ++e2) {
              ph[e2] = twU[mm];
              pl[e2] = twU[H + mm];
              mm += ky[i];
              if (mm >= H) mm -= H;
            }
          }
          ah[i] = __builtin_bit_cast(half8, dh);
          al[i] = __builtin_bit_cast(half8, dl);
        }
      }
      mma_units(ni, ah, al, Bs, ks);
    }
  };
  auto gemm = [&](int rd, const _Float16* Bs) {
    (void)rd;
    if (nUnits == NU) gemm_n(std::integral_constant<int, NU>{}, Bs);
    else gemm_n(std::integral_constant<int, NU - 1>{}, Bs);
  };

  if constexpr (Sh::MR) {
    // Row transforms on MFMA.  Pass A: Y[l][a] = sum_m z[l + 8m] w16^(a m) for
    // the 16-row tiles (l, round), the frame samples offset by -2048 so that
    // 12-bit values are exact in fp16 (a wave whose words hold larger values
    // adds the exact fp16 remainder as a third product).  Pass A leaves its
    // output in the exchange, grouped by residue pair; pass B multiplies each
    // pair's 16 Y values of a row by the pair's matrix (8-point DFT, twiddles,
    // r2c unpack), which yields the row spectrum at 16 columns, straight in
    // the B buffer's layout (lane (o, q) holds rows 4q .. 4q + 3 of column o).
    // Round i's exchange and B buffer are buffer i & 1: one barrier per round.
    const _Float16* mrf = reinterpret_cast<const _Float16*>(tabs + w3_mr_off(H, L, M));
    const bool nyq = c0 + w > NR;   // the window holds column NR (then not column 0)
    half8 bA[2][2], bP[2][2][2];    // fragments: 16-point DFT, the wave's pairs wave and wave + 4
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int hl = 0; hl < 2; ++hl) bA[nt][hl] = *reinterpret_cast<const half8*>(mrf + (nt * 2 + hl) * 512 + lane * 8);
    int kc[2], kpos[2];   // window column of this lane's pass-B output, its B buffer position
    float dca[2];
    // the offset taken off every sample leaves column 0 short by 2048 W
    const float dcfix = 2048.f * (float)(2 * NR) * a.xscale;
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
      const int v = (wave + 4 * pp == 0 && nyq) ? 8 : wave + 4 * pp;
      const _Float16* m = mrf + (1 + v) * kMrMatHalfs + lane * 8;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) bP[pp][nt][hl] = *reinterpret_cast<const half8*>(m + (nt * 2 + hl) * 512);
      const int k = kmr(v, lr);
      kc[pp] = k - c0;
      kpos[pp] = (unsigned)kc[pp] < (unsigned)w ? (kc[pp] & 15) * U + (kc[pp] >> 4) : w;   // w: not written
      dca[pp] = k == 0 ? dcfix : 0.f;
    }
    // this lane's pass-A column a = lr in the exchange: pair and position in it
    const int pa = (lr == 0 || lr == 8) ? 0 : (lr < 8 ? lr : 16 - lr);
    const int wa = lr >= 8 ? 1 : 0;
    constexpr int EBUF = RR * Sh::ERP;
    const floatx4 zero4 = {0.f, 0.f, 0.f, 0.f};

    // pass A of round rd (fr holds its words; round rd + 1's are loaded here)
    // into exchange buffer rd & 1
    struct PassA {
      bool big;
      half8 ah[2];
      float xs[2][4][2];
      floatx4 acc2[2][2];
    };
    auto passA_main = [&](int rd, PassA& st) {
      unsigned any = 0;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) any |= fr[t][e];
      // a sample >= 4096 in the wave's words: the exact fp16 remainder as a third product
      st.big = __builtin_amdgcn_ballot_w64((any & 0xF000F000u) != 0u) != 0;
      half8 (&ah)[2] = st.ah;
      float (&xs)[2][4][2] = st.xs;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        uint4 hw;
        unsigned* hp = reinterpret_cast<unsigned*>(&hw);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          xs[t][e][0] = (float)(fr[t][e] & 0xffffu) - 2048.f;
          xs[t][e][1] = (float)(fr[t][e] >> 16) - 2048.f;
          hp[e] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(xs[t][e][0], xs[t][e][1]));
        }
        ah[t] = __builtin_bit_cast(half8, hw);
      }
      load_mr(rd + 1 < nrd ? rd + 1 : nrd - 1);   // (past the end: a row already read, unused)
      floatx4 (&acc2)[2][2] = st.acc2;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc2[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[t], bA[nt][0], zero4, 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
          acc2[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[t], bA[nt][1], acc2[t][nt], 0, 0, 0);
    };
    // a wave whose samples reach 4096: the exact fp16 remainder as a third
    // product (a branch of its own after the round's other work)
    auto passA_fix = [&](PassA& st) {
      if (st.big) {
        const half8 (&ah)[2] = st.ah;
        const float (&xs)[2][4][2] = st.xs;
        floatx4 (&acc2)[2][2] = st.acc2;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          uint4 lw;
          unsigned* lp = reinterpret_cast<unsigned*>(&lw);
#pragma unroll
          for (int e = 0; e < 4; ++e) {   // the remainder: an integer below 32 in magnitude (exact)
            const fp16x2 hv = __builtin_bit_cast(fp16x2, reinterpret_cast<const unsigned*>(&ah[t])[e]);
            lp[e] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(xs[t][e][0] - (float)hv.x,
                                                                           xs[t][e][1] - (float)hv.y));
          }
          const half8 al = __builtin_bit_cast(half8, lw);
#pragma unroll
          for (int nt = 0; nt < 2; ++nt)
            acc2[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bA[nt][0], acc2[t][nt], 0, 0, 0);
        }
      }
    };
    // lane (a, q) holds Y[l][a] of tile rows 4q + e: round row 4 wave + 2t +
    // (q >> 1), l = 4 (q & 1) + e
    auto passA_store = [&](int rd, const PassA& st) {
      const floatx4 (&acc2)[2][2] = st.acc2;
      unsigned char* eb = rowBase + (rd & 1) * EBUF + (4 * wave + (q >> 1)) * Sh::ERP + pa * Sh::EPP + wa * 8 +
                          64 * (q & 1);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          *reinterpret_cast<float2*>(eb + 2 * t * Sh::ERP + 16 * e) = make_float2(acc2[t][0][e], acc2[t][1][e]);
    };
    auto passA = [&](int rd) {
      PassA st;
      passA_main(rd, st);
      passA_fix(st);
      passA_store(rd, st);
    };
    // pass B of round rd (exchange buffer rd & 1), pairs wave and wave + 4;
    // the window columns go to B buffer rd & 1 as fp16 hi / lo
    auto pairs = [&](int rd) {
      const unsigned char* eb = rowBase + (rd & 1) * EBUF + lr * Sh::ERP + 32 * q;
      _Float16* Bs = Bs0 + (rd & 1) * BSTRIDE;
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        const float4* ep = reinterpret_cast<const float4*>(eb + (wave + 4 * pp) * Sh::EPP);
        const float4 e0 = ep[0], e1 = ep[1];
        uint4 hw, lw;
        split_f16x2(e0.x, e0.y, hw.x, lw.x);
        split_f16x2(e0.z, e0.w, hw.y, lw.y);
        split_f16x2(e1.x, e1.y, hw.z, lw.z);
        split_f16x2(e1.z, e1.w, hw.w, lw.w);
        const half8 ah = __builtin_bit_cast(half8, hw), al = __builtin_bit_cast(half8, lw);
        floatx4 c[2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bP[pp][nt][0], zero4, 0, 0, 0);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bP[pp][nt][1], c[nt], 0, 0, 0);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) c[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bP[pp][nt][0], c[nt], 0, 0, 0);
        uint4 xh, xl;
        unsigned* hp = reinterpret_cast<unsigned*>(&xh);
        unsigned* lp = reinterpret_cast<unsigned*>(&xl);
#pragma unroll
        for (int e = 0; e < 4; ++e) split_f16x2(c[0][e] + dca[pp], c[1][e], hp[e], lp[e]);
        // only the window's columns are written (lanes outside it would all
        // hit the spare column, one address: a bank conflict per write)
        if (kpos[pp] < w) {
          _Float16* bp = Bs + kpos[pp] * PK + 8 * q;
          *reinterpret_cast<uint4*>(bp) = xh;
          *reinterpret_cast<uint4*>(bp + (w + 1) * PK) = xl;
        }
      }
    };

    // The column product's A operands come precomputed per window (every
    // frame of the batch uses the same ones): unit u of 16-row block yb at
    // atab + (yb 2 J + u) 128, 16-byte words hi then lo, lane-contiguous, read
    // one round ahead from L2.
    const uint4* atab = reinterpret_cast<const uint4*>(tabs + w3_atab_off(H, L, M)) +
                        (size_t)p * (H / 16) * (2 * J) * 128 + lane;
    auto load_a = [&](auto ni, int yb, uint4 (&an)[NU][2]) {
      constexpr int NI = decltype(ni)::value;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const uint4* s = atab + (size_t)(yb * 2 * J + wv + 4 * i) * 128;
        an[i][0] = s[0];
        an[i][1] = s[64];
      }
    };
    auto gemm_a = [&](auto ni, const uint4 (&an)[NU][2], const _Float16* Bs) {
      constexpr int NI = decltype(ni)::value;
      half8 ah[NI > 0 ? NI : 1], al[NI > 0 ? NI : 1];
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        ah[i] = __builtin_bit_cast(half8, an[i][0]);
        al[i] = __builtin_bit_cast(half8, an[i][1]);
      }
      mma_units(ni, ah, al, Bs, 0);
    };
    // (nrd >= 2: H is at least 32).  A round's pass B, column product and
    // pass A touch different buffers; up to pass A's rare correction they form
    // one block, so their MFMA and VALU work can interleave.
    auto rounds = [&](auto ni) {
      uint4 an[NU][2];   // A operands of the next column product
      load_a(ni, 0, an);
      __syncthreads();   // tables staged
      passA(0);
      passA(1);
      __syncthreads();
      pairs(0);
      __syncthreads();
      for (int rd = 0; rd + 2 < nrd; ++rd) {
        PassA st;
        pairs(rd + 1);
        gemm_a(ni, an, Bs0 + (rd & 1) * BSTRIDE);
        load_a(ni, rd + 1, an);
        passA_main(rd + 2, st);
        // the scheduler's small-GEMM interleave of LDS reads with MFMAs
        __builtin_amdgcn_iglp_opt(0);
        passA_fix(st);
        passA_store(rd + 2, st);
        __syncthreads();
      }
      pairs(nrd - 1);
      gemm_a(ni, an, Bs0 + (nrd & 1) * BSTRIDE);
      load_a(ni, nrd - 1, an);
      __syncthreads();
      gemm_a(ni, an, Bs0 + ((nrd - 1) & 1) * BSTRIDE);
      __syncthreads();
    };
    if (nUnits == NU) rounds(std::integral_constant<int, NU>{});
    else rounds(std::integral_constant<int, NU - 1>{});
  } else {
    // Two B buffers: round rd's rows go to buffer rd & 1, so one barrier per
    // round separates the writers of a buffer from its readers (a buffer is
    // rewritten two rounds later, after every wave has passed the barrier that
    // follows its last reads).
    __syncthreads();   // tables staged
    rows(0, Bs0);
    __syncthreads();
    for (int rd = 0; rd < nrd; ++rd) {
      if (rd + 1 < nrd) rows(rd + 1, Bs0 + ((rd + 1) & 1) * BSTRIDE);
      gemm(rd, Bs0 + (rd & 1) * BSTRIDE);
      __syncthreads();
    }
  }

  // The epilogue's first row pass needs the window factor and its first
  // column pass the reference wave (both from L2): their loads are issued
  // here, so their latency passes while the plane is written.
  const int t16 = lane >> 2, l4 = lane & 3;   // transform within a wave pass, lane within the transform
  constexpr int MMx = w / 4, MMy = h / 4, USy = MMy / 4;
  const char* fac = reinterpret_cast<const char*>(factor + (size_t)(lam_in(g, b) * g.P + p) * (h * w));
  const char* rp = reinterpret_cast<const char*>(ref + (size_t)(lam_in(g, b) * g.P + p) * (h * w));
  auto load_factor = [&](int ps, float2 (&f)[MMx]) {
    const unsigned fo = (unsigned)((16 * ps + t16) * w + l4) * 8u;
#pragma unroll
    for (int m = 0; m < MMx; ++m) f[m] = *reinterpret_cast<const float2*>(fac + fo + 32u * m);
  };
  // column passes deal lanes l4-major (lane = 16 l4 + t16): a 64-bit access to
  // one column's rows l4 + 4 m then touches each LDS bank at most twice per
  // half wave on the plane's pitch (t16-major: four times)
  const int tc = lane & 15, lc = lane >> 4;
  auto load_ref = [&](int ps, float2 (&r)[USy][4]) {
    const unsigned ro = (unsigned)(lc * USy * w + 16 * ps + tc) * 8u;
#pragma unroll
    for (int s2 = 0; s2 < USy; ++s2)
#pragma unroll
      for (int bb = 0; bb < 4; ++bb)
        r[s2][bb] = *reinterpret_cast<const float2*>(rp + ro + (unsigned)((s2 + MMy * bb) * w) * 8u);
  };
  // the epilogue's passes and overlap units rotate with the frame as well (a
  // wave with a second pass sits on a different SIMD for every other frame)
  const int we = (wv + 2) & 3;
  float2 fcur[MMx], rcur[USy][4];
  if (we < J) load_factor(we, fcur);
  if (we < U) load_ref(we, rcur);

  // window plane: F / xscale
  const float inv = 1.f / a.xscale;
  float* fl = reinterpret_cast<float*>(fbuf);
#pragma unroll
  for (int i = 0; i < NU; ++i) {
    const int unit = wv + 4 * i;
    if (unit < 2 * J) {
      const int jt = unit >> 1, part = unit & 1;
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) fl[2 * ((16 * jt + 4 * q + e) * WP + 16 * u + lr) + part] = acc[i][u][e] * inv;
    }
  }
  __syncthreads();
  // inverse transforms along x (rows), window factor on load; a wave with a
  // second pass loads its factor rows at the start of the first
  {
    constexpr int MM = MMx, US = MM / 4;
    const bool hasCal = a.cal.p != nullptr;
    const float2 cb = a.cal.at(p, b);
    for (int ps = we; ps < J; ps += 4) {
      const int j = 16 * ps + t16;
      float2* row = fbuf + j * WP;
      float2 fnx[MM];
      const bool more = ps + 4 < J;
      if (more) load_factor(ps + 4, fnx);
      float2 v[MM];
#pragma unroll
      for (int m = 0; m < MM; ++m) v[m] = cmul(row[l4 + 4 * m], fcur[m]);
      if (hasCal) {
#pragma unroll
        for (int m = 0; m < MM; ++m) v[m] = cmul(v[m], cb);
      }
      dft_any<MM, true>(v);
#pragma unroll
      for (int aa = 1; aa < MM; ++aa) {
        const float2 tw = twix[l4 * aa];
        v[aa] = cmul(v[aa], make_float2(tw.x, -tw.y));
      }
      wave_lds_sync();
#pragma unroll
      for (int aa = 0; aa < MM; aa += 2)
        *reinterpret_cast<float4*>(row + l4 * MM + aa) = make_float4(v[aa].x, v[aa].y, v[aa + 1].x, v[aa + 1].y);
      wave_lds_sync();
      float2 z[US][4];
#pragma unroll
      for (int s2 = 0; s2 < US; ++s2)
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) z[s2][ll] = row[ll * MM + l4 * US + s2];
      wave_lds_sync();
#pragma unroll
      for (int s2 = 0; s2 < US; ++s2) {
        dftR<4, true>(z[s2]);
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) row[l4 * US + s2 + MM * bb] = z[s2][bb];
      }
      if (more) {
#pragma unroll
        for (int m = 0; m < MM; ++m) fcur[m] = fnx[m];
      }
    }
  }
  __syncthreads();
  // inverse transforms along y (columns), tilt on the way out
  const bool ovl = a.hgx16 != nullptr || a.hgx32 != nullptr;
  float vmax = 0.f;
  float* red = reinterpret_cast<float*>(twiy + h);   // [4] per-wave peaks
  {
    constexpr int MM = MMy, US = USy;
    // reference wave (loaded a pass ahead) and fields through uniform bases
    // and 32-bit lane offsets
    char* out = reinterpret_cast<char*>(fields + (size_t)(out_pos(g, b) * g.P + p) * (h * w));
    for (int ps = we; ps < U; ps += 4) {
      const int c = 16 * ps + tc;
      float2* col = fbuf + c;
      const unsigned ro = (unsigned)(lc * US * w + c) * 8u;   // element (k, c) at ro + (k - lc US) w 8
      float2 rnx[US][4];
      const bool more = ps + 4 < U;
      if (more) load_ref(ps + 4, rnx);
      float2 v[MM];
#pragma unroll
      for (int m = 0; m < MM; ++m) v[m] = col[(lc + 4 * m) * WP];
      dft_any<MM, true>(v);
#pragma unroll
      for (int aa = 1; aa < MM; ++aa) {
        const float2 tw = twiy[lc * aa];
        v[aa] = cmul(v[aa], make_float2(tw.x, -tw.y));
      }
      wave_lds_sync();
#pragma unroll
      for (int aa = 0; aa < MM; ++aa) col[(lc * MM + aa) * WP] = v[aa];
      wave_lds_sync();
      float2 z[US][4];
#pragma unroll
      for (int s2 = 0; s2 < US; ++s2)
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) z[s2][ll] = col[(ll * MM + lc * US + s2) * WP];
      wave_lds_sync();
#pragma unroll
      for (int s2 = 0; s2 < US; ++s2) {
        dftR<4, true>(z[s2]);
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          const int k = lc * US + s2 + MM * bb;
          const float2 tt = cmul(z[s2][bb], rcur[s2][bb]);
          const float2 V = make_float2(-tt.x, tt.y);
          *reinterpret_cast<float2*>(out + ro + (unsigned)((s2 + MM * bb) * w) * 8u) = V;
          if (ovl) {   // the field stays in the plane for the overlap
            col[k * WP] = V;
            vmax = fmaxf(vmax, fmaxf(fabsf(V.x), fabsf(V.y)));
          }
        }
      }
      if (more) {
#pragma unroll
        for (int s2 = 0; s2 < US; ++s2)
#pragma unroll
          for (int bb = 0; bb < 4; ++bb) rcur[s2][bb] = rnx[s2][bb];
      }
    }
  }
  if (a.hgx16) w3_overlap<U, J>(a, fbuf, vmax, red, p, out_pos(g, b), we, lane);
  else if (a.hgx32) {
    __syncthreads();   // the whole plane is written before any wave reads it
    w3_overlap32<U, J>(a, fbuf, w3_smem + a.ovlOff, p, out_pos(g, b), we, lane);
  }
}

template <int L, int M, int U, int J, bool SQ, bool R>
__global__ __launch_bounds__(kW3Threads, 2) void k_window_v3(W3Args a, const unsigned short* __restrict__ frames,
                                                          const float2* __restrict__ tabs,
                                                          const float2* __restrict__ factor,
                                                          const float2* __restrict__ ref,
                                                          float2* __restrict__ fields) {
  w3_body<L, M, U, J, SQ, R>(a, frames, tabs, factor, ref, fields, blockIdx.x);
}

// One handle's batch in a multi-batch launch: its arguments, its kernel
// shape (index into W3_SHAPES) and its P B work items.
struct W3Item {
  W3Args a;
  const unsigned short* frames;
  const float2* tabs;
  const float2* factor;
  const float2* ref;
  float2* fields;
  int shape;
  int items;
};
struct W3Multi {
  W3Item it[kW3MaxBatches];
  int n;       // batches
  int total;   // work items of all batches
};

// Tables of the v3 kernel, from FP64: [H] cos and sin of 2 pi m / H as fp16
// hi/lo pairs {(cos_hi, sin_hi), (cos_lo, sin_lo)}; [L][M + 1]
// exp(-2 pi i l a / NR); [M] exp(-2 pi i a / W).
// Column-product A operands of the MFMA row path, per window: unit (jt, part)
// of 16-row block yb, lane (lr, q) holding window row ky = (16 jt + lr + SY +
// H / 2) mod H against rows y = 16 yb + 4 q + e (e = 0..3): the staged table's
// fp16 hi / lo words of w_H^(ky y), part 1 as (-sin, cos).  [P][H / 16][2 J]
// [hi, lo][64] uint4.  Reads the first H entries of tab (k_w3_tables).
__global__ void k_w3_atab(Geometry g, const float2* __restrict__ tab, uint4* __restrict__ at) {
  const int H = g.H, J = g.h / 16, nyb = H / 16;
  const int n = g.P * nyb * 2 * J * 64;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int lane = e & 63;
  int r = e >> 6;
  const int unit = r % (2 * J);
  r /= 2 * J;
  const int yb = r % nyb, p = r / nyb;
  const int lr = lane & 15, q = lane >> 4, jt = unit >> 1, part = unit & 1;
  int ky = (16 * jt + lr + g.SY[p] + H / 2) % H;
  if (ky < 0) ky += H;
  const uint2* t2 = reinterpret_cast<const uint2*>(tab);
  unsigned hw[4], lw[4];
#pragma unroll
  for (int e2 = 0; e2 < 4; ++e2) {
    const int y = 16 * yb + 4 * q + e2;
    const uint2 t = t2[(int)(((long long)ky * y) % H)];
    hw[e2] = part ? (__builtin_amdgcn_alignbit(t.x, t.x, 16) ^ 0x8000u) : t.x;
    lw[e2] = part ? (__builtin_amdgcn_alignbit(t.y, t.y, 16) ^ 0x8000u) : t.y;
  }
  uint4* dst = at + (size_t)((p * nyb + yb) * 2 * J + unit) * 128;
  dst[lane] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
  dst[64 + lane] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

__global__ void k_w3_tables(int H, int L, int M, float2* __restrict__ tab) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int NR = L * M;
  double ang;   // in units of pi
  if (e < H) {
    double sn, cs;
    sincospi(2.0 * e / H, &sn, &cs);
    const _Float16 ch = (_Float16)cs, sh = (_Float16)sn;
    const _Float16 cl = (_Float16)(cs - (double)ch), sl = (_Float16)(sn - (double)sh);
    const half2v hi = {ch, sh}, lo = {cl, sl};
    reinterpret_cast<uint2*>(tab)[e] = make_uint2(__builtin_bit_cast(unsigned, hi), __builtin_bit_cast(unsigned, lo));
    return;
  } else if ((e -= H) < L * (M + 1)) {
    const int ll = e / (M + 1), aa = e - ll * (M + 1);
    ang = -2.0 * (double)((ll * aa) % NR) / NR;
  } else if ((e -= L * (M + 1)) < M) {
    ang = -1.0 * e / NR;
  } else {
    return;
  }
  double sn, cs;
  sincospi(ang, &sn, &cs);
  tab[blockIdx.x * blockDim.x + threadIdx.x] = make_float2((float)cs, (float)sn);
}

// MFMA fragments of the MR row transform (128-point rows), from FP64, as fp16
// hi/lo pairs laid out [matrix][n-tile][hi, lo][lane][8 k-slots]: lane
// (n, q) of a 16x16x32 B operand holds column n, k-slots 8q .. 8q + 7.
//
// Matrix 0, the 16-point DFT of pass A (times 2^-5): k-slot 8q + j is part
// j & 1 (re, im) of z[l + 8 m], m = 4q + (j >> 1); column a of n-tile 0 / 1 is
// Re / Im of Y[l][a] = sum_m z[l + 8 m] w16^(a m).
//
// Matrices 1 + v, pass B of residue pair v (v = 0..7; 8: pair 0 with the
// Nyquist column in place of column 0): k-slot 8q + j is part j & 1 of
// Y[l][a], complex index c = 4q + (j >> 1), l = c >> 1, a the pair's residue
// c & 1 (pair v >= 1: v, 16 - v; pair 0: 0, 8).  Column o of n-tile 0 / 1 is
// Re / Im of X[k], k = kmr(v, o), with Z[k] = sum_l w128^(l k) Y[l][k mod 16]
// and X[k] = (Z[k] + conj Z[128 - k]) / 2 - i w256^k (Z[k] - conj Z[128 - k]) / 2,
// times xscale / 2^-5: the r2c unpack of the 256 real samples packed into
// z[n] = f[2n] + i f[2n + 1].  (w_N = exp(-2 pi i / N))
__global__ void k_w3_mr_tables(double pairScale, _Float16* __restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kMrMats * 2 * 64 * 8) return;
  const int j = t & 7, lane = (t >> 3) & 63, nt = (t >> 9) & 1, mat = t >> 10;
  const int q = lane >> 4, n = lane & 15, pin = j & 1;
  double v;
  if (mat == 0) {
    const int m = 4 * q + (j >> 1);
    double s, c;
    sincospi(2.0 * ((n * m) % 16) / 16.0, &s, &c);
    v = nt == 0 ? (pin ? s : c) : (pin ? c : -s);
    v *= kMrScaleA;
  } else {
    const int pv = mat - 1, p = pv == 8 ? 0 : pv;
    const int ci = 4 * q + (j >> 1), l = ci >> 1, which = ci & 1;
    const int ain = p == 0 ? (which ? 8 : 0) : (which ? 16 - p : p);
    const int k = kmr(pv, n), kp = (128 - k) & 127;
    // dZ = w128^(l k) Y for the unit Y = 1 (re) or i (im)
    auto dz = [&](int kk, double& re, double& im) {
      re = im = 0.0;
      if ((kk & 15) != ain) return;
      double s, c;
      sincospi(-2.0 * ((l * kk) % 128) / 128.0, &s, &c);   // w128^(l kk) = c + i s
      if (pin == 0) { re = c; im = s; }
      else { re = -s; im = c; }
    };
    double zr, zi, pr, pi;
    dz(k, zr, zi);
    dz(kp, pr, pi);
    pi = -pi;                                   // conj Z[128 - k]
    const double er = 0.5 * (zr + pr), ei = 0.5 * (zi + pi);
    const double dr = 0.5 * (zr - pr), di = 0.5 * (zi - pi);
    // -i w256^k (d): w256^k = c - i s
    double s, c;
    sincospi(2.0 * (k % 256) / 256.0, &s, &c);
    const double wr = c, wi = -s;
    const double mr = di, mi = -dr;            // -i d
    const double xr = er + wr * mr - wi * mi, xi = ei + wr * mi + wi * mr;
    v = (nt == 0 ? xr : xi) * pairScale;
  }
  const _Float16 hi = (_Float16)v;
  const _Float16 lo = (_Float16)(v - (double)hi);
  _Float16* m = dst + (size_t)mat * kMrMatHalfs + (size_t)nt * (2 * 64 * 8);
  m[lane * 8 + j] = hi;
  m[64 * 8 + lane * 8 + j] = lo;
}

// (L, M) of a row transform of nr complex points; false if not specialised
static bool w3_lm(int nr, int& L, int& M) {
  if (nr == 128) { L = 8; M = 16; return true; }
  if (nr == 64) { L = 4; M = 16; return true; }
  if (nr == 96) { L = 4; M = 24; return true; }
  return false;
}

// the power of two keeping the row spectra of uint16 frames inside fp16:
// |X| <= W 65535 < 2^(16 + log2 W), scaled into [2^14, 2^15) at most
static float w3_xscale(int W) {
  int lw = 0;
  while ((1 << lw) < W) ++lw;
  return ldexpf(1.f, 15 - (16 + lw));
}

size_t window_v3_table_elems(const Geometry& g) {
  int L, M;
  if (!w3_lm(g.W / 2, L, M)) return 0;
  if (L == 8 && M == 16) return (size_t)w3_atab_off(g.H, L, M) + w3_atab_elems(g.P, g.H, g.h / 16);
  return (size_t)g.H + (size_t)L * (M + 1) + M;
}

// Staged row transform of 256- and 512-wide FFT windows (N = 16 T complex
// points of packed real pairs, digHolo.cpp:1855-1872 for the r2c layout), in
// registers: T lanes per row, each taking points t + T m (m = 0..15) straight
// from the frame, a 16-point DFT over m, twiddles w_N^(t k1), a transpose
// through LDS within the row's lanes (a wave holds 64 / T rows: no block
// barrier), T-point DFTs over t for k1 = t, t + T, .. giving Z[k1 + 16 k2],
// then the r2c unpack of the w window columns c0 .. c0 + w - 1 from Z[k] and
// Z[N - k].  grid (ceil(H / (256 / T)), P, frames), 256 threads.
template <int T>
__global__ __launch_bounds__(256) void k_row_fft_reg(Geometry g, const float2* __restrict__ tw,
                                                     const float2* __restrict__ twW, const void* __restrict__ frames,
                                                     int u16, float2* __restrict__ rowspec) {
  constexpr int N = 16 * T, ROWS = 256 / T;
  constexpr int RP = 16 * (T + 1) + 1;   // a row's LDS: [16 k1][T + 1], then Z[k] at 17 (k / 16) + k % 16
  __shared__ float2 stw[N];
  __shared__ float2 ex[ROWS * RP];
  const int b = blockIdx.z + g.b0, p = blockIdx.y;
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + r;
  for (int e = threadIdx.x; e < N; e += 256) stw[e] = tw[e];
  const bool live = y < g.H;
  const size_t rowOff = (size_t)b * g.frameH * g.frameW + (size_t)(g.cy[p] + (live ? y : 0)) * g.frameW + g.cx[p];
  float2 z[16];
  if (u16) {
    const unsigned short* f = static_cast<const unsigned short*>(frames) + rowOff;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const ushort2 v = *reinterpret_cast<const ushort2*>(f + 2 * (t + T * m));
      z[m] = make_float2((float)v.x, (float)v.y);
    }
  } else {
    const float* f = static_cast<const float*>(frames) + rowOff;
#pragma unroll
    for (int m = 0; m < 16; ++m) z[m] = *reinterpret_cast<const float2*>(f + 2 * (t + T * m));
  }
  dft_any<16, false>(z);
  __syncthreads();   // twiddles staged
  float2* xr = ex + r * RP;
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) xr[(T + 1) * k1 + t] = k1 ? cmul(z[k1], stw[t * k1]) : z[0];
  wave_lds_sync();
  constexpr int NK = 16 / T;   // k1 values per lane
  float2 u[NK][T];
#pragma unroll
  for (int i = 0; i < NK; ++i) {
#pragma unroll
    for (int tt = 0; tt < T; ++tt) u[i][tt] = xr[(T + 1) * (t + T * i) + tt];
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < NK; ++i) {
    dft_any<T, false>(u[i]);   // u[i][k2] = Z[t + T i + 16 k2]
#pragma unroll
    for (int k2 = 0; k2 < T; ++k2) xr[17 * k2 + t + T * i] = u[i][k2];
  }
  wave_lds_sync();
  if (!live) return;
  auto Zat = [&](int k) { return xr[17 * (k >> 4) + (k & 15)]; };
  const int c0 = g.c0[p], w = g.w;
  float2* out = rowspec + ((size_t)(b * g.P + p) * g.H + y) * w;
  for (int kk = t; kk < w; kk += T) {
    const int k = c0 + kk;   // k <= N (the Nyquist column)
    const float2 zk = Zat(k == N ? 0 : k);
    float2 zc = Zat((k == 0 || k == N) ? 0 : N - k);
    zc.y = -zc.y;
    const float2 ev = cscale(cadd(zk, zc), 0.5f);
    const float2 od = mul_mi<false>(cscale(csub(zk, zc), 0.5f));   // (zk - zc)/(2i)
    out[kk] = cadd(ev, cmul(twW[c0 + kk], od));
  }
}

hipError_t launch_row_fft_reg(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                              int frame_is_u16, float2* rowspec, hipStream_t s) {
  if (pl.n != 128 && pl.n != 256) return hipErrorNotSupported;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    if (pl.n == 256)
      k_row_fft_reg<16><<<dim3((g.H + 15) / 16, g.P, nb), 256, 0, s>>>(gc, pl.tw, twW, frames, frame_is_u16, rowspec);
    else
      k_row_fft_reg<8><<<dim3((g.H + 31) / 32, g.P, nb), 256, 0, s>>>(gc, pl.tw, twW, frames, frame_is_u16, rowspec);
    return hipGetLastError();
  });
}

// Staged column stage of 256- and 512-row FFT windows, the column transform
// in registers: 16 lanes per column (lane c + 16 t), each taking band rows
// t + 16 m (m < H / 16), an H/16-point DFT over m, the twiddles w_H^(t k1), a
// transpose through LDS in halves of 16 k1, and 16-point DFTs over t giving
// Z[k1 + (H / 16) k2] for k1 = t, t + 16, ...  Only the h
// window rows go to the plane (window row j is spectrum row (j + SY + H/2) mod
// H), which takes the exchange's LDS; then as k_col_window_iy: times the
// window factor (loaded while the transform runs), the inverse transform along
// y, the window plane out.  grid (ceil(w / 16), P, frames), 256 threads,
// h <= 256; about 40 KB of LDS.
constexpr int kC512Cols = 16, kC512Sel = 16;
constexpr int kC512XS = 16 * 17 + 1;          // exchange floats2 per column: [16 k1][17]
template <int H>
__global__ __launch_bounds__(256) void k_col_reg_window_iy(Geometry g, FFTPlan py, const float2* __restrict__ twc_g,
                                                          const float2* __restrict__ rowspec,
                                                          const float2* __restrict__ factor,
                                                          float2* __restrict__ win, int chunkY, BatchCal cal,
                                                          unsigned* __restrict__ peak) {
  constexpr int NM = H / 16, NH = H / 256;   // points per lane, halves of the exchange
  extern __shared__ float2 smem[];
  const int h = g.h, w = g.w, LD = pad_stride(h);
  float2* twc = smem;       // [H] column-transform twiddles
  float2* twy = smem + H;   // [h] inverse twiddles
  float2* A = twy + h;      // the exchange, then window row j of column c at c LD + pad32(j)
  const int b = blockIdx.z + g.b0, p = blockIdx.y, i0 = blockIdx.x * kC512Cols;
  const int cols = min(kC512Cols, w - i0);
  const int c = threadIdx.x & 15, t = threadIdx.x >> 4;
  if (peak && blockIdx.x == 0 && threadIdx.x == 0) peak[out_pos(g, b) * g.P + p] = 0u;
  for (int e = threadIdx.x; e < H; e += 256) twc[e] = twc_g[e];
  for (int e = threadIdx.x; e < h; e += 256) twy[e] = py.tw[e];
  const float2* in = rowspec + (size_t)(b * g.P + p) * H * w + i0;
  float2 z[NM];
  if (c < cols) {
#pragma unroll
    for (int m = 0; m < NM; ++m) z[m] = in[(size_t)(t + 16 * m) * w + c];
  } else {
#pragma unroll
    for (int m = 0; m < NM; ++m) z[m] = make_float2(0.f, 0.f);
  }
  // window row j is spectrum row (j + SY + H/2) mod H
  int sy = (g.SY[p] + H / 2) % H;
  if (sy < 0) sy += H;
  const float2* fac = factor + (size_t)(lam_in(g, b) * g.P + p) * h * w + i0;
  const unsigned magC = magic_div((unsigned)cols);
  float2 sel[kC512Sel];
#pragma unroll
  for (int s = 0; s < kC512Sel; ++s) {
    const int e = threadIdx.x + s * 256;
    if (e < h * cols) {
      const int j = fdiv(e, magC), cc = e - j * cols;
      sel[s] = fac[(size_t)j * w + cc];
    }
  }
  dft_any<NM, false>(z);
  __syncthreads();   // twiddles staged
  float2* xc = A + c * kC512XS;
  float2 u[NH][16];
#pragma unroll
  for (int hh = 0; hh < NH; ++hh) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int k1 = 16 * hh + k;
      xc[17 * k + t] = k1 ? cmul(z[k1], twc[t * k1]) : z[0];
    }
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < 16; ++tt) u[hh][tt] = xc[17 * t + tt];
    __syncthreads();   // read before the next half (or the plane) is written
    dft_any<16, false>(u[hh]);   // u[hh][k2] = Z[t + 16 hh + NM k2]
  }
#pragma unroll
  for (int hh = 0; hh < NH; ++hh)
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
      int j = t + 16 * hh + NM * k2 - sy;
      if (j < 0) j += H;
      if (j < h) A[c * LD + pad32(j)] = u[hh][k2];
    }
  __syncthreads();
  const float2 cb = cal.at(p, b);
#pragma unroll
  for (int s = 0; s < kC512Sel; ++s) {   // each element read and written by one thread
    const int e = threadIdx.x + s * 256;
    if (e < h * cols) {
      const int j = fdiv(e, magC), cc = e - j * cols;
      float2* a = A + cc * LD + pad32(j);
      *a = cmul(cmul(*a, sel[s]), cb);
    }
  }
  __syncthreads();
  lds_fft_any<true, 256>(A, cols, LD, py, twy, chunkY);
  float2* out = win + (size_t)(out_pos(g, b) * g.P + p) * h * w + i0;
  for (int e = threadIdx.x; e < h * cols; e += 256) {
    const int j = fdiv(e, magC), cc = e - j * cols;
    out[(size_t)j * w + cc] = A[cc * LD + pad32(j)];
  }
}

hipError_t launch_col_reg_window_iy(const Geometry& g, const FFTPlan& pc, const FFTPlan& py, const float2* rowspec,
                                    const float2* factor, float2* win, const BatchCal& cal, hipStream_t s,
                                    unsigned* peak) {
  const int chunkY = ip_chunk(py, 256);
  if ((g.H != 256 && g.H != 512) || pc.n != g.H || py.n != g.h || g.h > 256 || chunkY < 1)
    return hipErrorNotSupported;
  const size_t lds = sizeof(float2) * ((size_t)g.H + g.h + (size_t)kC512Cols * std::max(kC512XS, pad_stride(g.h)));
  const void* kern = g.H == 512 ? reinterpret_cast<const void*>(k_col_reg_window_iy<512>)
                                : reinterpret_cast<const void*>(k_col_reg_window_iy<256>);
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    const dim3 grid((g.w + kC512Cols - 1) / kC512Cols, g.P, nb);
    if (g.H == 512)
      k_col_reg_window_iy<512><<<grid, 256, lds, s>>>(gc, py, pc.tw, rowspec, factor, win, chunkY, cal, peak);
    else
      k_col_reg_window_iy<256><<<grid, 256, lds, s>>>(gc, py, pc.tw, rowspec, factor, win, chunkY, cal, peak);
    return hipGetLastError();
  });
}

// Staged inverse transform along x of w = 16 M point rows in registers, then
// the reference wave: 16 lanes per row (four rows per wave), lane t taking
// points t + 16 m (m < M), an M-point inverse DFT over m, the twiddles
// conj(w_w^(t k1)), a transpose through LDS inside the row's lanes, and on
// lanes t < M a 16-point inverse DFT giving X[t + M k2]; then V = -conj(X ref)
// (X itself when ref is null) out, and the field's largest |Re|, |Im| into
// peak (one atomicMax of the float bits per block: one per wave contended).  grid (ceil(h / 16), P,
// frames), 256 threads.
template <int M>
__global__ __launch_bounds__(256) void k_ifft_rows_reg(Geometry g, const float2* __restrict__ tw,
                                                       const float2* __restrict__ win, const float2* __restrict__ ref,
                                                       float2* __restrict__ fields, unsigned* __restrict__ peak) {
  constexpr int N = 16 * M, RP = 17 * M;   // a row's LDS: [M][17] exchange, then X[k] at k + k / 16
  __shared__ float2 stw[N];
  __shared__ float2 ex[16 * RP];
  const int b = blockIdx.z + g.b0, p = blockIdx.y;
  const int r = threadIdx.x >> 4, t = threadIdx.x & 15;
  const int y = blockIdx.x * 16 + r;
  for (int e = threadIdx.x; e < N; e += 256) stw[e] = tw[e];
  const bool live = y < g.h;
  const size_t base = ((size_t)(b * g.P + p) * g.h + (live ? y : 0)) * N;
  float2 z[M];
#pragma unroll
  for (int m = 0; m < M; ++m) z[m] = win[base + t + 16 * m];
  dft_any<M, true>(z);
  __syncthreads();   // twiddles staged
  float2* xr = ex + r * RP;
#pragma unroll
  for (int k1 = 0; k1 < M; ++k1) xr[17 * k1 + t] = k1 ? cmul(z[k1], twiddle<true>(stw, t * k1)) : z[0];
  wave_lds_sync();
  float2 u[16];
  if (t < M) {
#pragma unroll
    for (int tt = 0; tt < 16; ++tt) u[tt] = xr[17 * t + tt];
  }
  wave_lds_sync();
  if (t < M) {
    dft_any<16, true>(u);
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
      const int k = t + M * k2;
      xr[k + (k >> 4)] = u[k2];
    }
  }
  wave_lds_sync();
  float mx = 0.f;
  if (live) {
    const float2* rp = ref ? ref + ((size_t)(lam_out(g, b) * g.P + p) * g.h + y) * N : nullptr;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const int k = t + 16 * j;
      float2 v = xr[k + (k >> 4)];
      if (rp) {
        const float2 tt = cmul(v, rp[k]);
        v = make_float2(-tt.x, tt.y);
        mx = fmaxf(mx, fmaxf(fabsf(tt.x), fabsf(tt.y)));
      }
      fields[base + k] = v;
    }
  }
  if (peak && ref) {   // one atomicMax per block; non-negative floats order as their bit patterns
    __shared__ float s_mx[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) s_mx[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
      atomicMax(peak + b * g.P + p, __float_as_uint(fmaxf(fmaxf(s_mx[0], s_mx[1]), fmaxf(s_mx[2], s_mx[3]))));
  }
}

hipError_t launch_ifft_rows_reg(const Geometry& g, const FFTPlan& pl, const float2* win, const float2* ref,
                                float2* fields, hipStream_t s, unsigned* peak) {
  if (pl.n != g.w || !ifft_rows_reg_width(g.w)) return hipErrorNotSupported;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    const dim3 grid((g.h + 15) / 16, g.P, nb);
    switch (g.w) {
      case 48: k_ifft_rows_reg<3><<<grid, 256, 0, s>>>(gc, pl.tw, win, ref, fields, peak); break;
      case 64: k_ifft_rows_reg<4><<<grid, 256, 0, s>>>(gc, pl.tw, win, ref, fields, peak); break;
      case 80: k_ifft_rows_reg<5><<<grid, 256, 0, s>>>(gc, pl.tw, win, ref, fields, peak); break;
      default: k_ifft_rows_reg<10><<<grid, 256, 0, s>>>(gc, pl.tw, win, ref, fields, peak); break;
    }
    return hipGetLastError();
  });
}

hipError_t launch_window_v3_tables(const Geometry& g, float2* tab, hipStream_t s) {
  int L, M;
  if (!w3_lm(g.W / 2, L, M)) return hipErrorNotSupported;
  const int n = g.H + L * (M + 1) + M;
  k_w3_tables<<<(n + 255) / 256, 256, 0, s>>>(g.H, L, M, tab);
  if (L == 8 && M == 16) {
    const int nf = kMrMats * 2 * 64 * 8;
    k_w3_mr_tables<<<(nf + 255) / 256, 256, 0, s>>>((double)w3_xscale(g.W) / kMrScaleA,
                                                     reinterpret_cast<_Float16*>(tab + w3_mr_off(g.H, L, M)));
    if (g.H % 16 == 0 && g.h % 16 == 0 && g.h > 0) {
      const int na = g.P * (g.H / 16) * 2 * (g.h / 16) * 64;
      k_w3_atab<<<(na + 255) / 256, 256, 0, s>>>(g, tab, reinterpret_cast<uint4*>(tab + w3_atab_off(g.H, L, M)));
    }
  }
  return hipGetLastError();
}

template <int L, int M, int U, int J>
static size_t w3_lds(const W3Args& a) {
  return (size_t)a.region + 8 * (size_t)(16 * U + 16 * J) + 16;   // + twiddles, per-wave peaks
}

// stage-2 tiles of the overlap (rows (m, part) x columns n with some m + n < G)
static int ov_tiles(int G) {
  const int Gp = (G + 15) & ~15, rt2 = (2 * Gp) / 16, ct2 = Gp / 16;
  int cnt = 0;
  for (int t = 0; t < rt2 * ct2; ++t)
    if (8 * (t / ct2) + 16 * (t % ct2) < G) ++cnt;
  return cnt;
}

bool window_v3_overlap_fits(const Geometry& g) {
  return g.G > 0 && ov_tiles(g.G) <= 4 * kOvMaxTiles && ((g.G + 15) & ~15) <= 64;
}

static int w3_ovl_bytes(const Geometry& g) {
  const int Gp = (g.G + 15) & ~15;
  return 2 * (2 * Gp * kOvPK32 * 4);   // the FP32 form's double-buffered stage-1 output
}

// LDS layout of a batch's launch (region, overlap buffers); returns the bytes
template <int L, int M, int U, int J, bool R = true>
static size_t w3_layout(W3Args& a) {
  using Sh = W3Shape<L, M, U, J, R>;
  a.region = w3_region(Sh::EBYTES, Sh::BBYTES, a.g.H, L, M, Sh::w, Sh::h);
  a.ovlOff = 8 * Sh::h * Sh::WP;
  if (a.hgx32) a.region = std::max(a.region, a.ovlOff + w3_ovl_bytes(a.g));   // the FP32 form's stage-1 buffers
  return w3_lds<L, M, U, J>(a);
}

template <int L, int M, int U, int J, bool R = true>
static hipError_t w3_launch(W3Args a, const void* frames, const float2* tab, const float2* factor,
                            const float2* ref, float2* fields, hipStream_t s) {
  using Sh = W3Shape<L, M, U, J, R>;
  const Geometry& g = a.g;
  if (g.H % Sh::RR) return hipErrorNotSupported;
  const size_t lds = w3_layout<L, M, U, J, R>(a);
  if (lds > 160 * 1024) return hipErrorNotSupported;
  auto kern = g.H == 2 * L * M ? k_window_v3<L, M, U, J, true, R> : k_window_v3<L, M, U, J, false, R>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  dim3 grid(g.P * g.B);
  kern<<<grid, kW3Threads, lds, s>>>(a, static_cast<const unsigned short*>(frames), tab, factor, ref, fields);
  return hipGetLastError();
}

size_t hg_split_elems(const Geometry& g, int n) { return (size_t)g.P * 2 * ((g.G + 15) & ~15) * n; }

hipError_t launch_hg_split(const Geometry& g, const float* hgx, const float* hgy, void* hgx16, void* hgy16,
                           hipStream_t s) {
  if (g.G <= 0) return hipSuccess;
  const int Gp = (g.G + 15) & ~15;
  const int nx = g.P * Gp * g.fw, ny = g.P * Gp * g.fh;
  k_hg_split<<<(nx + 255) / 256, 256, 0, s>>>(hgx, g.P, g.G, g.fw, static_cast<_Float16*>(hgx16));
  k_hg_split<<<(ny + 255) / 256, 256, 0, s>>>(hgy, g.P, g.G, g.fh, static_cast<_Float16*>(hgy16));
  return hipGetLastError();
}

// Calls f<L, M, U, J>() for the specialised shapes; returns false otherwise.
#define W3_SHAPES(X)                                                            \
  X(8, 16, 5, 5) X(8, 16, 4, 4) X(8, 16, 3, 3) X(8, 16, 6, 6) X(4, 16, 3, 3) \
  X(4, 16, 4, 4) X(4, 24, 4, 4)

static bool w3_common(const Geometry& g, int& L, int& M) {
  if (!w3_lm(g.W / 2, L, M)) return false;
  if (g.w % 16 || g.h % 16 || g.H > 1024) return false;
  if (L == 8 && M == 16 && g.H < 32) return false;   // the MFMA row path runs at least two 16-row rounds
  if ((g.frameW & 1) || (g.cx[0] & 1) || (g.P > 1 && (g.cx[1] & 1))) return false;   // dword row loads
  for (int p = 0; p < g.P; ++p)
    if (g.c0[p] < 0 || g.c0[p] + g.w > g.W / 2 + 1) return false;
  return true;
}

bool window_v3_supported(const FusedArgs& a) {
  const Geometry& g = a.g;
  if (g.fw != g.w || g.fh != g.h) return false;   // full-resolution fields: staged kernels
  int L, M;
  if (!w3_common(g, L, M)) return false;
  const int U = g.w / 16, J = g.h / 16;
#define W3_MATCH(LL, MM, UU, JJ)                                                             \
  if (L == LL && M == MM && U == UU && J == JJ)                                              \
    return g.H % W3Shape<LL, MM, UU, JJ>::RR == 0 &&                                         \
           w3_region(W3Shape<LL, MM, UU, JJ>::EBYTES, W3Shape<LL, MM, UU, JJ>::BBYTES, g.H, LL, MM, \
                     16 * UU, 16 * JJ) + 8 * (16 * UU + 16 * JJ) <= 160 * 1024;
  W3_SHAPES(W3_MATCH)
#undef W3_MATCH
  return false;
}

// Kernel arguments of one batch (false: the shape is not specialised).
static bool w3_args(const FusedArgs& fa, const void* hgx16, const void* hgy16, float2* coefs, const BatchCal& cal,
                    const float* hgx32, const float* hgy32, W3Args& a, int& L, int& M) {
  if (!window_v3_supported(fa)) return false;
  if ((hgx16 || hgx32) && !window_v3_overlap_fits(fa.g)) return false;
  const Geometry& g = fa.g;
  w3_lm(g.W / 2, L, M);
  a = W3Args{};
  a.g = g;
  a.ix = fa.ix;
  a.iy = fa.iy;
  a.hgx16 = static_cast<const _Float16*>(hgx16);
  a.hgy16 = static_cast<const _Float16*>(hgy16);
  a.hgx32 = hgx16 ? nullptr : hgx32;
  a.hgy32 = hgx16 ? nullptr : hgy32;
  a.coefs = coefs;
  a.cal = cal;
  a.magH = magic_div((unsigned)g.H);
  if (a.magH == 0) return false;
  // |X| <= W * 65535 < 2^(16 + log2 W): scale into [2^14, 2^15) at most
  int lw = 0;
  while ((1 << lw) < g.W) ++lw;
  a.xscale = ldexpf(1.f, 15 - (16 + lw));
  return true;
}

// rowMode 1: the register row transform where the MFMA one exists
template <int L, int M, int U, int J>
static hipError_t w3_launch_rows(int rowMode, const W3Args& a, const void* frames, const float2* tab,
                                 const float2* factor, const float2* ref, float2* fields, hipStream_t s) {
  if constexpr (L == 8 && M == 16) {
    if (rowMode == 1) return w3_launch<L, M, U, J, false>(a, frames, tab, factor, ref, fields, s);
  }
  return w3_launch<L, M, U, J, true>(a, frames, tab, factor, ref, fields, s);
}

bool window_v3_row_choice(const Geometry& g) { return g.W == 256; }

hipError_t launch_window_v3(const FusedArgs& fa, const void* frames_u16, const float2* tab, const float2* factor,
                            const float2* ref, float2* fields, const void* hgx16, const void* hgy16, float2* coefs,
                            const BatchCal& cal, hipStream_t s, const float* hgx32, const float* hgy32, int rowMode) {
  W3Args a;
  int L, M;
  if (!w3_args(fa, hgx16, hgy16, coefs, cal, hgx32, hgy32, a, L, M)) return hipErrorNotSupported;
  const Geometry& g = fa.g;
  const int U = g.w / 16, J = g.h / 16;
#define W3_CALL(LL, MM, UU, JJ) \
  if (L == LL && M == MM && U == UU && J == JJ)                                                  \
    return w3_launch_rows<LL, MM, UU, JJ>(rowMode, a, frames_u16, tab, factor, ref, fields, s);
  W3_SHAPES(W3_CALL)
#undef W3_CALL
  return hipErrorNotSupported;
}

// Several batches (different handles, shapes and sizes) in one launch:
// a one-dimensional grid holds the (pol, frame) work items of the batches one
// batch after the other, so no workgroup is idle.  Workgroups are dispatched
// in order, so the batches run back to back with no launch boundary or tail
// between them.  Square windows and the shapes of
// W3_MULTI_SHAPES (a kernel holding more of the specialised bodies spills).
#define W3_MULTI_SHAPES(X) X(4, 16, 3, 3) X(4, 16, 4, 4) X(4, 24, 4, 4) X(8, 16, 4, 4) X(8, 16, 5, 5)

__global__ __launch_bounds__(kW3Threads, 2) void k_window_v3_multi(W3Multi m) {
  // the batch of this workgroup: the item ranges of the batches follow each
  // other, [base, end) is that of batch k
  int k = 0, base = 0, end = m.it[0].items;
#pragma unroll
  for (int i = 1; i < kW3MaxBatches; ++i) {
    if (i < m.n && (int)blockIdx.x >= end) {
      base = end;
      end += m.it[i].items;
      k = i;
    }
  }
  k = __builtin_amdgcn_readfirstlane(k);
  const W3Item& t = m.it[k];
  const int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x - base);
  if (item >= t.items) return;
  int sid = 0;
#define W3_MULTI_CASE(LL, MM, UU, JJ)                                                                  \
  if (t.shape == sid) w3_body<LL, MM, UU, JJ, true>(t.a, t.frames, t.tabs, t.factor, t.ref, t.fields, item); \
  ++sid;
  W3_MULTI_SHAPES(W3_MULTI_CASE)
#undef W3_MULTI_CASE
}

// The shape index of a batch in W3_MULTI_SHAPES (-1: none), with its LDS bytes.
static int w3_multi_shape(W3Args& a, int L, int M, size_t& bytes) {
  const Geometry& g = a.g;
  if (g.H != g.W) return -1;   // square windows: the compile-time column length
  const int U = g.w / 16, J = g.h / 16;
  int sid = 0, shape = -1;
#define W3_MULTI_SHAPE(LL, MM, UU, JJ)                                                             \
  if (shape < 0 && L == LL && M == MM && U == UU && J == JJ && g.H % W3Shape<LL, MM, UU, JJ>::RR == 0) { \
    shape = sid;                                                                                  \
    bytes = w3_layout<LL, MM, UU, JJ>(a);                                                         \
  }                                                                                               \
  ++sid;
  W3_MULTI_SHAPES(W3_MULTI_SHAPE)
#undef W3_MULTI_SHAPE
  return bytes <= 160 * 1024 ? shape : -1;
}

bool window_v3_multi_supported(const W3Launch& b) {
  W3Args a;
  int L, M;
  size_t bytes = 0;
  if (!w3_args(b.fa, b.hgx16, b.hgy16, b.coefs, b.cal, b.hgx32, b.hgy32, a, L, M)) return false;
  return w3_multi_shape(a, L, M, bytes) >= 0;
}

hipError_t launch_window_v3_multi(const W3Launch* b, int n, hipStream_t s) {
  static_assert(sizeof(W3Multi) <= 4000, "kernel arguments too large");
  if (n < 1 || n > kW3MaxBatches) return hipErrorInvalidValue;
  W3Multi m{};
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    W3Item& t = m.it[i];
    int L, M;
    size_t bytes = 0;
    if (!w3_args(b[i].fa, b[i].hgx16, b[i].hgy16, b[i].coefs, b[i].cal, b[i].hgx32, b[i].hgy32, t.a, L, M))
      return hipErrorNotSupported;
    t.shape = w3_multi_shape(t.a, L, M, bytes);
    if (t.shape < 0) return hipErrorNotSupported;
    lds = std::max(lds, bytes);
    t.frames = static_cast<const unsigned short*>(b[i].frames);
    t.tabs = b[i].tab;
    t.factor = b[i].factor;
    t.ref = b[i].ref;
    t.fields = b[i].fields;
    t.items = t.a.g.P * t.a.g.B;
    m.total += t.items;
  }
  m.n = n;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_window_v3_multi),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  k_window_v3_multi<<<m.total, kW3Threads, lds, s>>>(m);
  return hipGetLastError();
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
