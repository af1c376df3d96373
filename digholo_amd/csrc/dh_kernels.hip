// HIP kernels of the batch-processing path (gfx950).
//
// Data flow of one batch (B frames, P polarisations):
//   frames (u16/f32) --k_row_fft--> rowspec[B][P][H][w]   r2c rows, only the w
//                                                          columns the Fourier
//                                                          window reads
//   rowspec --k_col_fft_window--> win[B][P][h][w]         column FFT, keep the h
//                                                          window rows, times the
//                                                          window factor
//   win --k_ifft_rows--> win (in place)                    inverse FFT along x
//   win --k_ifft_cols_tilt--> fields[B][P][h][w]           inverse FFT along y,
//                                                          times reference wave,
//                                                          V = -conj(E ref)
//   fields --k_overlap_mfma (dh_overlap.hip)--> coefs[B][P][M]   separable HG overlap
//
// All transforms are batched mixed-radix Stockham FFTs staged in LDS
// (radix 8/4/2/3/5 butterflies, any other prime by a direct DFT).
#include <algorithm>

#include "dh_fft_ip.h"

namespace dh {

// One radix-R Stockham butterfly: reads src[j + r n/R], applies the stage
// twiddles, writes dst[(j/Ns) Ns R + j%Ns + r Ns].
template <int R, bool INV>
__device__ __forceinline__ void stage_bf(const float2* S, float2* D, int j, int jm, int nbf, int Ns,
                                         int step, const float2* tw) {
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = S[j + r * nbf];
  if (jm) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<INV>(tw, jm * r * step));
  }
  dftR<R, INV>(v);
  const int idxD = (j - jm) * R + jm;
#pragma unroll
  for (int r = 0; r < R; ++r) D[idxD + r * Ns] = v[r];
}

// Radix without a dedicated butterfly: direct DFT from LDS.
template <bool INV>
__device__ void stage_bf_generic(const float2* S, float2* D, int R, int n, int j, int jm, int nbf, int Ns,
                                 int step, const float2* tw) {
  const int nR = n / R;
  const int idxD = (j - jm) * R + jm;
  for (int k = 0; k < R; ++k) {
    float2 acc = make_float2(0.f, 0.f);
    for (int r = 0; r < R; ++r) {
      float2 v = S[j + r * nbf];
      if (r && jm) v = cmul(v, twiddle<INV>(tw, jm * r * step));
      acc = cadd(acc, cmul(v, twiddle<INV>(tw, ((r * k) % R) * nR)));
    }
    D[idxD + k * Ns] = acc;
  }
}

// nb transforms of length pl.n, transform t at a[t*n .. t*n+n).  Uses b as the
// ping-pong buffer; returns whichever buffer holds the result.  Must be called
// by the whole block; ends with a barrier.
template <bool INV>
__device__ float2* lds_fft(float2* a, float2* b, int nb, const FFTPlan& pl) {
  const int n = pl.n;
  int Ns = 1;
  float2* src = a;
  float2* dst = b;
  for (int s = 0; s < pl.nstages; ++s) {
    const int R = pl.radix[s];
    const int nbf = n / R;
    const int total = nbf * nb;
    const int step = n / (Ns * R);
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
      const int t = q / nbf;
      const int j = q - t * nbf;
      const int jm = j % Ns;
      const float2* S = src + t * n;
      float2* D = dst + t * n;
      switch (R) {
        case 8: stage_bf<8, INV>(S, D, j, jm, nbf, Ns, step, pl.tw); break;
        case 4: stage_bf<4, INV>(S, D, j, jm, nbf, Ns, step, pl.tw); break;
        case 2: stage_bf<2, INV>(S, D, j, jm, nbf, Ns, step, pl.tw); break;
        case 3: stage_bf<3, INV>(S, D, j, jm, nbf, Ns, step, pl.tw); break;
        case 5: stage_bf<5, INV>(S, D, j, jm, nbf, Ns, step, pl.tw); break;
        default: stage_bf_generic<INV>(S, D, R, n, j, jm, nbf, Ns, step, pl.tw); break;
      }
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// 1. frame ingest + r2c FFT along x, keeping the Fourier-window columns
// ---------------------------------------------------------------------------
// grid (ceil(H/RB), P, B).  Row y of window p of frame b is packed as
// z[k] = x[2k] + i x[2k+1] (k < W/2), transformed, and unpacked with
// X[k] = (Z[k] + conj Z[n-k])/2 - i/2 exp(-2 pi i k/W) (Z[k] - conj Z[n-k]).
__global__ __launch_bounds__(kThreads) void k_row_fft(Geometry g, FFTPlan pl, const float2* __restrict__ twW,
                                                      const void* __restrict__ frames, int u16,
                                                      float2* __restrict__ rowspec, int RB) {
  extern __shared__ float2 smem[];
  const int n = pl.n;
  float2* A = smem;
  float2* Bf = smem + RB * n;
  const int b = blockIdx.z + g.b0, p = blockIdx.y, y0 = blockIdx.x * RB;
  const int rows = min(RB, g.H - y0);
  const size_t fbase = (size_t)b * g.frameH * g.frameW + (size_t)(g.cy[p] + y0) * g.frameW + g.cx[p];
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
    const int r = e / n, k = e - r * n;
    const size_t off = fbase + (size_t)r * g.frameW + 2 * k;
    float2 z;
    if (u16) {
      const ushort2 v = *reinterpret_cast<const ushort2*>(static_cast<const unsigned short*>(frames) + off);
      z = make_float2((float)v.x, (float)v.y);
    } else {
      z = *reinterpret_cast<const float2*>(static_cast<const float*>(frames) + off);
    }
    A[e] = z;
  }
  __syncthreads();
  const float2* Z = lds_fft<false>(A, Bf, rows, pl);
  const int c0 = g.c0[p];
  float2* out = rowspec + ((size_t)(b * g.P + p) * g.H + y0) * g.w;
  for (int e = threadIdx.x; e < rows * g.w; e += blockDim.x) {
    const int r = e / g.w, kk = e - r * g.w;
    const int k = c0 + kk;                   // k <= n (= W/2, the Nyquist column)
    const float2 zk = Z[r * n + (k == n ? 0 : k)];
    float2 zc = Z[r * n + ((k == 0 || k == n) ? 0 : n - k)];
    zc.y = -zc.y;
    const float2 ev = cscale(cadd(zk, zc), 0.5f);
    const float2 od = mul_mi<false>(cscale(csub(zk, zc), 0.5f));   // (zk - zc)/(2i)
    out[e] = cadd(ev, cmul(twW[k], od));
  }
}

// ---------------------------------------------------------------------------
// 2. FFT along y of the w kept columns, pick the h window rows, apply the
//    Fourier-window factor (mask * fill-factor envelope * centring ramp * norm)
// ---------------------------------------------------------------------------
// grid (ceil(w/CB), P, B)
__global__ __launch_bounds__(kThreads) void k_col_fft_window(Geometry g, FFTPlan pl,
                                                             const float2* __restrict__ rowspec,
                                                             const float2* __restrict__ factor,
                                                             float2* __restrict__ win, int CB, BatchCal cal) {
  extern __shared__ float2 smem[];
  const int n = pl.n;
  float2* A = smem;
  float2* Bf = smem + CB * n;
  const int b = blockIdx.z + g.b0, p = blockIdx.y, i0 = blockIdx.x * CB;
  const int cols = min(CB, g.w - i0);
  const float2* in = rowspec + (size_t)(b * g.P + p) * g.H * g.w + i0;
  for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
    const int y = e / cols, c = e - y * cols;
    A[c * n + y] = in[(size_t)y * g.w + c];
  }
  __syncthreads();
  const float2* F = lds_fft<false>(A, Bf, cols, pl);
  const float2* fac = factor + (size_t)(lam_in(g, b) * g.P + p) * g.h * g.w + i0;
  float2* out = win + (size_t)(out_pos(g, b) * g.P + p) * g.h * g.w + i0;
  const int sy = g.SY[p] + n / 2;
  const float2 cb = cal.at(p, b);
  for (int e = threadIdx.x; e < g.h * cols; e += blockDim.x) {
    const int j = e / cols, c = e - j * cols;
    int row = (j + sy) % n;
    if (row < 0) row += n;
    out[(size_t)j * g.w + c] = cmul(cmul(F[c * n + row], fac[(size_t)j * g.w + c]), cb);
  }
}

// FFT along y of every column of the R2C rows, in place (the whole Fourier
// plane, for its analysis).  grid (ceil(w/CB), P, B)
__global__ __launch_bounds__(kThreads) void k_col_fft_full(Geometry g, FFTPlan pl, float2* __restrict__ spec,
                                                           int CB) {
  extern __shared__ float2 smem[];
  const int n = pl.n;
  float2* A = smem;
  float2* Bf = smem + CB * n;
  const int b = blockIdx.z + g.b0, p = blockIdx.y, i0 = blockIdx.x * CB;
  const int cols = min(CB, g.w - i0);
  float2* base = spec + (size_t)(b * g.P + p) * g.H * g.w + i0;
  for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
    const int y = e / cols, c = e - y * cols;
    A[c * n + y] = base[(size_t)y * g.w + c];
  }
  __syncthreads();
  const float2* F = lds_fft<false>(A, Bf, cols, pl);
  for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
    const int y = e / cols, c = e - y * cols;
    base[(size_t)y * g.w + c] = F[c * n + y];
  }
}

// ---------------------------------------------------------------------------
// 3. inverse FFT along x, in place.  grid (ceil(h/RB), P, B)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_ifft_rows(Geometry g, FFTPlan pl, float2* __restrict__ win,
                                                        int RB) {
  extern __shared__ float2 smem[];
  const int n = pl.n;
  float2* A = smem;
  float2* Bf = smem + RB * n;
  const int b = blockIdx.z + g.b0, p = blockIdx.y, y0 = blockIdx.x * RB;
  const int rows = min(RB, g.h - y0);
  float2* base = win + ((size_t)(b * g.P + p) * g.h + y0) * g.w;
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) A[e] = base[e];
  __syncthreads();
  const float2* X = lds_fft<true>(A, Bf, rows, pl);
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) base[e] = X[e];
}

// ---------------------------------------------------------------------------
// 4. inverse FFT along y, remove the reference wave: V = -conj(E * ref).
//    grid (ceil(w/CB), P, B)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_ifft_cols_tilt(Geometry g, FFTPlan pl,
                                                             const float2* __restrict__ win,
                                                             const float2* __restrict__ ref,
                                                             float2* __restrict__ fields, int CB) {
  extern __shared__ float2 smem[];
  const int n = pl.n;
  float2* A = smem;
  float2* Bf = smem + CB * n;
  const int b = blockIdx.z + g.b0, p = blockIdx.y, i0 = blockIdx.x * CB;
  const int cols = min(CB, g.w - i0);
  const size_t base = (size_t)(b * g.P + p) * g.h * g.w + i0;
  for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
    const int y = e / cols, c = e - y * cols;
    A[c * n + y] = win[base + (size_t)y * g.w + c];
  }
  __syncthreads();
  const float2* E = lds_fft<true>(A, Bf, cols, pl);
  if (!ref) {   // the field before tilt removal (the IFFT step on its own)
    for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
      const int j = e / cols, c = e - j * cols;
      fields[base + (size_t)j * g.w + c] = E[c * n + j];
    }
    return;
  }
  const float2* rp = ref + (size_t)(lam_out(g, b) * g.P + p) * g.h * g.w + i0;
  for (int e = threadIdx.x; e < n * cols; e += blockDim.x) {
    const int j = e / cols, c = e - j * cols;
    const float2 t = cmul(E[c * n + j], rp[(size_t)j * g.w + c]);
    fields[base + (size_t)j * g.w + c] = make_float2(-t.x, t.y);
  }
}

// ---------------------------------------------------------------------------
// Staged path on in-place transforms (every radix 2/3/4/5/8).  Each block
// stages its transforms' twiddles in LDS, index arithmetic is multiply-high
// division, and the inverse transform along y runs in the column kernel, so
// the window plane makes one trip through HBM fewer than in the kernels above:
//   frames  --k_row_fft_ip-----> rowspec[B][P][H][w]
//   rowspec --k_col_window_iy--> win[B][P][h][w]     column FFT, window rows,
//                                                    factor, inverse FFT along y
//   win     --k_ifft_rows_tilt-> fields[B][P][h][w]  inverse FFT along x,
//                                                    V = -conj(E ref)
// ---------------------------------------------------------------------------
constexpr int kRowThreads = 512;   // row and column kernels: one stage pass covers twice the
constexpr int kColThreads = 512;   // transforms of a 256-thread block
constexpr int kSel = 8;            // window values per thread held across the row selection
constexpr int kColV = 16;          // column kernel: values per thread in one transform stage
constexpr int kLd = 8;             // global loads per thread in flight in the copy loops

// for e in [0, cnt): put(e, get(e)), kLd loads issued before the first is used
template <int NT = kThreads, class Get, class Put>
__device__ __forceinline__ void batched(int cnt, Get&& get, Put&& put) {
  for (int e0 = threadIdx.x; e0 < cnt; e0 += kLd * NT) {
    float2 v[kLd];
#pragma unroll
    for (int k = 0; k < kLd; ++k) {
      const int e = e0 + k * NT;
      if (e < cnt) v[k] = get(e);
    }
#pragma unroll
    for (int k = 0; k < kLd; ++k) {
      const int e = e0 + k * NT;
      if (e < cnt) put(e, v[k]);
    }
  }
}

// grid (ceil(H/RB), P, B); rows as in k_row_fft.  NC: the transform length
// when it is a power of two with a compile-time transform, else 0.
template <int NC>
__global__ __launch_bounds__(kRowThreads) void k_row_fft_ip(Geometry g, FFTPlan pl, const float2* __restrict__ twW,
                                                         const void* __restrict__ frames, int u16,
                                                         float2* __restrict__ rowspec, int RB, int chunk,
                                                         unsigned magN, unsigned magW) {
  extern __shared__ float2 smem[];
  const int n = NC ? NC : pl.n, w = g.w;
  float2* tw = smem;          // [n] transform twiddles
  float2* tu = smem + n;      // [w] unpack twiddles exp(-2 pi i k / W) of the window columns
  float2* A = tu + w;         // [RB][LD], point k of row r at r LD + pad32(k)
  const int LD = pad_stride(n);
  const int b = blockIdx.z + g.b0, p = blockIdx.y, y0 = blockIdx.x * RB;
  const int rows = min(RB, g.H - y0);
  const int c0 = g.c0[p];
  for (int e = threadIdx.x; e < n; e += kRowThreads) tw[e] = pl.tw[e];
  for (int e = threadIdx.x; e < w; e += kRowThreads) tu[e] = twW[c0 + e];
  const size_t fbase = (size_t)b * g.frameH * g.frameW + (size_t)(g.cy[p] + y0) * g.frameW + g.cx[p];
  auto rowof = [&](int e) { return NC ? (int)((unsigned)e / (unsigned)NC) : fdiv(e, magN); };
  auto put = [&](int e, float2 v) {
    const int r = rowof(e);
    A[r * LD + pad32(e - r * n)] = v;
  };
  if (u16) {
    const unsigned short* f = static_cast<const unsigned short*>(frames) + fbase;
    batched<kRowThreads>(rows * n, [&](int e) {
      const int r = rowof(e), k = e - r * n;
      const ushort2 v = *reinterpret_cast<const ushort2*>(f + (size_t)r * g.frameW + 2 * k);
      return make_float2((float)v.x, (float)v.y);
    }, put);
  } else {
    const float* f = static_cast<const float*>(frames) + fbase;
    batched<kRowThreads>(rows * n, [&](int e) {
      const int r = rowof(e), k = e - r * n;
      return *reinterpret_cast<const float2*>(f + (size_t)r * g.frameW + 2 * k);
    }, put);
  }
  __syncthreads();
  if constexpr (NC > 0)
    lds_fft_pow2<NC, false, kRowThreads>(A, rows, LD, tw);
  else
    lds_fft_ip<false, kRowThreads, true>(A, rows, LD, 1, false, pl, tw, chunk);
  float2* out = rowspec + ((size_t)(b * g.P + p) * g.H + y0) * w;
  for (int e = threadIdx.x; e < rows * w; e += kRowThreads) {
    const int r = fdiv(e, magW), kk = e - r * w;
    const int k = c0 + kk;                   // k <= n (= W/2, the Nyquist column)
    const float2 zk = A[r * LD + pad32(k == n ? 0 : k)];
    float2 zc = A[r * LD + pad32((k == 0 || k == n) ? 0 : n - k)];
    zc.y = -zc.y;
    const float2 ev = cscale(cadd(zk, zc), 0.5f);
    const float2 od = mul_mi<false>(cscale(csub(zk, zc), 0.5f));   // (zk - zc)/(2i)
    out[e] = cadd(ev, cmul(tu[kk], od));
  }
}

// grid (ceil(w/CB), P, B).  Loads and stores move the strip's CB * 8
// contiguous bytes per row; in LDS the strip is column-major, each column
// padded.  h * CB <= kColThreads * kSel (the selected window is held in
// registers while it moves to the top of the strip).  NC: H when the column
// transform has a compile-time form (a power of two), else 0.
template <int NC>
__global__ __launch_bounds__(kColThreads) void k_col_window_iy(Geometry g, FFTPlan pc, FFTPlan py,
                                                            const float2* __restrict__ rowspec,
                                                            const float2* __restrict__ factor,
                                                            float2* __restrict__ win, int CB, int chunkC,
                                                            int chunkY, BatchCal cal, unsigned* __restrict__ peak) {
  extern __shared__ float2 smem[];
  const int H = NC ? NC : g.H, h = g.h, w = g.w;
  float2* twc = smem;
  float2* twy = smem + H;
  float2* A = twy + h;        // [CB][LD], point y of column c at c LD + pad32(y)
  const int LD = pad_stride(H);
  const int b = blockIdx.z + g.b0, p = blockIdx.y, i0 = blockIdx.x * CB;
  const int cols = min(CB, w - i0);
  const unsigned magC = magic_div((unsigned)cols);
  if (peak && blockIdx.x == 0 && threadIdx.x == 0) peak[out_pos(g, b) * g.P + p] = 0u;
  for (int e = threadIdx.x; e < H; e += kColThreads) twc[e] = pc.tw[e];
  for (int e = threadIdx.x; e < h; e += kColThreads) twy[e] = py.tw[e];
  const float2* in = rowspec + (size_t)(b * g.P + p) * H * w + i0;
  batched<kColThreads>(H * cols, [&](int e) {
    const int y = fdiv(e, magC), c = e - y * cols;
    return in[(size_t)y * w + c];
  }, [&](int e, float2 v) {
    const int y = fdiv(e, magC), c = e - y * cols;
    A[c * LD + pad32(y)] = v;
  });
  // window row j is spectrum row (j + SY + H/2) mod H; its factor is loaded
  // while the column transform runs
  int sy = (g.SY[p] + H / 2) % H;
  if (sy < 0) sy += H;
  const float2* fac = factor + (size_t)(lam_in(g, b) * g.P + p) * h * w + i0;
  float2 sel[kSel];
#pragma unroll
  for (int s = 0; s < kSel; ++s) {
    const int e = threadIdx.x + s * kColThreads;
    if (e < h * cols) {
      const int j = fdiv(e, magC), c = e - j * cols;
      sel[s] = fac[(size_t)j * w + c];
    }
  }
  __syncthreads();
  if constexpr (NC > 0)
    lds_fft_pow2<NC, false, kColThreads, kColV>(A, cols, LD, twc);
  else
    lds_fft_ip<false, kColThreads, true, kColV>(A, cols, LD, 1, false, pc, twc, chunkC);
  const float2 cb = cal.at(p, b);
#pragma unroll
  for (int s = 0; s < kSel; ++s) {
    const int e = threadIdx.x + s * kColThreads;
    if (e < h * cols) {
      const int j = fdiv(e, magC), c = e - j * cols;
      int row = j + sy;
      if (row >= H) row -= H;
      sel[s] = cmul(cmul(A[c * LD + pad32(row)], sel[s]), cb);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kSel; ++s) {
    const int e = threadIdx.x + s * kColThreads;
    if (e < h * cols) {
      const int j = fdiv(e, magC), c = e - j * cols;
      A[c * LD + pad32(j)] = sel[s];
    }
  }
  __syncthreads();
  lds_fft_any<true, kColThreads>(A, cols, LD, py, twy, chunkY);
  float2* out = win + (size_t)(out_pos(g, b) * g.P + p) * h * w + i0;
  for (int e = threadIdx.x; e < h * cols; e += kColThreads) {
    const int j = fdiv(e, magC), c = e - j * cols;
    out[(size_t)j * w + c] = A[c * LD + pad32(j)];
  }
}

// grid (ceil(h/RB), P, B); win holds the plane already transformed along y.
// ref null: the field before the reference wave is removed.
__global__ __launch_bounds__(kThreads) void k_ifft_rows_tilt(Geometry g, FFTPlan pl, const float2* __restrict__ win,
                                                             const float2* __restrict__ ref,
                                                             float2* __restrict__ fields, int RB, int chunk,
                                                             unsigned* __restrict__ peak) {
  extern __shared__ float2 smem[];
  const int n = pl.n;   // = w
  float2* tw = smem;
  float2* A = smem + n;        // [RB][LD], point k of row r at r LD + pad32(k)
  const int LD = pad_stride(n);
  const unsigned magN = magic_div((unsigned)n);
  auto at = [&](int e) -> float2& {
    const int r = fdiv(e, magN);
    return A[r * LD + pad32(e - r * n)];
  };
  const int b = blockIdx.z + g.b0, p = blockIdx.y, y0 = blockIdx.x * RB;
  const int rows = min(RB, g.h - y0);
  for (int e = threadIdx.x; e < n; e += kThreads) tw[e] = pl.tw[e];
  const size_t base = ((size_t)(b * g.P + p) * g.h + y0) * n;
  batched(rows * n, [&](int e) { return win[base + e]; }, [&](int e, float2 v) { at(e) = v; });
  __syncthreads();
  lds_fft_any<true, kThreads>(A, rows, LD, pl, tw, chunk);
  if (!ref) {
    for (int e = threadIdx.x; e < rows * n; e += kThreads) fields[base + e] = at(e);
    return;
  }
  const float2* rp = ref + ((size_t)(lam_out(g, b) * g.P + p) * g.h + y0) * n;
  float mx = 0.f;
  batched(rows * n, [&](int e) { return rp[e]; }, [&](int e, float2 r) {
    const float2 t = cmul(at(e), r);
    fields[base + e] = make_float2(-t.x, t.y);
    mx = fmaxf(mx, fmaxf(fabsf(t.x), fabsf(t.y)));
  });
  if (peak) {   // one atomicMax per block; non-negative floats order as their bit patterns
    __shared__ float s_mx[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) s_mx[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < kThreads / 64; ++i) mx = fmaxf(mx, s_mx[i]);
      atomicMax(peak + b * g.P + p, __float_as_uint(mx));
    }
  }
}

// ---------------------------------------------------------------------------
// Step kernels (the pipeline run one stage at a time)
// ---------------------------------------------------------------------------
// win[out(b)][p][j][c] = spec[b][p][row(j)][c] * factor[lam(b)][p][j][c], spec
// the column-transformed band of the w window columns ([B][P][H][w]).
__global__ void k_select_window(Geometry g, const float2* __restrict__ spec, const float2* __restrict__ factor,
                                float2* __restrict__ win, BatchCal cal) {
  const size_t hw = (size_t)g.h * g.w, total = (size_t)g.B * g.P * hw;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t bp = e / hw;
    const int p = (int)(bp % g.P);
    const int rem = (int)(e - bp * hw);
    const int j = rem / g.w, c = rem - j * g.w;
    int row = (j + g.SY[p] + g.H / 2) % g.H;
    if (row < 0) row += g.H;
    const int b = (int)(bp / g.P);
    const float2 fac = factor[(size_t)(lam_in(g, b) * g.P + p) * hw + rem];
    win[((size_t)out_pos(g, b) * g.P + p) * hw + rem] = cmul(cmul(spec[(bp * g.H + row) * g.w + c], fac), cal.at(p, b));
  }
}

// Full-resolution IFFT input (digHolo.cpp:1855-1872, :1938-1950, :2042-2065):
// the window keeps its own frequency position in a zero W x H plane.
__global__ void k_scatter_full(Geometry g, const float2* __restrict__ win, float2* __restrict__ plane) {
  const size_t n = (size_t)g.H * g.W, total = (size_t)g.B * g.P * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t bp = e / n;
    const int p = (int)(bp % g.P);
    const int rem = (int)(e - bp * n);
    const int row = rem / g.W, col = rem - row * g.W;
    int j = (row - g.SY[p] - g.H / 2) % g.H;
    if (j < 0) j += g.H;
    const int i = col - g.c0[p];
    float2 v = make_float2(0.f, 0.f);
    if (j < g.h && i >= 0 && i < g.w) v = win[(bp * g.h + j) * g.w + i];
    plane[e] = v;
  }
}

hipError_t launch_scatter_full(const Geometry& g, const float2* win, float2* plane, hipStream_t s) {
  const size_t n = (size_t)g.B * g.P * g.H * g.W;
  k_scatter_full<<<stride_blocks(n, 256), 256, 0, s>>>(g, win, plane);
  return hipGetLastError();
}

// uint16 frames given transposed (digHolo.cpp:1206-1229): out[f][i H + j] = in[f][j W + i]
__global__ void k_u16_transpose(const unsigned short* __restrict__ in, unsigned short* __restrict__ out, int W,
                                int H, size_t frames) {
  const size_t n = (size_t)W * H, total = frames * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t f = e / n;
    const int o = (int)(e - f * n);
    const int i = o / H, j = o - i * H;
    out[e] = in[f * n + (size_t)j * W + i];
  }
}

// In-phase averaging of A Fourier windows per (batch element, pol)
// (digHolo.cpp:1885-1928, :2066-2114): W = o_0 / A; for a >= 1,
// u = phi / |phi| with phi = sum_pixels W conj(o_a), W += o_a u / A.  o_a is the
// calibrated window of frame f(b, a) (layout by avgMode, digHolo.cpp:8661-8672).
// One 256-thread block per (b, p); W is held in registers (up to kAvgPer per thread).
constexpr int kAvgPer = 32;
__global__ __launch_bounds__(256) void k_average_windows(Geometry g, int A, int mode, int frames,
                                                         const float2* __restrict__ winAll,
                                                         float2* __restrict__ win, BatchCal cal) {
  __shared__ float2 red[256 / 64];
  const int p = blockIdx.x, b = blockIdx.y + g.b0, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int hw = g.h * g.w;
  const float inv = 1.f / (float)A;
  const float2 cb = cal.at(p, b);
  float2 Wv[kAvgPer];
#pragma unroll
  for (int s = 0; s < kAvgPer; ++s) Wv[s] = make_float2(0.f, 0.f);
  for (int a = 0; a < A; ++a) {
    int f = mode == 0 ? b * A + a : (mode == 1 ? b + a * g.B : b + a);
    if (f >= frames) f = frames - 1;   // guarded on the host; never past the buffer
    const float2* o = winAll + (size_t)(f * g.P + p) * hw;
    float2 u = make_float2(1.f, 0.f);
    if (a > 0) {
      float2 phi = make_float2(0.f, 0.f);
#pragma unroll
      for (int s = 0; s < kAvgPer; ++s) {
        const int e = tid + 256 * s;
        if (e < hw) {
          const float2 oa = cmul(o[e], cb);
          phi.x += Wv[s].x * oa.x + Wv[s].y * oa.y;   // W conj(o)
          phi.y += Wv[s].y * oa.x - Wv[s].x * oa.y;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        phi.x += __shfl_xor(phi.x, off);
        phi.y += __shfl_xor(phi.y, off);
      }
      __syncthreads();   // red[] of the previous frame has been read by everyone
      if (lane == 0) red[wave] = phi;
      __syncthreads();
      phi = make_float2(red[0].x + red[1].x + red[2].x + red[3].x, red[0].y + red[1].y + red[2].y + red[3].y);
      const float mag = sqrtf(phi.x * phi.x + phi.y * phi.y);
      u = mag > 0.f ? make_float2(phi.x / mag, phi.y / mag) : make_float2(1.f, 0.f);
    }
#pragma unroll
    for (int s = 0; s < kAvgPer; ++s) {
      const int e = tid + 256 * s;
      if (e < hw) Wv[s] = cadd(Wv[s], cscale(cmul(cmul(o[e], cb), u), inv));
    }
  }
  float2* out = win + (size_t)(b * g.P + p) * hw;
#pragma unroll
  for (int s = 0; s < kAvgPer; ++s) {
    const int e = tid + 256 * s;
    if (e < hw) out[e] = Wv[s];
  }
}

hipError_t launch_u16_transpose(const void* in, void* out, int W, int H, size_t frames, hipStream_t s) {
  const size_t n = frames * W * H;
  k_u16_transpose<<<stride_blocks(n, 256), 256, 0, s>>>(static_cast<const unsigned short*>(in),
                                                        static_cast<unsigned short*>(out), W, H, frames);
  return hipGetLastError();
}

hipError_t launch_average_windows(const Geometry& g, int A, int mode, int frames, const float2* winAll, float2* win,
                                  const BatchCal& cal, hipStream_t s) {
  if (g.h * g.w > 256 * kAvgPer) return hipErrorNotSupported;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_average_windows<<<dim3(g.P, nb), 256, 0, s>>>(gc, A, mode, frames, winAll, win, cal);
    return hipGetLastError();
  });
}

// fields = -conj(fields * ref), in place; fields in output order, ref of the
// position's wavelength
__global__ void k_apply_tilt(Geometry g, const float2* __restrict__ ref, float2* __restrict__ fields) {
  const size_t hw = (size_t)g.h * g.w, total = (size_t)g.B * g.P * hw;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t bp = e / hw;
    const int p = (int)(bp % g.P);
    const int lam = lam_out(g, (int)(bp / g.P));
    const float2 t = cmul(fields[e], ref[((size_t)lam * g.P + p) * hw + e % hw]);
    fields[e] = make_float2(-t.x, t.y);
  }
}

hipError_t launch_select_window(const Geometry& g, const float2* spec, const float2* factor, float2* win,
                                const BatchCal& cal, hipStream_t s) {
  const size_t n = (size_t)g.B * g.P * g.h * g.w;
  k_select_window<<<stride_blocks(n, 256), 256, 0, s>>>(g, spec, factor, win, cal);
  return hipGetLastError();
}

hipError_t launch_apply_tilt(const Geometry& g, const float2* ref, float2* fields, hipStream_t s) {
  const size_t n = (size_t)g.B * g.P * g.h * g.w;
  k_apply_tilt<<<stride_blocks(n, 256), 256, 0, s>>>(g, ref, fields);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 tables: Fourier-window factor, reference wave, HG basis
// ---------------------------------------------------------------------------
__device__ __forceinline__ double xcsc(double t) { return t == 0.0 ? 1.0 : t / sin(t); }

// factor[p][j][i] = mask * S(kx) S(ky) exp(i(kx bcx + ky bcy)) * norm, with the
// FFT-order frequency of the r2c row/column each window pixel reads;
// ref[p][j][i] = exp(i(phi_x(i) + phi_y(j))).
__global__ void k_tables(TableArgs a, const uint8_t* __restrict__ mask, float2* __restrict__ factor,
                         float2* __restrict__ ref) {
  const Geometry& g = a.g;
  const int nf = g.P * g.h * g.w;          // window factor over the Fourier window
  const int nr = g.P * g.fh * g.fw;        // reference wave over the field plane
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nf + nr; idx += gridDim.x * blockDim.x) {
    double sn, cs;
    if (idx < nf) {
      const int p = idx / (g.h * g.w);
      const int rem = idx - p * g.h * g.w;
      const int j = rem / g.w, i = rem - j * g.w;
      const PolParams& pp = a.pp[p];
      const int col = g.c0[p] + i;
      int row = (j + g.SY[p] + g.H / 2) % g.H;
      if (row < 0) row += g.H;
      const double kx = col * pp.dkx;
      const double ky = (row <= g.H / 2 ? row : row - g.H) * pp.dky;
      double env = pp.norm;
      if (pp.fill) env *= xcsc(0.5 * kx * pp.px) * xcsc(0.5 * ky * pp.px);
      sincos(kx * pp.bcx + ky * pp.bcy, &sn, &cs);
      const double m = mask[idx] ? env : 0.0;
      factor[idx] = make_float2((float)(m * cs), (float)(m * sn));
    } else {
      const int e = idx - nf;
      const int p = e / (g.fh * g.fw);
      const int rem = e - p * g.fh * g.fw;
      const int j = rem / g.fw, i = rem - j * g.fw;
      const PolParams& pp = a.pp[p];
      const double x = (i - g.fw / 2) * pp.pxx;
      const double y = (j - g.fh / 2) * pp.pxy;
      const double phi = pp.kxr * x + pp.pistx + pp.focal * x * x + pp.kyr * y + pp.pisty + pp.focal * y * y;
      sincos(phi, &sn, &cs);
      ref[e] = make_float2((float)cs, (float)sn);
    }
  }
}

// HG mode n at coordinate t: psi_n(sqrt2 t / waist) sqrt(dt sqrt2 / waist), by
// the three-term recurrence with the mantissa renormalised every step and the
// binary exponent carried separately (no under/overflow at high order).
__global__ void k_hg_basis(TableArgs a, float* __restrict__ hgx, float* __restrict__ hgy) {
  const Geometry& g = a.g;
  const int per = g.fw + g.fh;              // the basis lives on the field plane
  const int total = g.P * per;
  const double rt2 = 1.41421356237309504880;
  const double pim4 = 0.75112554446494248286;   // pi^-1/4
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
    const double u = rt2 * t / pp.waist;
    const double nrm = sqrt(dt * rt2 / pp.waist);
    const double gexp = -0.5 * u * u;
    float* out = isx ? (hgx + (size_t)p * g.G * g.fw + i) : (hgy + (size_t)p * g.G * g.fh + i);
    const size_t stride = N;
    double m0 = pim4, e0 = 0.0;
    double m1 = rt2 * pim4 * u, e1 = 0.0;
    if (g.G > 0) out[0] = (float)(nrm * m0 * exp(gexp));
    if (g.G > 1) out[stride] = (float)(nrm * m1 * exp(gexp));
    for (int n = 2; n < g.G; ++n) {
      const double c1 = sqrt(2.0 / n), c2 = sqrt((n - 1.0) / n);
      double mn = c1 * u * m1 - c2 * m0 * exp2(e0 - e1);
      int fe;
      mn = frexp(mn, &fe);
      const double en = e1 + fe;
      m0 = m1;
      e0 = e1;
      m1 = mn;
      e1 = en;
      out[n * stride] = (float)(nrm * mn * exp(gexp + en * ln2));
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int lds_limit() { return 64 * 1024; }

template <typename K>
static hipError_t set_lds(K kern, size_t bytes) {
  if (bytes > 64 * 1024)
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
  return hipSuccess;
}

// transforms per block so that two ping-pong buffers fit comfortably
static inline int batch_per_block(int n, int cap_elems) {
  int nb = cap_elems / n;
  return nb < 1 ? 1 : nb;
}

hipError_t launch_tables(const TableArgs& a, const uint8_t* d_mask, float2* d_factor, float2* d_ref,
                         float* d_hgx, float* d_hgy, hipStream_t s) {
  const int total = a.g.P * (a.g.h * a.g.w + a.g.fh * a.g.fw);
  k_tables<<<(total + 255) / 256, 256, 0, s>>>(a, d_mask, d_factor, d_ref);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || a.g.G == 0 || !d_hgx) return e;
  const int t2 = a.g.P * (a.g.fw + a.g.fh);
  k_hg_basis<<<(t2 + 63) / 64, 64, 0, s>>>(a, d_hgx, d_hgy);
  return hipGetLastError();
}

hipError_t launch_hg_basis(const TableArgs& a, float* d_hgx, float* d_hgy, hipStream_t s) {
  if (a.g.G == 0) return hipSuccess;
  const int t2 = a.g.P * (a.g.fw + a.g.fh);
  k_hg_basis<<<(t2 + 63) / 64, 64, 0, s>>>(a, d_hgx, d_hgy);
  return hipGetLastError();
}

hipError_t launch_row_fft(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                          int frame_is_u16, float2* rowspec, hipStream_t s) {
  const int RB = std::min(g.H, batch_per_block(pl.n, 2048));
  const size_t lds = 2 * (size_t)RB * pl.n * sizeof(float2);
  hipError_t e = set_lds(k_row_fft, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_row_fft<<<dim3((g.H + RB - 1) / RB, g.P, nb), kThreads, lds, s>>>(gc, pl, twW, frames, frame_is_u16, rowspec, RB);
    return hipGetLastError();
  });
}

hipError_t launch_col_fft_window(const Geometry& g, const FFTPlan& pl, const float2* rowspec,
                                 const float2* factor, float2* win, const BatchCal& cal, hipStream_t s) {
  const int CB = std::min(16, std::min(g.w, batch_per_block(pl.n, 4096)));
  const size_t lds = 2 * (size_t)CB * pl.n * sizeof(float2);
  hipError_t e = set_lds(k_col_fft_window, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_col_fft_window<<<dim3((g.w + CB - 1) / CB, g.P, nb), kThreads, lds, s>>>(gc, pl, rowspec, factor, win, CB, cal);
    return hipGetLastError();
  });
}

hipError_t launch_col_fft_full(const Geometry& g, const FFTPlan& pl, float2* spec, hipStream_t s) {
  const int CB = std::min(16, std::min(g.w, batch_per_block(pl.n, 4096)));
  const size_t lds = 2 * (size_t)CB * pl.n * sizeof(float2);
  hipError_t e = set_lds(k_col_fft_full, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_col_fft_full<<<dim3((g.w + CB - 1) / CB, g.P, nb), kThreads, lds, s>>>(gc, pl, spec, CB);
    return hipGetLastError();
  });
}

hipError_t launch_ifft_rows(const Geometry& g, const FFTPlan& pl, float2* win, hipStream_t s) {
  const int RB = std::min(g.h, batch_per_block(pl.n, 2048));
  const size_t lds = 2 * (size_t)RB * pl.n * sizeof(float2);
  hipError_t e = set_lds(k_ifft_rows, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_ifft_rows<<<dim3((g.h + RB - 1) / RB, g.P, nb), kThreads, lds, s>>>(gc, pl, win, RB);
    return hipGetLastError();
  });
}

// In-place staged kernels: rows per block of the row transforms, columns per
// block of the column kernel (0 when a transform has no in-place butterfly or
// the strip does not fit).
static int ip_rows_per_block(const FFTPlan& pl, int total, int nt = kThreads) {
  const int chunk = ip_chunk(pl, nt);
  if (chunk < 1) return 0;
  return std::min(total, std::max(chunk, 4096 / pl.n));
}

static int ip_cols_per_block(const Geometry& g, const FFTPlan& pc, const FFTPlan& py) {
  if (ip_chunk(pc, kColThreads, kColV) < 1 || ip_chunk(py, kColThreads) < 1) return 0;
  for (int cb = std::min(16, g.w); cb >= 1; --cb) {
    const size_t lds = sizeof(float2) * ((size_t)g.H + g.h + (size_t)pad_stride(g.H) * cb);
    if (g.h * cb <= kColThreads * kSel && lds <= 80 * 1024) return cb;
  }
  return 0;
}

bool staged_ip_supported(const Geometry& g, const FFTPlan& row, const FFTPlan& col, const FFTPlan& ix,
                         const FFTPlan& iy) {
  return ip_rows_per_block(row, g.H, kRowThreads) > 0 && ip_rows_per_block(ix, g.h) > 0 && ix.n == g.w && iy.n == g.h &&
         ip_cols_per_block(g, col, iy) > 0;
}

hipError_t launch_row_fft_ip(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                             int frame_is_u16, float2* rowspec, hipStream_t s) {
  if (pl.n == 128 || pl.n == 256) return launch_row_fft_reg(g, pl, twW, frames, frame_is_u16, rowspec, s);
  const int RB = ip_rows_per_block(pl, g.H, kRowThreads);
  if (RB < 1) return hipErrorNotSupported;
  const size_t lds = sizeof(float2) * ((size_t)pl.n + g.w + (size_t)pad_stride(pl.n) * RB);
  auto kern = k_row_fft_ip<0>;
  switch (pl.n) {
    case 64: kern = k_row_fft_ip<64>; break;
    case 128: kern = k_row_fft_ip<128>; break;
    case 256: kern = k_row_fft_ip<256>; break;
    case 512: kern = k_row_fft_ip<512>; break;
    case 1024: kern = k_row_fft_ip<1024>; break;
    default: break;
  }
  hipError_t e = set_lds(kern, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    kern<<<dim3((g.H + RB - 1) / RB, g.P, nb), kRowThreads, lds, s>>>(gc, pl, twW, frames, frame_is_u16, rowspec,
                                                                      RB, ip_chunk(pl, kRowThreads),
                                                                      magic_div((unsigned)pl.n),
                                                                      magic_div((unsigned)g.w));
    return hipGetLastError();
  });
}

hipError_t launch_col_window_iy(const Geometry& g, const FFTPlan& pc, const FFTPlan& py, const float2* rowspec,
                                const float2* factor, float2* win, const BatchCal& cal, hipStream_t s,
                                unsigned* peak) {
  if ((g.H == 256 || g.H == 512) && pc.n == g.H && py.n == g.h && g.h <= 256 && ip_chunk(py, 256) >= 1)
    return launch_col_reg_window_iy(g, pc, py, rowspec, factor, win, cal, s, peak);
  const int CB = ip_cols_per_block(g, pc, py);
  if (CB < 1 || py.n != g.h) return hipErrorNotSupported;
  const size_t lds = sizeof(float2) * ((size_t)g.H + g.h + (size_t)pad_stride(g.H) * CB);
  auto kern = k_col_window_iy<0>;
  switch (pc.n) {
    case 64: kern = k_col_window_iy<64>; break;
    case 128: kern = k_col_window_iy<128>; break;
    case 256: kern = k_col_window_iy<256>; break;
    case 512: kern = k_col_window_iy<512>; break;
    case 1024: kern = k_col_window_iy<1024>; break;
    default: break;
  }
  hipError_t e = set_lds(kern, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    kern<<<dim3((g.w + CB - 1) / CB, g.P, nb), kColThreads, lds, s>>>(gc, pc, py, rowspec, factor, win, CB,
                                                                      ip_chunk(pc, kColThreads, kColV),
                                                                      ip_chunk(py, kColThreads), cal, peak);
    return hipGetLastError();
  });
}

hipError_t launch_ifft_rows_tilt(const Geometry& g, const FFTPlan& pl, const float2* win, const float2* ref,
                                 float2* fields, hipStream_t s, unsigned* peak) {
  if (pl.n == g.w && ifft_rows_reg_width(g.w)) return launch_ifft_rows_reg(g, pl, win, ref, fields, s, peak);
  const int RB = ip_rows_per_block(pl, g.h);
  if (RB < 1 || pl.n != g.w) return hipErrorNotSupported;
  const size_t lds = sizeof(float2) * ((size_t)pl.n + (size_t)pad_stride(pl.n) * RB);
  hipError_t e = set_lds(k_ifft_rows_tilt, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_ifft_rows_tilt<<<dim3((g.h + RB - 1) / RB, g.P, nb), kThreads, lds, s>>>(gc, pl, win, ref, fields, RB,
                                                                              ip_chunk(pl, kThreads),
                                                                              ref ? peak : nullptr);
    return hipGetLastError();
  });
}

hipError_t launch_ifft_cols_tilt(const Geometry& g, const FFTPlan& pl, const float2* win, const float2* ref,
                                 float2* fields, hipStream_t s) {
  const int CB = std::min(16, std::min(g.w, batch_per_block(pl.n, 4096)));
  const size_t lds = 2 * (size_t)CB * pl.n * sizeof(float2);
  hipError_t e = set_lds(k_ifft_cols_tilt, lds);
  if (e != hipSuccess) return e;
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_ifft_cols_tilt<<<dim3((g.w + CB - 1) / CB, g.P, nb), kThreads, lds, s>>>(gc, pl, win, ref, fields, CB);
    return hipGetLastError();
  });
}

}  // namespace dh
