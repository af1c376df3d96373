// Separable Hermite-Gaussian overlap on MFMA (gfx950).
//
//   stage 1:  T[(y,c)][m] = sum_x V_c[y][x] HGx[m][x]        (c = re/im)
//   stage 2:  C[(m,c)][n] += sum_y T[(y,c)][m] HGy[n][y]      -> coef(m,n), m+n < G
//
// One workgroup (8 waves) per pol and frame slot walks the frames b =
// slot, slot + slots, ... (the basis stays in LDS for all of them) and each
// field in chunks of 16 field rows; the next chunk's field values are loaded
// into registers while the current chunk is multiplied.  Per chunk, stage 1 is
// a (32 x Gp x w) product over the chunk's rows grouped by part (rows 0-15 the
// real parts of field rows 0-15, rows 16-31 the imaginary parts), so each lane
// of a stage-1 tile receives T_c[m][yy] for four consecutive yy, stored as one
// packed write into the A operand of stage 2, a (2Gp x Gp x 16) product
// accumulated in registers across chunks.  Tiles of stage 2 that hold no mode
// with m+n < G are skipped.
//
// Two arithmetic forms:
//  * FP32:   v_mfma_f32_16x16x4_f32 on float operands;
//  * SPLIT:  each operand x = hi + lo with hi = fp16(x), lo = fp16(x - hi);
//            hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 (stage 1) and
//            v_mfma_f32_16x16x16_f16 (stage 2, K = the 16 rows of a chunk),
//            FP32 accumulate.
//            The field is scaled by a power of two per (frame, pol) so that its
//            largest magnitude lies in (128, 256]; every partial sum then stays
//            far inside the fp16 range, and the scale is removed at the end.
//
// MFMA operand layout used (16x16 output tile, 64 lanes, q = lane/16):
//   A: lane holds row lane%16, k-slots kk(q, j);  B: column lane%16, same slots;
//   D: lane holds column lane%16, rows 4q + i (i = 0..3).
// A and B use the same k-slot mapping, so the contraction is independent of how
// the hardware orders k inside an instruction.
#include <algorithm>

#include "dh_internal.h"

namespace dh {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// (xr, xi) -> fp16 pairs hi = (fp16(xr), fp16(xi)), lo = fp16 of the exact
// residuals, both rounded to nearest even (cvt_pk + fma_mix: the values of
// convert / subtract / convert).  hi is an ordinary conversion, so the
// compiler keeps the wait states an MFMA result needs before a VALU reads it;
// the inline v_fma_mix depend on hi and so come after it (inline assembly
// operands are not checked for that hazard).
__device__ __forceinline__ void split_pair(float xr, float xi, unsigned& hi, unsigned& lo) {
  const half2v hv = {(_Float16)xr, (_Float16)xi};
  hi = __builtin_bit_cast(unsigned, hv);
  asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %0, %1, -1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=&v"(lo)
      : "v"(hi), "v"(xr), "v"(xi));
}

constexpr int kChunk = 16;          // field rows per chunk (K of stage 2)

template <bool SPLIT>
struct OvT {
  using E = typename std::conditional<SPLIT, _Float16, float>::type;
};

__device__ __forceinline__ void split16(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// row pitch (elements) of an LDS matrix with K columns: pad by 8 elements to
// spread 16 consecutive rows over the banks
__host__ __device__ constexpr int pitch_of(int K) { return K + 8; }
// row pitch of the stage-2 A operand (16 chunk rows): 20 elements keep its
// 8- and 16-byte accesses aligned and 16 consecutive rows on distinct banks
constexpr int kPitch2 = 20;

template <bool SPLIT, int TPW, int NW>
__global__ __launch_bounds__(64 * NW) void k_overlap_mfma(Geometry g, const float2* __restrict__ fields,
                                                             const float* __restrict__ hgx,
                                                             const float* __restrict__ hgy,
                                                             float2* __restrict__ coefs,
                                                             const unsigned* __restrict__ peak) {
  using E = typename OvT<SPLIT>::E;
  extern __shared__ __align__(16) unsigned char ov_smem[];
  const int p = blockIdx.x;
  const int G = g.G, w = g.w, h = g.h, M = g.M;
  const int Gp = (G + 15) & ~15;
  const int pw = pitch_of(w), ph = pitch_of(h), pk = kPitch2;
  const int nL = SPLIT ? 2 : 1;   // hi (+ lo) planes
  E* Bx = reinterpret_cast<E*>(ov_smem);          // [nL][Gp][pw]   HGx[m][x]
  E* By = Bx + nL * Gp * pw;                      // [nL][Gp][ph]   HGy[n][y]
  E* A1 = By + nL * Gp * ph;                      // [nL][32][pw]   V chunk, rows (c, yy)
  E* A2 = A1 + nL * 32 * pw;                      // [nL][2Gp][pk]  T chunk, rows (m, c)
  constexpr int NT = 64 * NW;
  __shared__ float s_red[NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, q = lane >> 4;

  // --- basis into LDS (rows m >= G are zero), once for every frame of the slot;
  //     eight loads per thread in flight before the first is stored ---
  auto stage_basis = [&](const float* src, E* dst, int len, int pitch) {
    constexpr int kB = 8;
    for (int e0 = tid; e0 < Gp * len; e0 += kB * NT) {
      float v[kB];
#pragma unroll
      for (int k = 0; k < kB; ++k) {
        const int e = e0 + k * NT;
        const int m = e / len, x = e - m * len;
        v[k] = (e < Gp * len && m < G) ? src[((size_t)p * G + m) * len + x] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < kB; ++k) {
        const int e = e0 + k * NT;
        if (e < Gp * len) {
          const int m = e / len, x = e - m * len;
          if (SPLIT) {
            _Float16 hi, lo;
            split16(v[k], hi, lo);
            dst[m * pitch + x] = hi;
            dst[Gp * pitch + m * pitch + x] = lo;
          } else {
            dst[m * pitch + x] = v[k];
          }
        }
      }
    }
  };
  stage_basis(hgx, Bx, w, pw);
  stage_basis(hgy, By, h, ph);

  // --- stage-2 tiles owned by this wave ---
  const int rt2 = (2 * Gp) / 16, ct2 = Gp / 16;
  int myTile[TPW];
  int nMine = 0;
  {
    int cnt = 0;
    for (int t = 0; t < rt2 * ct2; ++t) {
      const int mt = t / ct2, nt = t - mt * ct2;
      if (8 * mt + 16 * nt >= G) continue;          // every (m, n) in the tile has m + n >= G
      if (cnt % NW == wave && nMine < TPW) myTile[nMine++] = t;
      ++cnt;
    }
  }
  const int rt1 = 2, ct1 = Gp / 16;
  // this thread's element pairs of a chunk: (2 e2, 2 e2 + 1), e2 = tid + NT j <
  // 8 w (w is even, so a pair never straddles rows); held in registers one chunk
  // ahead when 8 w <= kPerT NT, else loaded as they are stored
  constexpr int kPerT = 4;
  const int nPer = (kChunk * w / 2 - tid + NT - 1) / NT;
  const bool ahead = kChunk * w / 2 <= kPerT * NT;
  for (int b = blockIdx.y; b < g.B; b += gridDim.y) {
  const float2* V = fields + (size_t)(b * g.P + p) * h * w;
  const float4* V4 = reinterpret_cast<const float4*>(V);

  // --- per-field power-of-two scale (split form only) ---
  float scale = 1.f;
  if (SPLIT) {
    float mx = 0.f;
    if (peak) {   // left by the kernel that wrote the field
      mx = __uint_as_float(peak[b * g.P + p]);
    } else {
      for (int e = tid; e < h * w; e += NT) {
        const float2 v = V[e];
        mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
      }
      for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
      __syncthreads();   // s_red of the previous frame has been read
      if (lane == 0) s_red[wave] = mx;
      __syncthreads();
      mx = s_red[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) mx = fmaxf(mx, s_red[i]);
    }
    if (mx > 0.f) {
      int ex;
      frexpf(mx, &ex);                 // mx in [2^(ex-1), 2^ex)
      scale = ldexpf(1.f, 8 - ex);     // max * scale in [128, 256)
    }
  }

  floatx4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};

  float4 nxt[kPerT];   // the next chunk's pairs, in flight during the current chunk
#pragma unroll
  for (int j = 0; j < kPerT; ++j)
    if (ahead && j < nPer) nxt[j] = V4[tid + NT * j];
  for (int y0 = 0; y0 < h; y0 += kChunk) {
    __syncthreads();   // previous chunk's A1/A2 readers are done (and basis/scale ready)
    // V chunk -> A1: real parts in rows yy, imaginary parts in rows 16 + yy;
    // one 32-bit (fp16 pair) or 64-bit (float pair) write per part and plane
    auto store_a1 = [&](int e2, float4 v) {
      const int e = 2 * e2;
      const int yy = e / w, x = e - yy * w;
      if (SPLIT) {
        unsigned rh, rl, ih, il;
        split_pair(v.x * scale, v.z * scale, rh, rl);
        split_pair(v.y * scale, v.w * scale, ih, il);
        *reinterpret_cast<unsigned*>(A1 + yy * pw + x) = rh;
        *reinterpret_cast<unsigned*>(A1 + 32 * pw + yy * pw + x) = rl;
        *reinterpret_cast<unsigned*>(A1 + (16 + yy) * pw + x) = ih;
        *reinterpret_cast<unsigned*>(A1 + 32 * pw + (16 + yy) * pw + x) = il;
      } else {
        *reinterpret_cast<float2*>(A1 + yy * pw + x) = make_float2(v.x, v.z);
        *reinterpret_cast<float2*>(A1 + (16 + yy) * pw + x) = make_float2(v.y, v.w);
      }
    };
    if (ahead) {
#pragma unroll
      for (int j = 0; j < kPerT; ++j) {
        if (j < nPer) {
          const int e2 = tid + NT * j;
          const float4 v = nxt[j];
          if (y0 + kChunk < h) nxt[j] = V4[(size_t)(y0 + kChunk) * (w / 2) + e2];
          store_a1(e2, v);
        }
      }
    } else {
      for (int e2 = tid; e2 < kChunk * w / 2; e2 += NT) store_a1(e2, V4[(size_t)y0 * (w / 2) + e2]);
    }
    __syncthreads();
    // stage 1: tiles (rt in 0..1) x (ct in 0..Gp/16)
    for (int t = wave; t < rt1 * ct1; t += NW) {
      const int rt = t / ct1, ct = t - rt * ct1;
      floatx4 d = {0.f, 0.f, 0.f, 0.f};
      const E* arow = A1 + (rt * 16 + lr) * pw;
      const E* brow = Bx + (ct * 16 + lr) * pw;
      if (SPLIT) {
        for (int k0 = 0; k0 < w; k0 += 32) {
          // k-slots: 8 consecutive k per lane group; a ragged last step (w % 32 == 16)
          // zero-fills the upper lane groups
          half8 ah, al, bh, bl;
          const int kb = k0 + 8 * q;
          if (kb < w) {   // 16-byte aligned: pitches, plane offsets and kb are multiples of 8
            ah = *reinterpret_cast<const half8*>(arow + kb);
            al = *reinterpret_cast<const half8*>(arow + 32 * pw + kb);
            bh = *reinterpret_cast<const half8*>(brow + kb);
            bl = *reinterpret_cast<const half8*>(brow + Gp * pw + kb);
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) ah[j] = al[j] = bh[j] = bl[j] = (_Float16)0.f;
          }
          d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, d, 0, 0, 0);
        }
      } else {
        // lane group q takes the k-slots q w/4 ... (q + 1) w/4 - 1 (the same for A and B),
        // four per 16-byte load
        const int kq = w >> 2;
        const float* ar = reinterpret_cast<const float*>(arow) + q * kq;
        const float* br = reinterpret_cast<const float*>(brow) + q * kq;
        for (int s4 = 0; s4 < kq; s4 += 4) {
          const float4 a4 = *reinterpret_cast<const float4*>(ar + s4);
          const float4 b4 = *reinterpret_cast<const float4*>(br + s4);
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, b4.x, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, b4.y, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, b4.z, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, b4.w, d, 0, 0, 0);
        }
      }
      // D[yy][m] of part c = rt, yy = 4q + i  ->  A2[(m, c)][4q .. 4q + 3], one write
      const int m = ct * 16 + lr;
      const int row2 = 2 * m + rt;
      if (SPLIT) {
        uint2 th, tl;
        split_pair(d[0], d[1], th.x, tl.x);
        split_pair(d[2], d[3], th.y, tl.y);
        *reinterpret_cast<uint2*>(A2 + row2 * pk + 4 * q) = th;
        *reinterpret_cast<uint2*>(A2 + 2 * Gp * pk + row2 * pk + 4 * q) = tl;
      } else {
        *reinterpret_cast<float4*>(A2 + row2 * pk + 4 * q) = make_float4(d[0], d[1], d[2], d[3]);
      }
    }
    __syncthreads();
    // stage 2: acc[(m,c)][n] += sum_yy A2[(m,c)][yy] HGy[n][y0+yy]
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      if (i < nMine) {
        const int t = myTile[i];
        const int mt = t / ct2, nt = t - mt * ct2;
        const E* arow = A2 + (mt * 16 + lr) * pk;
        const E* brow = By + (nt * 16 + lr) * ph + y0;
        if (SPLIT) {
          // K = 16 = the chunk's rows: lane group q carries k-slots 4q .. 4q + 3
          const half4 ah = *reinterpret_cast<const half4*>(arow + 4 * q);
          const half4 al = *reinterpret_cast<const half4*>(arow + 2 * Gp * pk + 4 * q);
          const half4 bh = *reinterpret_cast<const half4*>(brow + 4 * q);
          const half4 bl = *reinterpret_cast<const half4*>(brow + Gp * ph + 4 * q);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, acc[i], 0, 0, 0);
        } else {
          // lane group q takes the k-slots 4q .. 4q + 3 of the chunk
          const float4 a4 = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(arow) + 4 * q);
          const float4 b4 = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(brow) + 4 * q);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, b4.x, acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, b4.y, acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, b4.z, acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, b4.w, acc[i], 0, 0, 0);
        }
      }
    }
  }

  // --- write the triangle m + n < G ---
  const float inv = 1.f / scale;
  float2* out = coefs + (size_t)(b * g.P + p) * M;
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    if (i < nMine) {
      const int t = myTile[i];
      const int mt = t / ct2, nt = t - mt * ct2;
      const int n = nt * 16 + lr;
#pragma unroll
      for (int ii = 0; ii < 4; ii += 2) {
        const int m = (mt * 16 + 4 * q + ii) >> 1;      // rows ii, ii+1 are (m, re), (m, im)
        if (n < G && m < G && m + n < G) {
          const int gsum = m + n;
          out[gsum * (gsum + 1) / 2 + n] = make_float2(acc[i][ii] * inv, acc[i][ii + 1] * inv);
        }
      }
    }
  }
  }   // frames of this slot
}

// float4 pairs of a field chunk (8 w of them, w <= 32 KS1) per thread of a
// workgroup of NT2 waves: at most 8
// halfs per row of the split overlap's fp16 chunk planes: the field width or
// the k-steps' width, whichever is larger, + 8 (rows 16 bytes apart mod 256)
__host__ __device__ constexpr int split_row_halfs(int w, int ks1) { return (w > 32 * ks1 ? w : 32 * ks1) + 8; }

__host__ __device__ constexpr int split_pairs_per_thread(int ks1, int nt2) {
  return (4 * ks1 + nt2 - 1) / nt2 < 8 ? (4 * ks1 + nt2 - 1) / nt2 : 8;
}

// Split-fp16 overlap with its basis operands in registers (fields of w <= 160,
// G <= 96): one wave per m-tile of 16 modes, both parts (re, im) of the field on
// the same wave, so the wave's HGx operand (KS1 k-steps of hi and lo) is loaded
// once and serves every chunk of the field and both parts.  Per 16-row chunk the
// field goes to LDS once (double-buffered, row pitch w + 2 complex: 16 rows of one
// 16-byte column fall on distinct bank groups), every wave splits its A operand
// into fp16 hi/lo as it reads it, and stage 2 runs from the stage-1 accumulators
// (their C layout is the A layout of v_mfma_f32_16x16x16_f16) against HGy
// fragments read from L2: no exchange through LDS, one barrier per chunk.  One
// (frame, pol) per workgroup.
// Workgroups per CU the register budget is bounded for: three (168 VGPRs) where
// the wave's basis, accumulators and prefetched chunk fit without spilling
// (8 KS1 + 8 NT2 + 4 pairs <= 80 registers, and fewer than six m-tiles; checked
// against the compiler's spill counts), else two.
__host__ __device__ constexpr int split_min_blocks(int ks1, int nt2) {
  return 8 * ks1 + 8 * nt2 + 4 * split_pairs_per_thread(ks1, nt2) <= 80 && nt2 < 6 ? 3 : 2;
}

template <int KS1, int NT2>
__global__ __launch_bounds__(64 * NT2, split_min_blocks(KS1, NT2)) void k_overlap_split(Geometry g, const float2* __restrict__ fields,
                                                            const _Float16* __restrict__ hgx16,
                                                            const _Float16* __restrict__ hgy16,
                                                            float2* __restrict__ coefs,
                                                            const unsigned* __restrict__ peak) {
  extern __shared__ __align__(16) unsigned char os_smem[];
  __shared__ float s_red[NT2];
  constexpr int NT = 64 * NT2;
  const int p = blockIdx.x, b = blockIdx.y + g.b0;
  const int G = g.G, Gp = (G + 15) & ~15, w = g.w, h = g.h;
  const int tid = threadIdx.x, lane = tid & 63;
  const int mt = __builtin_amdgcn_readfirstlane(tid >> 6);   // this wave's m-tile
  const int lr = lane & 15, q = lane >> 4;
  const _Float16* bx = hgx16 + (size_t)p * 2 * Gp * w;
  const _Float16* by = hgy16 + (size_t)p * 2 * Gp * h;

  // stage-1 basis operand: HGx rows 16 mt + lr, k-slots 32 ks + 8 q .. + 7
  half8 b1h[KS1], b1l[KS1];
#pragma unroll
  for (int ks = 0; ks < KS1; ++ks) {
    const int kb = 32 * ks + 8 * q;
    if (kb < w) {
      const _Float16* bp = bx + (16 * mt + lr) * w + kb;
      b1h[ks] = *reinterpret_cast<const half8*>(bp);
      b1l[ks] = *reinterpret_cast<const half8*>(bp + Gp * w);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) b1h[ks][e] = b1l[ks][e] = (_Float16)0.f;
    }
  }

  // complex pairs as native 4-vectors (a struct-typed float4 copy would keep the
  // prefetch registers in scratch)
  const floatx4* V4 = reinterpret_cast<const floatx4*>(fields + (size_t)(b * g.P + p) * h * w);
  // per-field power of two: the peak magnitude into [128, 256)
  float mx;
  if (peak) {   // left by the kernel that wrote the field
    mx = __uint_as_float(peak[b * g.P + p]);
  } else {
    mx = 0.f;
    for (int e = tid; e < h * w / 2; e += NT) {
      const floatx4 v = V4[e];
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) s_red[mt] = mx;
    __syncthreads();
    mx = s_red[0];
#pragma unroll
    for (int i = 1; i < NT2; ++i) mx = fmaxf(mx, s_red[i]);
  }
  float scale = 1.f;
  if (mx > 0.f) {
    int ex;
    frexpf(mx, &ex);
    scale = ldexpf(1.f, 8 - ex);
  }

  // Each chunk (16 field rows) is split once into fp16 hi / lo planes [re hi,
  // re lo, im hi, im lo][16 rows][SP halfs] in LDS (double-buffered; columns
  // w .. 32 KS1 - 1 zero), so every wave reads its MFMA operands with one
  // 16-byte load per plane.  This thread's complex pairs e2 = tid + NT j (j <
  // kPer, which the launcher checks covers the chunk) and their LDS offsets
  // are the same for every chunk; the next chunk's pairs are held in registers.
  constexpr int kPer = split_pairs_per_thread(KS1, NT2);
  const int n4 = 8 * w, J = h / 16;
  const int SP = split_row_halfs(w, KS1), PS = 16 * SP;
  _Float16* pl = reinterpret_cast<_Float16*>(os_smem);   // [2][4][16][SP]
  for (int e = tid; e < 2 * 4 * 16 * (32 * KS1 - w); e += NT) {
    const int c = 32 * KS1 - w, row = e / c;
    pl[row * SP + w + (e - row * c)] = (_Float16)0.f;
  }
  int lofs[kPer], gofs[kPer];
  floatx4 nxt[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e2 = tid + NT * j, yy = (2 * e2) / w;
    lofs[j] = e2 < n4 ? yy * SP + 2 * e2 - yy * w : -1;
    gofs[j] = e2 < n4 ? e2 : n4 - 1;
    nxt[j] = V4[gofs[j]];   // chunk 0
  }
  auto stage = [&](int bi) {   // nxt (scaled, split) into buffer bi
    _Float16* pb = pl + bi * 4 * PS;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (lofs[j] >= 0) {
        unsigned rh, rl, ih, il;
        split_pair(nxt[j][0] * scale, nxt[j][2] * scale, rh, rl);   // re of columns x, x + 1
        split_pair(nxt[j][1] * scale, nxt[j][3] * scale, ih, il);   // im
        *reinterpret_cast<unsigned*>(pb + lofs[j]) = rh;
        *reinterpret_cast<unsigned*>(pb + PS + lofs[j]) = rl;
        *reinterpret_cast<unsigned*>(pb + 2 * PS + lofs[j]) = ih;
        *reinterpret_cast<unsigned*>(pb + 3 * PS + lofs[j]) = il;
      }
    }
  };
  stage(0);
#pragma unroll
  for (int j = 0; j < kPer; ++j) nxt[j] = V4[(J > 1 ? n4 : 0) + gofs[j]];

  floatx4 acc0[NT2], acc1[NT2];
#pragma unroll
  for (int nt = 0; nt < NT2; ++nt) acc0[nt] = acc1[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  // stage-2 basis of chunk ch, HGy[16 nt + lr][16 ch + 4 q .. + 3] (rows >= G
  // zero), one chunk ahead of its use
  half4 b2nh[NT2], b2nl[NT2];
  auto load_b2 = [&](int ch) {
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
      if (16 * mt + 16 * nt < G) {
        const _Float16* brow = by + (16 * nt + lr) * h + 16 * ch + 4 * q;
        b2nh[nt] = *reinterpret_cast<const half4*>(brow);
        b2nl[nt] = *reinterpret_cast<const half4*>(brow + Gp * h);
      }
    }
  };
  load_b2(0);
  for (int ch = 0; ch < J; ++ch) {
    __syncthreads();   // chunk ch is in its buffer; the other buffer's readers are done
    const _Float16* cb = pl + (ch & 1) * 4 * PS + lr * SP + 8 * q;
    // stage 1, both parts: T_c[yy][m] = sum_x V_c[yy][x] HGx[m][x]
    floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks) {
      const half8 ahr = *reinterpret_cast<const half8*>(cb + 32 * ks);
      const half8 alr = *reinterpret_cast<const half8*>(cb + PS + 32 * ks);
      const half8 ahi = *reinterpret_cast<const half8*>(cb + 2 * PS + 32 * ks);
      const half8 ali = *reinterpret_cast<const half8*>(cb + 3 * PS + 32 * ks);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahr, b1h[ks], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, b1h[ks], d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahr, b1l[ks], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, b1l[ks], d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(alr, b1h[ks], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ali, b1h[ks], d1, 0, 0, 0);
    }
    // stage-2 basis of this chunk (loaded a chunk ago); the next chunk's from L2
    half4 b2h[NT2], b2l[NT2];
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
      b2h[nt] = b2nh[nt];
      b2l[nt] = b2nl[nt];
    }
    if (ch + 1 < J) load_b2(ch + 1);
    // the next chunk into the other buffer (loaded a chunk ago); the one after in flight
    if (ch + 1 < J) {
      stage((ch + 1) & 1);
      const size_t c2 = (size_t)(ch + 2 < J ? ch + 2 : ch + 1) * n4;
#pragma unroll
      for (int j = 0; j < kPer; ++j) nxt[j] = V4[c2 + gofs[j]];
    }
    // stage 2 from registers: C_c[m][n] += sum_yy T_c[m][yy] HGy[n][yy]
    uint2 t0h, t0l, t1h, t1l;
    split_pair(d0[0], d0[1], t0h.x, t0l.x);
    split_pair(d0[2], d0[3], t0h.y, t0l.y);
    split_pair(d1[0], d1[1], t1h.x, t1l.x);
    split_pair(d1[2], d1[3], t1h.y, t1l.y);
    const half4 a0h = __builtin_bit_cast(half4, t0h), a0l = __builtin_bit_cast(half4, t0l);
    const half4 a1h = __builtin_bit_cast(half4, t1h), a1l = __builtin_bit_cast(half4, t1l);
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
      if (16 * mt + 16 * nt < G) {
        acc0[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a0h, b2h[nt], acc0[nt], 0, 0, 0);
        acc1[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a1h, b2h[nt], acc1[nt], 0, 0, 0);
        acc0[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a0h, b2l[nt], acc0[nt], 0, 0, 0);
        acc1[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a1h, b2l[nt], acc1[nt], 0, 0, 0);
        acc0[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a0l, b2h[nt], acc0[nt], 0, 0, 0);
        acc1[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(a1l, b2h[nt], acc1[nt], 0, 0, 0);
      }
    }
  }
  // the triangle m + n < G; lane (n, q) holds m = 16 mt + 4 q .. + 3
  const float inv = 1.f / scale;
  float2* out = coefs + (size_t)(b * g.P + p) * g.M;
#pragma unroll
  for (int nt = 0; nt < NT2; ++nt) {
    const int n = 16 * nt + lr;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int m = 16 * mt + 4 * q + e;
      if (m + n < G) {
        const int gs = m + n;
        out[gs * (gs + 1) / 2 + n] = make_float2(acc0[nt][e] * inv, acc1[nt][e] * inv);
      }
    }
  }
}

template <int KS1, int NT2>
static hipError_t split_launch(const Geometry& g, const float2* fields, const void* hgx16, const void* hgy16,
                               float2* coefs, hipStream_t s, const unsigned* peak) {
  if (8 * g.w > split_pairs_per_thread(KS1, NT2) * 64 * NT2) return hipErrorNotSupported;
  const size_t lds = (size_t)2 * 4 * 16 * split_row_halfs(g.w, KS1) * sizeof(_Float16);
  return for_frame_chunks(g.B, [&](int b0, int nb) {
    Geometry gc = g;
    gc.b0 = b0;
    k_overlap_split<KS1, NT2><<<dim3(g.P, nb), 64 * NT2, lds, s>>>(gc, fields, static_cast<const _Float16*>(hgx16),
                                                                   static_cast<const _Float16*>(hgy16), coefs, peak);
    return hipGetLastError();
  });
}

template <int KS1>
static hipError_t split_dispatch_nt(const Geometry& g, const float2* fields, const void* hgx16, const void* hgy16,
                                    float2* coefs, hipStream_t s, const unsigned* peak) {
  switch ((g.G + 15) / 16) {
    case 1: return split_launch<KS1, 1>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 2: return split_launch<KS1, 2>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 3: return split_launch<KS1, 3>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 4: return split_launch<KS1, 4>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 5: return split_launch<KS1, 5>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 6: return split_launch<KS1, 6>(g, fields, hgx16, hgy16, coefs, s, peak);
    default: return hipErrorNotSupported;
  }
}

bool overlap_split_supported(const Geometry& g) {
  return g.G > 0 && g.G <= 96 && g.w % 16 == 0 && g.h % 16 == 0 && g.w >= 16 && g.w <= 160 && g.h >= 16;
}

hipError_t launch_overlap_split(const Geometry& g, const float2* fields, const void* hgx16, const void* hgy16,
                                float2* coefs, hipStream_t s, const unsigned* peak) {
  if (!overlap_split_supported(g) || !hgx16 || !hgy16) return hipErrorNotSupported;
  switch ((g.w + 31) / 32) {
    case 1: return split_dispatch_nt<1>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 2: return split_dispatch_nt<2>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 3: return split_dispatch_nt<3>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 4: return split_dispatch_nt<4>(g, fields, hgx16, hgy16, coefs, s, peak);
    case 5: return split_dispatch_nt<5>(g, fields, hgx16, hgy16, coefs, s, peak);
    default: return hipErrorNotSupported;
  }
}

static size_t overlap_lds(const Geometry& g, bool split) {
  const int Gp = (g.G + 15) & ~15;
  const size_t e = split ? 2 : 4;
  const int nL = split ? 2 : 1;
  return e * nL * ((size_t)Gp * pitch_of(g.w) + (size_t)Gp * pitch_of(g.h) + 32 * (size_t)pitch_of(g.w) +
                   2 * (size_t)Gp * kPitch2);
}

// waves per workgroup (8 measured faster than 4 on both the 80-point and the
// 160-point fields: more waves per SIMD to cover the LDS and barrier latency)
constexpr int kOvWaves = 8;

static int tiles_per_wave(const Geometry& g, int nw) {
  const int Gp = (g.G + 15) & ~15;
  const int rt2 = (2 * Gp) / 16, ct2 = Gp / 16;
  int cnt = 0;
  for (int t = 0; t < rt2 * ct2; ++t) {
    const int mt = t / ct2, nt = t - mt * ct2;
    if (8 * mt + 16 * nt < g.G) ++cnt;
  }
  return (cnt + nw - 1) / nw;
}

template <bool SPLIT, int TPW, int NW>
static hipError_t ov_launch(const Geometry& g, const float2* fields, const float* hgx, const float* hgy,
                            float2* coefs, hipStream_t s, const unsigned* peak) {
  const size_t lds = overlap_lds(g, SPLIT);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_overlap_mfma<SPLIT, TPW, NW>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  // as many workgroups as stay resident (two per CU on 256 CUs, one when the LDS
  // holds only one), each walking several frames with the basis staged once
  const int perCU = lds > 80 * 1024 ? 1 : 2;
  const int slots = std::max(1, std::min(g.B, 256 * perCU / std::max(1, g.P)));
  dim3 grid(g.P, slots);
  k_overlap_mfma<SPLIT, TPW, NW><<<grid, 64 * NW, lds, s>>>(g, fields, hgx, hgy, coefs, peak);
  return hipGetLastError();
}

template <bool SPLIT, int NW>
static hipError_t ov_dispatch_nw(const Geometry& g, const float2* fields, const float* hgx, const float* hgy,
                                 float2* coefs, int tpwWanted, hipStream_t s, const unsigned* peak) {
  const int tpw = std::max(tiles_per_wave(g, NW), tpwWanted);
  if (tpw <= 2) return ov_launch<SPLIT, 2, NW>(g, fields, hgx, hgy, coefs, s, peak);
  if (tpw <= 4) return ov_launch<SPLIT, 4, NW>(g, fields, hgx, hgy, coefs, s, peak);
  if (tpw <= 8) return ov_launch<SPLIT, 8, NW>(g, fields, hgx, hgy, coefs, s, peak);
  if (tpw <= 16) return ov_launch<SPLIT, 16, NW>(g, fields, hgx, hgy, coefs, s, peak);
  if (tpw <= 32) return ov_launch<SPLIT, 32, NW>(g, fields, hgx, hgy, coefs, s, peak);
  return hipErrorInvalidValue;
}

template <bool SPLIT>
static hipError_t ov_dispatch(const Geometry& g, const float2* fields, const float* hgx, const float* hgy,
                              float2* coefs, int tpwWanted, hipStream_t s, const unsigned* peak) {
  return ov_dispatch_nw<SPLIT, kOvWaves>(g, fields, hgx, hgy, coefs, tpwWanted, s, peak);
}

int overlap_min_tiles_per_wave(const Geometry& g, int split) {
  (void)split;
  return tiles_per_wave(g, kOvWaves);
}

hipError_t launch_overlap_mfma(const Geometry& g, const float2* fields, const float* hgx, const float* hgy,
                               float2* coefs, int split, hipStream_t s, int tpw, const unsigned* peak,
                               const void* hgx16, const void* hgy16) {
  if (g.w % 16 || g.h % kChunk || g.G <= 0) return hipErrorInvalidValue;
  if (split && hgx16 && hgy16) {
    const hipError_t e = launch_overlap_split(g, fields, hgx16, hgy16, coefs, s, peak);
    if (e != hipErrorNotSupported) return e;
  }
  if (overlap_lds(g, split != 0) > 160 * 1024) return hipErrorInvalidValue;
  return split ? ov_dispatch<true>(g, fields, hgx, hgy, coefs, tpw, s, peak)
               : ov_dispatch<false>(g, fields, hgx, hgy, coefs, tpw, s, nullptr);
}

}  // namespace dh
