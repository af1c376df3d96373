// Shared declarations between the C API host code (dh_api.cpp) and the HIP
// kernels (dh_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dh {

constexpr int kMaxPol = 2;
constexpr int kMaxStages = 12;
// Frames of one launch along grid y or z (the grid extent there is limited to
// 65535): kernels that index frames by blockIdx.y / blockIdx.z add
// Geometry::b0, and their launchers cover a batch in chunks of this many.
constexpr int kMaxGridFrames = 65535;
// Frames of one batch (averaged frames included).
constexpr int kMaxBatchFrames = 1 << 24;

// Bytes of the per-GPU scratch array of the allocation-free batch entry point.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

// One 1-D FFT length factored for the Stockham kernels.
struct FFTPlan {
  int n;                      // transform length (complex points)
  int nstages;
  int radix[kMaxStages];
  // q / d == __umulhi(q, mag) for q < 2^20 (mag = ceil(2^32 / d); 0 when d == 1)
  unsigned magNbf[kMaxStages];  // d = n / radix[s]
  unsigned magNs[kMaxStages];   // d = radix[0] * ... * radix[s-1]
  const float2* tw;           // n twiddles exp(-2 pi i k / n), device memory
};

__host__ __device__ inline unsigned magic_div(unsigned d) {
  return d <= 1 ? 0u : (unsigned)((0xFFFFFFFFull / d) + 1ull);
}

// Everything the kernels need about the windows of one batch.
struct Geometry {
  int B, P;                   // batch, polarisations
  int frameW, frameH;
  int W, H;                   // FFT window
  int w, h;                   // IFFT window (the Fourier-plane window that is inverse transformed)
  int fw, fh;                 // reconstructed field plane: w x h (low-res IFFT), W x H (full resolution)
  int cx[kMaxPol], cy[kMaxPol];   // FFT window corners in the frame
  int c0[kMaxPol];            // first r2c column read by the IFFT window (SX - W/2)
  int SY[kMaxPol];            // IFFT window row start (shifted-spectrum coordinates)
  int G, M;                   // HG groups and modes per pol
  // wavelength sweep: L wavelengths; frame and output ordering, 0 wavelength
  // the fast axis, 1 the slow one; subB = B / L frames per wavelength
  int L, ordIn, ordOut, subB;
  int b0;                     // first frame of a launch covering part of the batch (grid y / z)
};

// Workgroups for an elementwise kernel that strides over n elements: one
// element per thread up to 2^22 workgroups, several per thread beyond (a
// launch's work-item count in one dimension must stay below 2^32).
inline unsigned stride_blocks(size_t n, int threads) {
  const size_t b = (n + threads - 1) / threads;
  return (unsigned)(b < ((size_t)1 << 22) ? b : ((size_t)1 << 22));
}

// f(b0, nb) for consecutive chunks of at most kMaxGridFrames frames of n.
template <class F>
inline hipError_t for_frame_chunks(int n, F&& f) {
  for (int b0 = 0; b0 < n; b0 += kMaxGridFrames) {
    const hipError_t e = f(b0, n - b0 < kMaxGridFrames ? n - b0 : kMaxGridFrames);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Wavelength index of input frame b, the batch position its field and
// coefficients take, and the wavelength index of output position o
// (digHolo.cpp:2656-2700).  One wavelength: index 0, position b.  The host
// admits a slow ordering only when L divides B, so positions form a permutation.
__host__ __device__ __forceinline__ int lam_in(const Geometry& g, int b) {
  if (g.L <= 1) return 0;
  return g.ordIn ? b / g.subB : b % g.L;
}
__host__ __device__ __forceinline__ int out_pos(const Geometry& g, int b) {
  if (g.L <= 1) return b;
  int lam, sub;
  if (g.ordIn) {
    sub = b % g.subB;
    lam = b / g.subB;
  } else {
    lam = b % g.L;
    sub = b / g.L;
  }
  return g.ordOut ? lam * g.subB + sub : sub * g.L + lam;
}
__host__ __device__ __forceinline__ int lam_out(const Geometry& g, int o) {
  if (g.L <= 1) return 0;
  return g.ordOut ? o / g.subB : o % g.L;
}

// Per-frame complex calibration of the Fourier window (digHolo.cpp:4469-4500,
// :1831-1840): cal[p % P][b % B]; p == nullptr means none.
struct BatchCal {
  const float2* p = nullptr;
  int P = 1, B = 1;
  __host__ __device__ float2 at(int pol, int b) const {
    return p ? p[(pol % P) * B + (b % B)] : make_float2(1.f, 0.f);
  }
};

// FP64 per-pol parameters of the Fourier-window factor and the reference wave.
struct PolParams {
  double dkx, dky;            // 2 pi / (px W), 2 pi / (px H)
  double bcx, bcy;            // beam centre relative to the FFT window centre (m)
  double px;                  // camera pixel (m)
  double norm;                // 1/sqrt(W H w h)
  int fill;                   // fill-factor correction on/off
  // reference wave: phi(x) = kx x + piston_x + focal x^2 (same for y)
  double kxr, kyr, pistx, pisty, focal;
  double pxx, pxy;            // reconstructed-field pixel pitch
  double waist;               // HG waist
};

struct TableArgs {
  Geometry g;
  PolParams pp[kMaxPol];
};

// Launchers (defined in dh_kernels.hip).  All enqueue on `s`.
// One wavelength's window factor and reference wave; d_hgx == nullptr skips the HG basis.
hipError_t launch_tables(const TableArgs& a, const uint8_t* d_mask, float2* d_factor, float2* d_ref,
                         float* d_hgx, float* d_hgy, hipStream_t s);
// The HG basis alone ([P][G][fw] and [P][G][fh]), as launch_tables builds it.
hipError_t launch_hg_basis(const TableArgs& a, float* d_hgx, float* d_hgy, hipStream_t s);
hipError_t launch_row_fft(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                          int frame_is_u16, float2* rowspec, hipStream_t s);
hipError_t launch_col_fft_window(const Geometry& g, const FFTPlan& pl, const float2* rowspec,
                                 const float2* factor, float2* win, const BatchCal& cal, hipStream_t s);
hipError_t launch_ifft_rows(const Geometry& g, const FFTPlan& pl, float2* win, hipStream_t s);
// Staged kernels on in-place transforms (radix 2/3/4/5/8 plans only): row
// FFT, then column FFT + window + inverse FFT along y, then inverse FFT along x
// + reference wave.  staged_ip_supported says whether the three apply.
bool staged_ip_supported(const Geometry& g, const FFTPlan& row, const FFTPlan& col, const FFTPlan& ix,
                         const FFTPlan& iy);
// 128- and 256-point rows (256- and 512-wide FFT windows) in registers; the
// staged row kernel's output.  hipErrorNotSupported for other lengths.
hipError_t launch_row_fft_reg(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                              int frame_is_u16, float2* rowspec, hipStream_t s);
// Column stage of 256- and 512-row FFT windows with the column transform in
// registers (h <= 256); hipErrorNotSupported for other shapes.
hipError_t launch_col_reg_window_iy(const Geometry& g, const FFTPlan& pc, const FFTPlan& py, const float2* rowspec,
                                   const float2* factor, float2* win, const BatchCal& cal, hipStream_t s,
                                   unsigned* peak);
// Inverse transform along x in registers for w = 48, 64, 80 or 160, then the
// reference wave; hipErrorNotSupported for other widths.
inline bool ifft_rows_reg_width(int w) { return w == 48 || w == 64 || w == 80 || w == 160; }
hipError_t launch_ifft_rows_reg(const Geometry& g, const FFTPlan& pl, const float2* win, const float2* ref,
                                float2* fields, hipStream_t s, unsigned* peak);
hipError_t launch_row_fft_ip(const Geometry& g, const FFTPlan& pl, const float2* twW, const void* frames,
                             int frame_is_u16, float2* rowspec, hipStream_t s);
// peak (optional, [B][P] float bits): the column kernel clears it, the row
// kernel leaves there the largest |Re|, |Im| of each output field, for the
// split overlap's scale.
hipError_t launch_col_window_iy(const Geometry& g, const FFTPlan& pc, const FFTPlan& py, const float2* rowspec,
                                const float2* factor, float2* win, const BatchCal& cal, hipStream_t s,
                                unsigned* peak = nullptr);
hipError_t launch_ifft_rows_tilt(const Geometry& g, const FFTPlan& pl, const float2* win, const float2* ref,
                                 float2* fields, hipStream_t s, unsigned* peak = nullptr);
// ref == nullptr: the inverse transform only (field before tilt removal)
hipError_t launch_ifft_cols_tilt(const Geometry& g, const FFTPlan& pl, const float2* win,
                                 const float2* ref, float2* fields, hipStream_t s);
// Step kernels: window rows of a column-transformed band times the window
// factor; tilt removal fields = -conj(fields * ref) in place.
hipError_t launch_select_window(const Geometry& g, const float2* spec, const float2* factor, float2* win,
                                const BatchCal& cal, hipStream_t s);
// Full-resolution IFFT input: win [B][P][h][w] placed at its own frequencies
// (row (j + SY + H/2) mod H, column c0 + i) of a zero [B][P][H][W] plane.
hipError_t launch_scatter_full(const Geometry& g, const float2* win, float2* plane, hipStream_t s);
// uint16 frames given transposed: out[f][i H + j] = in[f][j W + i] (W, H the frame dims)
hipError_t launch_u16_transpose(const void* in, void* out, int W, int H, size_t frames, hipStream_t s);
// In-phase averaging of A windows per output (avgMode 0 sequential, 1 interlaced,
// 2 sequential sweep with one wavelength): winAll [frames][P][h][w] -> win [g.B][P][h][w]
hipError_t launch_average_windows(const Geometry& g, int A, int mode, int frames, const float2* winAll, float2* win,
                                  const BatchCal& cal, hipStream_t s);
hipError_t launch_apply_tilt(const Geometry& g, const float2* ref, float2* fields, hipStream_t s);
// Fused window path: frame ingest -> r2c rows -> Fourier window (DIT over row
// residues) -> window factor -> 2-D IFFT -> tilt -> fields.  Returns
// hipErrorNotSupported when the shape does not fit its LDS/register budget
// (the caller then uses the staged kernels above).
struct FusedArgs {
  Geometry g;
  FFTPlan row, col, ix, iy;   // W/2, H/R, w, h
  int R, Hc;                  // row residues, points per residue (H = R * Hc)
  // set by fused_prepare: rows per row batch, and the most transforms of each
  // plan one in-place pass may hold
  int RB, chunkCol, chunkIx, chunkIy;
};
// Fills the batch/chunk sizes; false if the shape does not fit the kernel.
bool fused_prepare(FusedArgs& a);
hipError_t launch_fused_window(const FusedArgs& a, const float2* twW, const float2* twH, const void* frames,
                               int frame_is_u16, const float2* factor, const float2* ref, float2* fields,
                               hipStream_t s);

// Fused window path with the column transform as a split-fp16 GEMM on MFMA
// (uint16 frames).  dtab: the column-DFT operands built by launch_coldft_table
// (coldft_table_elems(g) halfs) for the same geometry.
bool window_mfma_prepare(FusedArgs& a);
size_t coldft_table_elems(const Geometry& g);
hipError_t launch_coldft_table(const Geometry& g, void* tab, hipStream_t s);
hipError_t launch_window_mfma(const FusedArgs& a, const float2* twW, const void* frames_u16, const void* dtab,
                              const float2* factor, const float2* ref, float2* fields, hipStream_t s);

// Size-specialised fused window path (uint16 frames, W/2 in {64, 96, 128},
// w and h multiples of 16 from a fixed list): register row transforms and the
// column transform on MFMA with D built from an LDS twiddle table.  tab holds
// window_v3_table_elems(g) complex values built by launch_window_v3_tables.
bool window_v3_supported(const FusedArgs& a);
size_t window_v3_table_elems(const Geometry& g);
hipError_t launch_window_v3_tables(const Geometry& g, float2* tab, hipStream_t s);
// With hgx16/hgy16 (split basis from launch_hg_split) and coefs set, the same
// launch also computes the split-fp16 HG overlap into coefs; with hgx32/hgy32
// (the FP32 basis) instead, the FP32 overlap.
// rowMode: 0 the MFMA row transform where it exists (256-wide windows), 1 the
// register one (window_v3_row_choice: whether the geometry has the choice).
hipError_t launch_window_v3(const FusedArgs& a, const void* frames_u16, const float2* tab, const float2* factor,
                            const float2* ref, float2* fields, const void* hgx16, const void* hgy16, float2* coefs,
                            const BatchCal& cal, hipStream_t s, const float* hgx32 = nullptr,
                            const float* hgy32 = nullptr, int rowMode = 0);
bool window_v3_row_choice(const Geometry& g);
bool window_v3_overlap_fits(const Geometry& g);
// One batch of a multi-batch launch of the size-specialised window kernel: the
// arguments launch_window_v3 takes.
constexpr int kW3MaxBatches = 4;
struct W3Launch {
  FusedArgs fa;
  const void* frames;
  const float2 *tab, *factor, *ref;
  float2* fields;
  const void *hgx16, *hgy16;
  float2* coefs;
  BatchCal cal;
  const float *hgx32, *hgy32;
  int rowMode;   // launch_window_v3's row transform choice
};
// n <= kW3MaxBatches batches with square FFT windows in one launch (batch =
// blockIdx.y); window_v3_multi_supported says whether a batch's shape is one
// the multi-batch kernel holds.
bool window_v3_multi_supported(const W3Launch& b);
hipError_t launch_window_v3_multi(const W3Launch* batches, int n, hipStream_t s);
// Split-fp16 basis [P][hi/lo][Gp][n] (Gp = G rounded up to 16) from hgx/hgy.
size_t hg_split_elems(const Geometry& g, int n);
hipError_t launch_hg_split(const Geometry& g, const float* hgx, const float* hgy, void* hgx16, void* hgy16,
                           hipStream_t s);

// Power-meter reductions (dh_metrics.hip).  out: [S][N][5] doubles per row
// {diagonal, off-diagonal, in-group/g, out-of-group/g, total power}.
hipError_t launch_metric_rows(const float2* coefs, int N, int M, int Mpol, int rowStride, int S, double* out,
                              hipStream_t st);
// FP64 Gram of matrix s: rows as vectors (N x N) or columns as vectors (M x M).
hipError_t launch_gram_f64(const float2* coefs, int N, int M, int rowStride, int s, int rowsAreVectors,
                           double2* G, hipStream_t st);
// MDL seed of the alignment gradient for one extreme singular triplet of
// matrix s: x is the Gram eigenvector (length N when rowsAreVectors, else M),
// y receives its image (A^H x or A x, over sig; length M or N), and
// g[b][s M + j] += coef conj(u[b]) v[j] with (u, v) = (x, y) or (y, x).
hipError_t launch_mdl_seed_pair(const float2* coefs, int N, int M, int rowStride, int s, int rowsAreVectors,
                                const double2* x, double sig, double coef, double2* y, float2* g, hipStream_t st);

// Plane analysis (dh_analysis.hip).  Statistics of Nx x Ny planes: columns
// [x0, x1) of rows other than skip0/skip1 are visited; include (Ny x Nx bytes,
// or null for all) masks the complex input; I = norm |V|^2.
struct PlaneArgs {
  int Nx, Ny, inStride;        // plane size, input row stride (elements)
  int x0, x1, skip0, skip1;    // visited columns, rows left out (-1: none)
  float norm;
  const float* x;              // [Nx] axis
  const float* y;              // [Ny] axis
  const float* yw;             // [Ny] axis with negative values wrapped
  const uint8_t* include;
};
// out [planes][8] = {sum I, sum x I, sum y I, sum yw I, sum I^2, peak, peak index, 0};
// pw (optional) [planes][Ny][Nx] receives I (0 where not visited or masked).
// Exactly one of inC (complex) / inR (intensity) is non-null.
hipError_t launch_plane_stats(const PlaneArgs& a, const float2* inC, const float* inR, int planes, size_t planeStride,
                              float* pw, double* out, hipStream_t s);
// sum[p * polStride + e] = sum_b pw[b][p][e]
hipError_t launch_batch_sum(const float* pw, int B, int P, int N, float* sum, size_t polStride, hipStream_t s);
// int16 export: re/im [B][P][hw] = round(scale[p][b] * V), scale [P][B + 1]
hipError_t launch_fields16(const float2* fields, int B, int P, int hw, const float* scale, short* re, short* im,
                           hipStream_t s);
// In-place FFT along y of all g.w columns of [B][P][H][g.w] R2C rows
hipError_t launch_col_fft_full(const Geometry& g, const FFTPlan& pl, float2* spec, hipStream_t s);

// AutoAlign parameter estimate (dh_align.hip): tilt search of one batch-summed
// R2C Fourier map [Ny][Nx] (float) -- DCT-II both axes, Gaussian filter with
// exponent (c px)^2 wx2 + (j px)^2 wy2, DCT-III both axes -- then the index of
// the largest filtered value with kx[i]^2 + ky[j]^2 > ad (-1: none).
// work: snap_work_elems(Ny, Nx) doubles.
size_t snap_work_elems(int Ny, int Nx);
hipError_t launch_snap_tilt_search(const float* map, int Ny, int Nx, float px, float wx2, float wy2, const float* kx,
                                   const float* ky, float ad, double* work, int* outIdx, hipStream_t s);

// Basis transforms and basis fields (dh_basis.hip).  LG: tab holds the
// per-group blocks U_g row-major one after another; out[r][g0 + k] =
// sum_c U_g[k][c] in[r][g0 + c] for rows r < rows.  Custom: out[r][j] =
// sum_{i < nIn} in[r * ldIn + i] T[i * ldT + j], j < nOut.  Basis fields:
// out[mode][p][y][x] = sum_i W[mode * ldW + i] HGx_{m_i}(x) HGy_{n_i}(y) on the
// g.w x g.h plane.
hipError_t launch_lg_apply(const float2* in, float2* out, size_t rows, int M, const float2* tab, hipStream_t s);
hipError_t launch_custom_apply(const float2* in, int ldIn, int rows, int nIn, const float2* T, int ldT, int nOut,
                               float2* out, hipStream_t s);
hipError_t launch_basis_synth(const Geometry& g, const float2* W, int modes, int ldW, const float* hgx,
                              const float* hgy, float2* out, hipStream_t s);

// Alignment-loss gradient (dh_grad.hip).  dhgx/dhgy: d/dwaist of the HG
// tables, same layout.  Back-projection per (frame, pol): partial[bp][0..5] =
// {sum Im(AV), sum x Im(AV), sum y Im(AV), sum r^2 Im(AV), sum Re(Aw V),
// sum |c|^2}, q1 = -V R, q2 = conj(A) R; pxy2[p] = {pxx, pxy}.  kspace:
// partial[bp][6..7] = sum_k {kx, ky} Im(X1 conj X2) with kx [P][fw], ky [P][fh].
hipError_t launch_hg_dwaist(const TableArgs& a, float* dhgx, float* dhgy, hipStream_t s);
size_t grad_backproject_lds(const Geometry& g);
// gExtra (nullable) [B][P][M] is added to the seed -conj(c)/B.
hipError_t launch_grad_backproject(const Geometry& g, double invB, const float2* coefs, const float2* gExtra,
                                   const float* hgx,
                                   const float* hgy, const float* dhgx, const float* dhgy, const float2* fields,
                                   const float2* ref, const double* pxy2, float2* q1, float2* q2, double* partial,
                                   hipStream_t s);
hipError_t launch_grad_kspace(const Geometry& g, const float2* x1, const float2* x2, const double* kx,
                              const double* ky, double* partial, hipStream_t s);
// out[p][k] = sum over frames of partial[b][p][k], k < 8
hipError_t launch_grad_reduce(int B, int P, const double* partial, double* out, hipStream_t s);

// Device address of the current GPU's scratch array (kNoAllocScratchBytes,
// 256-byte aligned), looked up on the first call per device (dh_noalloc.hip).
hipError_t noalloc_scratch(void** ptr);

// Separable HG overlap on MFMA: split = 0 plain FP32, 1 split-FP16 (hi+lo).
// tpw: stage-2 tiles per wave of the launch (0: the fewest that cover the
// triangle, overlap_min_tiles_per_wave; larger values give each wave more).
// peak: per-field largest |Re|, |Im| as float bits ([B][P]) when the producing
// kernel has left it, else null (the split form then finds it itself).
// hgx16/hgy16: the fp16 hi/lo basis copies (launch_hg_split); when given, the
// split form runs on k_overlap_split where that covers the shape.
hipError_t launch_overlap_mfma(const Geometry& g, const float2* fields, const float* hgx, const float* hgy,
                               float2* coefs, int split, hipStream_t s, int tpw = 0,
                               const unsigned* peak = nullptr, const void* hgx16 = nullptr,
                               const void* hgy16 = nullptr);
int overlap_min_tiles_per_wave(const Geometry& g, int split);
// Split-fp16 overlap with the basis operands held in registers (one wave per
// 16-mode tile, one (frame, pol) per workgroup) from the fp16 hi/lo basis
// copies of launch_hg_split; fields with w <= 160 and G <= 96 (w, h multiples of
// 16), else hipErrorNotSupported.
bool overlap_split_supported(const Geometry& g);
hipError_t launch_overlap_split(const Geometry& g, const float2* fields, const void* hgx16, const void* hgy16,
                                float2* coefs, hipStream_t s, const unsigned* peak);

// Reference-wave calibration plane (dh_refcal.hip).
// tw[m] = exp(-2 pi i m / N), FP64, m < N
hipError_t launch_twiddle_f64(int N, double2* tw, hipStream_t s);
// out(o, k) = sum_{n < N} in(o, n) exp(sign 2 pi i f_k n / N) with in(o, n) at
// in[o inO + n inN], out(o, k) at out[o outO + k outK]; fk == nullptr: f_k = k
hipError_t launch_cdft_axis(const float2* in, size_t inO, size_t inN, float2* out, size_t outO, size_t outK,
                            int nOuter, int N, int K, const int* fk, int sign, const double2* tw, hipStream_t s);
// weighted window Y [h][w] * wt [h][w] into a zeroed fh x fw field plane
hipError_t launch_cal_place(const float2* Y, const float* wt, int w, int h, int fw, int fh, int lowRes, int SX, int SY,
                            int W, int H, float2* plane, hipStream_t s);
// a[0..n) = 0 (a kernel, so it may be captured in a graph)
hipError_t launch_zero_complex(float2* a, size_t n, hipStream_t s);
// a[q] *= b[q % nb] for nq planes of n elements
hipError_t launch_cmul_planes(float2* a, const float2* b, size_t n, int nq, int nb, hipStream_t s);

}  // namespace dh
