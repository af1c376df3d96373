// digHolo-compatible C API over the HIP batch-processing kernels.
//
// Host responsibilities: handle registry and configuration (validation and
// return conventions of the reference's EXT_C wrappers, digHolo.cpp:13413-15415),
// window geometry (digHolo.cpp:7267-7391, :7868-8048, evaluated in float32 like
// the reference so the same pixels are chosen), the Fourier-window mask
// (float32, :1950-2003), FFT plans and twiddle tables, device buffers, and the
// launch sequence on the handle's HIP stream.  There is no CPU compute path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/digHolo.h"
#include "dh_eigen.h"
#include "dh_internal.h"

namespace {

constexpr float PI_F = 3.14159265358979323846f;        // digHolo.cpp:93
constexpr double PI_D = 3.14159265358979323846;
const float UNIT_ANGLE_F = PI_F / 180.0f;               // digHolo.h:192
constexpr int PIXEL_QUANTA = 16;                         // digHolo.cpp:99
constexpr int POLCOUNTMAX = 2;                           // digHolo.cpp:101

// Defaults of digHoloConfig (digHolo.cpp:3327-3425, :3440-3478).
struct Config {
  int frameWidth = 320, frameHeight = 256;
  float pixelSize = 20e-6f;
  float apertureSize = 0.4f;
  int resolutionIdx = 0;
  float waist[2] = {510.f, 510.f};
  float tiltX[2] = {-0.75f, -0.75f};
  float tiltY[2] = {0.75f, 0.75f};
  float defocus[2] = {0.f, 0.f};
  float beamCX[2] = {0.f, 0.f}, beamCY[2] = {0.f, 0.f};
  int batchCount = 0, avgCount = 1, avgMode = 0;
  float wavelengthCentre = 1565e-9f;
  // wavelength sweep (digHolo.cpp:6578-6680): the axis the next batch uses;
  // fewer than two entries means the centre wavelength alone (:7855-7863)
  std::vector<float> wavelengths;
  int wavelengthOrdering[2] = {0, 0};   // input frames, output fields/coefficients
  int polLockTilt = 0, polLockWaist = 1, polLockDefocus = 0;
  int polCount = 1;
  int aaTilt = 1, aaDefocus = 1, aaCentre = 1, aaWaist = 1, aaFourierRadius = 1;
  int aaPolIndependence = 0, aaMulConjTrans = 0;
  float aaTol = 0.f;
  int aaMode = 0, aaGoalIdx = 0;
  int fillFactor = 1;
  int fftWindowSizeX = 0, fftWindowSizeY = 0;
  int maxMG = 0;
  int basisType = 0;
  int verbosity = 0;
  int threadCount = (int)std::thread::hardware_concurrency();
  int precision = DIGHOLO_PRECISION_FP32;
  // batch calibration (digHolo.cpp:4469-4530): [calP][calB] complex, enabled flag
  std::vector<float2> cal;
  int calP = 0, calB = 0, calEnabled = 0;
  unsigned calVersion = 0;
  // custom basis (digHolo.cpp:4750-4765): column-major basisOut x basisIn
  std::vector<float2> basisFull;
  int basisIn = 0, basisOut = 0;
  unsigned basisVersion = 0;
  // reference-wave calibration (digHolo.cpp:5309-5520): [L][height][width],
  // stored inverted and normalised the way the setters leave it
  std::vector<float2> refCal;
  int refCalW = 0, refCalH = 0, refCalL = 0, refCalEnabled = 0;
  unsigned refCalVersion = 0;
};

// Device and pinned buffers free themselves when their handle is deleted.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t ensure(size_t count) {
    if (count <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

template <typename T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  hipError_t ensure(size_t count) {
    if (count <= n) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

struct FFTEntry {
  DevBuf<float2> tw;
  dh::FFTPlan plan{};
};

// Everything derived from the configuration that the kernels consume.
struct Derived {
  dh::Geometry g{};
  dh::TableArgs tab{};                  // wavelength 0
  std::vector<float> lam;               // the wavelength axis, L entries
  std::vector<dh::TableArgs> tabLam;    // every wavelength, tabLam[0] == tab
  // reference-wave calibration plane (active only when enabled and sized like the frame)
  struct Cal {
    bool active = false;
    unsigned version = 0;
    int Lc = 0;
    int SX = 0, SY = 0;                 // Fourier window of the calibration (zero tilt)
    std::vector<int> col[2], row[2];    // transform columns/rows each window pixel reads
    std::vector<float> wt;              // [Lc][P][h][w] mask * envelope * norm
    std::vector<int> outPos;            // plane position of calibration frame c
  } cal;
  int SX[2]{}, SY[2]{};
  std::vector<uint8_t> mask;
  float pxx = 0.f, pxy = 0.f;  // float32 reconstructed pixel pitch (axes export)
};

struct PowerMeterOut {
  double v[DIGHOLO_METRIC_COUNT];
};

struct Handle {
  Config cfg;
  const void* frames = nullptr;
  int frameU16 = 0;
  int transpose = 0;

  bool devInit = false;
  hipStream_t ownStream = nullptr;
  hipStream_t stream = nullptr;

  DevBuf<char> staging;
  const void* stagedFrames = nullptr;   // host buffer the staging copy was made from
  size_t stagedBytes = 0;
  bool holdStaging = false;             // reuse the staging copy (set only inside AutoAlign)
  bool holdCrop = false, holdWin = false;   // AutoAlign: keep FFT crops / Fourier window
  Derived heldCrop, heldWin;
  DevBuf<float2> rowspec, win, fields, coefs, factor, ref, twW;
  DevBuf<unsigned> fieldPeak;   // [B][P] largest |Re|, |Im| of each field (staged path), float bits
  DevBuf<float> hgx, hgy;
  DevBuf<uint8_t> mask;
  std::map<int, std::unique_ptr<FFTEntry>> ffts;
  int twWn = 0;

  bool tablesValid = false;
  Derived tabKey;  // derived state the device tables were built from
  Derived cur;     // derived state of the last processed batch
  bool haveBatch = false;

  PinnedBuf<float2> hCoefs, hFields;

  // power meter: device scratch and the last metrics [metric][lambda + 1] in dB
  DevBuf<double> metricRows;
  DevBuf<double2> gram;
  DevBuf<float2> coefsBackup, coefsMix;   // AutoAlign: best-so-far coefficients per pol
  std::vector<float> aaMetrics;   // DIGHOLO_METRIC_COUNT x (lambdaCount + 1)
  int aaLambdaCount = 0;
  std::vector<PowerMeterOut> pmLam;   // the last power meter call, per wavelength
  std::vector<float> lamAxis;         // wavelength axis handed out by digHoloConfigGetWavelengths
  bool coefsHostValid = false, fieldsHostValid = false;
  bool coefsDevValid = false;   // coefs belong to the published fields
  // pipeline run step by step (ProcessFFT -> ProcessIFFT -> ProcessRemoveTilt ->
  // ProcessBasisExtractCoefs): 0 none, 1 Fourier band in rowspec, 2 untilted
  // field in fields, 3 published
  int stepState = 0;
  Derived stepD;
  int stepA = 1, stepMode = 0;
  dh::BatchCal stepCal;
  const void* stepFrames = nullptr;
  bool stepU16 = false;
  std::vector<float> xAxis, yAxis;

  // Plane analysis, computed on request from the last batch's device state:
  // plane 0 the R2C Fourier plane of the frames, plane 1 the fields.
  const void* lastFrames = nullptr;   // device frames the last batch read
  int lastFrameU16 = 0;
  bool analysisValid[2] = {false, false};
  std::vector<float> analysis[2];     // [7][P][B + 1]
  std::vector<float> anaSum[2];       // [P][W * H], Ny * Nx used per pol
  std::vector<float> anaX[2], anaY[2];
  DevBuf<float2> spec0;
  DevBuf<float> anaPw, anaSumDev, anaAxes, anaScale;
  DevBuf<uint8_t> anaInc;
  DevBuf<double> anaStats;
  DevBuf<short> f16;                  // int16 fields [2][B][P][h][w]
  PinnedBuf<short> hF16;
  bool f16Valid = false;

  // Stage timing: one set of 6 events per batch enqueued since the last
  // collection, so back-to-back batches need no host synchronisation.
  // window path: 0 staged kernels, 1 fused (FFT columns), 2 fused with the
  // column transform on MFMA (uint16 frames), 3 size-specialised fused kernel
  // (register row transforms, uint16 frames); each falls back to the next lower
  int windowMode = 3;
  int lastPath = -1;        // path the last batch took
  DevBuf<uint16_t> colTab;  // fp16 hi/lo column-DFT operands of path 2
  DevBuf<float2> w3Tab;     // twiddle tables of path 3
  DevBuf<uint16_t> tstage;  // uint16 frames after the transpose pre-pass
  DevBuf<float2> winAll;    // every frame's Fourier window when averaging
  DevBuf<float2> plane;     // full-resolution IFFT input [B][P][H][W]
  DevBuf<float2> calDev;    // batch calibration on the device
  unsigned calDevVersion = ~0u;
  DevBuf<uint16_t> hgx16, hgy16;   // split-fp16 basis of the fused overlap
  // path 3 computes the coefficients in the window kernel: -1 for the split
  // form only (the faster choice on the measured shapes), 0 never, 1 in either precision
  int fusedOverlap = -1;
  bool lastFusedOverlap = false;
  // row transform of path 3 on 256-wide windows: 0 MFMA, 1 registers
  int rowMode = 0;

  // AutoAlign parameter estimate: batch-summed Fourier maps [P][Ny][Nx],
  // DCT work space, peak index, axes; far-field copies of the fields and maps
  DevBuf<float> snapMap, snapAxes, ffMap;
  DevBuf<double> snapWork, snapStats;
  DevBuf<int> snapIdx;
  DevBuf<float2> ffBuf;

  // alignment gradient: waist-derivative basis, -V R and conj(A) R, their
  // inverse transforms, per-frame partial sums; the placement of the last call
  DevBuf<float> dhgx, dhgy;
  DevBuf<float2> gq1, gq2, gx1, gx2;
  DevBuf<double> gPart, gSum, gAxes;
  DevBuf<float2> gExtra;
  DevBuf<double2> mdlX, mdlY;   // MDL seed: Gram eigenvectors and their images
  Derived gradHeld;
  bool gradHeldValid = false;

  // basis transform of the HG coefficients (LG / custom): the coefficients the
  // caller and the power meter see are coefsT with Mout modes per pol when
  // outT, else the HG ones
  DevBuf<float2> coefsT, lgTab, customT, basisW, basisFieldsDev;
  PinnedBuf<float2> hBasisFields;
  int lgTabG = -1;
  unsigned customTVersion = ~0u;
  bool outT = false;
  int Mout = 0;
  const float2* outCoefs() const { return outT ? coefsT.p : coefs.p; }
  float2* outCoefs() { return outT ? coefsT.p : coefs.p; }

  // autotuning: 0 off, 1 tune each new shape once, 2 remembered choices only;
  // ovTpw: stage-2 tiles per wave of the separate overlap kernel (0: fewest)
  int autoTune = 0;
  int ovTpw = 0;
  bool tuning = false, capturing = false;
  unsigned tabVersion = 0;  // incremented whenever the device tables are rebuilt

  // reference-wave calibration: the stored frames on the device, transform
  // work space, the calibration plane [Lc][P][fh][fw] of the current tables
  DevBuf<float2> refCalSrc, calA, calY, calT, calT2, calPlane;
  DevBuf<int> calIdx;
  DevBuf<float> calWt;
  DevBuf<double2> calTw;
  unsigned refCalSrcVersion = ~0u;
  std::vector<int> calIdxHost;
  bool calPlaneValid = false;
  int calPlaneL = 0;
  PinnedBuf<float2> hCal;

  // multi-batch launch (ProcessBatchMultiDevice): while `collect` is set, a
  // batch that takes the size-specialised window kernel (HG basis) records
  // that launch -- and the separate overlap that follows it -- instead of
  // issuing them, for one launch of all the handles' window kernels
  bool collect = false;
  bool preferV3 = false;   // the specialised window kernel wherever it applies, whatever the tuned choice
  bool pendValid = false, pendOverlap = false, pendSplit = false;
  dh::W3Launch pend{};
  dh::Geometry pendGF{};
  // afterMulti: recorded on another handle's stream after a multi-batch launch
  // that wrote this handle's outputs; the next API call on this handle makes
  // its stream wait on it (after_multi).  fence: marks this stream's own work
  // for a multi-batch launch on another stream.
  hipEvent_t afterMulti = nullptr;
  hipEvent_t fence = nullptr;

  bool stageTiming = false;
  std::vector<std::array<hipEvent_t, 6>> evPool;
  size_t evUsed = 0;
  float stageMs[5] = {};   // average over the collected batches
  int stageCount = 0;
};

std::vector<Handle*> g_handles;     // index = handle; nullptr once destroyed
std::mutex g_registry;
std::mutex g_multiMu;               // multi-batch launch state (MultiGraph, Handle::afterMulti)

// Orders the handle's stream after the last multi-batch launch that wrote its
// outputs from another stream.
void after_multi(Handle& h) {
  std::lock_guard<std::mutex> lk(g_multiMu);
  if (h.afterMulti) (void)hipStreamWaitEvent(h.stream, h.afterMulti, 0);
  h.afterMulti = nullptr;
}

// The handle without the ordering that get() applies.
Handle* get_raw(int idx) {
  std::lock_guard<std::mutex> lk(g_registry);
  if (idx < 0 || idx >= (int)g_handles.size()) return nullptr;
  return g_handles[idx];
}

// The handle behind an API call: any work the call queues on its stream
// follows the multi-batch launches that wrote its outputs.
Handle* get(int idx) {
  Handle* h = get_raw(idx);
  if (h && h->afterMulti) after_multi(*h);
  return h;
}

void log_err(const Handle*, const char* what, hipError_t e) {
  fprintf(stderr, "digHolo: %s failed: %s\n", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// geometry (float32 exactly as the reference evaluates it)
// ---------------------------------------------------------------------------
float k_factor(int n, float px) {
  const float pixelPeriod = (float)(2.0 * (double)px);
  return (float)(2 * (2.0 * PI_D / pixelPeriod) / n);     // digHolo.cpp:7726-7729
}
float k_axis(int i, int n, float kf) {                    // digHolo.cpp:2843-2867
  return i <= n / 2 ? kf * (float)i : kf * (float)(i - n);
}

int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Returns false (and an error code) if the configuration cannot be processed.
// FFT window size and per-pol corners (digHolo.cpp:7283-7377)
bool fft_window(const Config& c, int& Wout, int& Hout, int cx[2], int cy[2], int& err) {
  err = DIGHOLO_ERROR_SUCCESS;
  const int P = c.polCount;
  int fwx = c.fftWindowSizeX > 0 ? c.fftWindowSizeX : c.frameWidth / P;
  int fwy = c.fftWindowSizeY > 0 ? c.fftWindowSizeY : c.frameHeight;
  int W = PIXEL_QUANTA * (fwx / PIXEL_QUANTA);
  int H = PIXEL_QUANTA * (fwy / PIXEL_QUANTA);
  W = std::max(W, 16);
  H = std::max(H, 16);
  W = std::min(W, c.frameWidth);
  H = std::min(H, c.frameHeight);
  if (W < 16 || H < 16) {
    err = DIGHOLO_ERROR_INVALIDDIMENSION;
    return false;
  }
  const float px = c.pixelSize;
  for (int p = 0; p < P; ++p) {
    const float CX = (c.beamCX[p] / px) + (float)(c.frameWidth / 2);
    const float CY = (c.beamCY[p] / px) + (float)(c.frameHeight / 2);
    int x = (int)std::round((CX - (float)(W / 2)) / 16.0f) * 16;
    int y = (int)std::round(CY - (float)(H / 2));
    x = std::max(x, 0);
    y = std::max(y, 0);
    if (x + W > c.frameWidth) x = c.frameWidth - W;
    if (y + H > c.frameHeight) y = c.frameHeight - H;
    cx[p] = x;
    cy[p] = y;
  }
  Wout = W;
  Hout = H;
  return true;
}

double xcsc_d(double t) { return t == 0.0 ? 1.0 : t / std::sin(t); }

// (output position, wavelength index) of element idx of a batch of B over L
// wavelengths (digHolo.cpp:2656-2700)
int wavelength_ordering_calc(int idx, int L, int B, int ordIn, int ordOut, int& lam) {
  const int subB = L > 0 ? B / L : 0;
  int sub;
  if (!ordIn) {
    lam = idx % L;
    sub = idx / L;
  } else if (subB) {
    sub = idx % subB;
    lam = idx / subB;
  } else {
    lam = sub = 0;
  }
  return ordOut ? lam * subB + sub : sub * L + lam;
}

// Reference-wave calibration plan (digHolo.cpp:6792-6850, :8253-8292): the
// calibration's Fourier window is placed for zero tilt over the calibration's
// wavelength range with its own size (:7868-8048, complex spectrum: no
// half-plane clamp), then the signal window's size is what is copied; each
// window pixel is weighted with the aperture of its frame's wavelength, the
// fill-factor envelope and the FFT/IFFT norm (:1808-2122 with no beam centre).
void derive_cal(const Config& c, Derived& d, float kfx, float kfy) {
  Derived::Cal& k = d.cal;
  k = Derived::Cal{};
  const dh::Geometry& g = d.g;
  if (!c.refCalEnabled || c.refCalL <= 0 || c.refCalW != c.frameWidth || c.refCalH != c.frameHeight) return;
  const int W = g.W, H = g.H, w = g.w, h = g.h, P = g.P, L = g.L;
  k.active = true;
  k.version = c.refCalVersion;
  k.Lc = c.refCalL;
  const float lamA = k.Lc == 1 ? c.wavelengthCentre : d.lam[0];
  const float lamB = k.Lc == 1 ? c.wavelengthCentre : d.lam[L - 1];
  const float k0max = 2 * PI_F / std::fmax(lamA, lamB);
  const float k0min = 2 * PI_F / std::fmin(lamA, lamB);
  const float dkx = k_axis(1, W, kfx) - k_axis(0, W, kfx);
  const float dky = k_axis(1, H, kfy) - k_axis(0, H, kfy);
  const float r0 = c.apertureSize * UNIT_ANGLE_F;
  const float zero = 0.f;
  const float maxEX = (float)((double)((zero + r0) * k0max / dkx) + W / 2.0);
  const float maxEY = (float)((double)((zero + r0) * k0max / dky) + H / 2.0);
  const float minSX = (float)((double)((zero - r0) * k0min / dkx) + W / 2.0);
  const float minSY = (float)((double)((zero - r0) * k0min / dky) + H / 2.0);
  const int wc = (int)std::round((double)(maxEX + minSX) / 4.0) * 2;
  const int hc = (int)std::round((double)(maxEY + minSY) / 4.0) * 2;
  int wcal = (int)(maxEX - minSX + 1.0f);
  int hcal = (int)(maxEY - minSY + 1.0f);
  wcal = std::max(16, (int)(16 * std::ceil((1.0f * wcal) / 16)));
  hcal = std::max(16, (int)(16 * std::ceil((1.0f * hcal) / 16)));
  int sx = (int)(wc - wcal / 2.0);
  if (sx < 0) sx = 0;
  if (sx + wcal > W) sx = W - wcal;
  const int sy = (int)(hc - hcal / 2.0);
  k.SX = sx;
  k.SY = sy;
  const float fullShift = std::fabs((k_axis(1, H, kfy) - k_axis(0, H, kfy)) * (float)H);
  const bool lowRes = g.fw == w && g.fh == h && c.resolutionIdx == 1;
  const float norm0 = std::sqrt((float)((1.0 * g.fw * g.fh) / (1.0 * W * H)));
  const double ifftLen = lowRes ? (double)w * h : (double)W * H;
  const float norm = (float)(norm0 / std::sqrt((double)W * H * ifftLen));
  const double hp = 0.5 * (double)c.pixelSize;
  for (int p = 0; p < P; ++p) {
    k.col[p].resize(w);
    k.row[p].resize(h);
    for (int i = 0; i < w; ++i) k.col[p][i] = imod(i + sx - W / 2, W);
    for (int j = 0; j < h; ++j) k.row[p][j] = imod(j + sy + H / 2, H);
  }
  k.wt.assign((size_t)k.Lc * P * h * w, 0.f);
  k.outPos.resize(k.Lc);
  for (int cc = 0; cc < k.Lc; ++cc) {
    int lam = 0;
    k.outPos[cc] = wavelength_ordering_calc(cc, L, k.Lc, g.ordIn, g.ordOut, lam);
    if (k.outPos[cc] < 0 || k.outPos[cc] >= k.Lc) k.outPos[cc] = cc;
    const float k0 = 2 * PI_F / d.lam[std::min(lam, L - 1)];
    const float kr = k0 * (float)std::sin((double)r0);
    const float r02 = kr * kr;
    for (int p = 0; p < P; ++p)
      for (int j = 0; j < h; ++j) {
        const int jSY = j + sy;
        const float shift = jSY < 0 ? fullShift : (jSY > H ? -fullShift : 0.0f);
        const float ky = k_axis(k.row[p][j], H, kfy);
        const float dk_y = ky - (zero + shift);
        const float dky2 = dk_y * dk_y;
        for (int i = 0; i < w; ++i) {
          const float kx = k_axis(k.col[p][i], W, kfx);
          const float dk_x = kx - zero;
          const float dr2 = dk_x * dk_x + dky2;
          if (!(dr2 <= r02)) continue;
          const double env = c.fillFactor ? xcsc_d((double)kx * hp) * xcsc_d((double)ky * hp) : 1.0;
          k.wt[(((size_t)cc * P + p) * h + j) * w + i] = (float)(env * norm);
        }
      }
  }
}

// crop: keep the camera-plane FFT windows of an earlier derivation; win: keep
// its Fourier window (position, size and mask).  AutoAlign holds them the way
// the reference skips its FFT and IFFT while tweaking (digHolo.cpp:12620-12640).
bool derive(const Config& c, int B, Derived& d, int& err, const Derived* crop = nullptr,
            const Derived* win = nullptr) {
  err = DIGHOLO_ERROR_SUCCESS;
  const int P = c.polCount;
  int W = 0, H = 0, cx[2] = {0, 0}, cy[2] = {0, 0};
  if (!fft_window(c, W, H, cx, cy, err)) return false;
  const float px = c.pixelSize;
  for (int p = 0; p < P && crop; ++p) {
    cx[p] = crop->g.cx[p];
    cy[p] = crop->g.cy[p];
  }
  // --- wavelength axis (digHolo.cpp:7823-7865): a single entry is the centre ---
  d.lam = c.wavelengths.size() > 1 ? c.wavelengths : std::vector<float>{c.wavelengthCentre};
  const int L = (int)d.lam.size();
  for (float v : d.lam)
    if (!(v > 0.f) || !std::isfinite(v)) {
      err = DIGHOLO_ERROR_INVALIDARGUMENT;
      return false;
    }
  const int ordIn = c.wavelengthOrdering[0] ? 1 : 0, ordOut = c.wavelengthOrdering[1] ? 1 : 0;
  // a slow ordering needs whole sweeps: otherwise frames would share output
  // positions or index past the axis
  if (L > 1 && (ordIn || ordOut) && B % L != 0) {
    err = DIGHOLO_ERROR_INVALIDDIMENSION;
    return false;
  }
  // --- IFFT window, digHolo.cpp:7868-8048: wide enough for both ends of the axis ---
  const float lamA = d.lam[0], lamB = d.lam[L - 1];
  const float k0max = 2 * PI_F / std::fmax(lamA, lamB);
  const float k0min = 2 * PI_F / std::fmin(lamA, lamB);
  const float kfx = k_factor(W, px), kfy = k_factor(H, px);
  const float dkx = k_axis(1, W, kfx) - k_axis(0, W, kfx);
  const float dky = k_axis(1, H, kfy) - k_axis(0, H, kfy);
  const float r0 = c.apertureSize * UNIT_ANGLE_F;
  int w = 0, h = 0, wc[2] = {0, 0}, hc[2] = {0, 0};
  for (int p = 0; p < P && !win; ++p) {
    const int pp = c.polLockTilt ? 0 : p;
    const float tx = c.tiltX[pp] * UNIT_ANGLE_F;
    const float ty = c.tiltY[pp] * UNIT_ANGLE_F;
    const float maxEX = (float)((double)((tx + r0) * k0max / dkx) + W / 2.0);
    const float maxEY = (float)((double)((ty + r0) * k0max / dky) + H / 2.0);
    const float minSX = (float)((double)((tx - r0) * k0min / dkx) + W / 2.0);
    const float minSY = (float)((double)((ty - r0) * k0min / dky) + H / 2.0);
    wc[p] = (int)std::round((double)(maxEX + minSX) / 4.0) * 2;
    hc[p] = (int)std::round((double)(maxEY + minSY) / 4.0) * 2;
    int wp = (int)(maxEX - minSX + 1.0f);
    int hp = (int)(maxEY - minSY + 1.0f);
    wp = (int)(16 * std::ceil((1.0f * wp) / 16));
    hp = (int)(16 * std::ceil((1.0f * hp) / 16));
    w = std::max(w, wp);
    h = std::max(h, hp);
  }
  w = std::max(w, 16);
  h = std::max(h, 16);
  if (win) {
    w = win->g.w;
    h = win->g.h;
  }
  if (w > W / 2 || h > H) {
    err = DIGHOLO_ERROR_INVALIDDIMENSION;   // Fourier window larger than the r2c half plane
    return false;
  }
  for (int p = 0; p < P; ++p) {
    if (win) {
      d.SX[p] = win->SX[p];
      d.SY[p] = win->SY[p];
      continue;
    }
    int sx = (int)(wc[p] - w / 2.0);
    if (sx < W / 2) sx = W / 2;
    if (sx < 0) sx = 0;
    if (sx + w > W) sx = W - w;
    d.SX[p] = sx;
    d.SY[p] = (int)(hc[p] - h / 2.0);   // not clamped: rows wrap around ky (digHolo.cpp:8025-8036)
  }
  dh::Geometry& g = d.g;
  g = dh::Geometry{};
  g.B = B;
  g.P = P;
  g.frameW = c.frameWidth;
  g.frameH = c.frameHeight;
  g.W = W;
  g.H = H;
  g.w = w;
  g.h = h;
  // field plane: the window itself (low-res IFFT) or the whole FFT window
  // (full resolution, digHolo.cpp:8081-8082, :8175-8189)
  const bool lowRes = c.resolutionIdx == 1;
  g.fw = lowRes ? w : W;
  g.fh = lowRes ? h : H;
  for (int p = 0; p < P; ++p) {
    g.cx[p] = cx[p];
    g.cy[p] = cy[p];
    g.c0[p] = d.SX[p] - W / 2;
    g.SY[p] = d.SY[p];
  }
  g.G = c.maxMG;
  g.M = g.G * (g.G + 1) / 2;
  g.L = L;
  g.ordIn = ordIn;
  g.ordOut = ordOut;
  g.subB = B / L;

  // --- Fourier-window mask in float32 (digHolo.cpp:1950-2003, :8647-8656),
  // [L][P][h][w]: the aperture moves with each wavelength, the window does not ---
  d.mask.assign((size_t)L * P * h * w, 0);
  if (win) d.mask = win->mask;
  const float fullShift = std::fabs((k_axis(1, H, kfy) - k_axis(0, H, kfy)) * (float)H);
  for (int lp = 0; lp < L * P && !win; ++lp) {
    const int l = lp / P, p = lp - l * P;
    const float k0 = 2 * PI_F / d.lam[l];
    const float tx = c.tiltX[p] * UNIT_ANGLE_F;
    const float ty = c.tiltY[p] * UNIT_ANGLE_F;
    const float kx0 = k0 * (float)std::sin((double)tx);
    const float ky0 = k0 * (float)std::sin((double)ty);
    const float kr = k0 * (float)std::sin((double)r0);
    const float r02 = kr * kr;
    for (int j = 0; j < h; ++j) {
      const int jSY = j + d.SY[p];
      const int row = imod(jSY + H / 2, H);
      const float shift = jSY < 0 ? fullShift : (jSY > H ? -fullShift : 0.0f);
      const float dk_y = k_axis(row, H, kfy) - (ky0 + shift);
      const float dky2 = dk_y * dk_y;
      for (int i = 0; i < w; ++i) {
        const int col = g.c0[p] + i;
        const float dk_x = k_axis(col, W, kfx) - kx0;
        const float dkx2 = dk_x * dk_x;
        const float dr2 = dkx2 + dky2;
        d.mask[((size_t)lp * h + j) * w + i] = dr2 <= r02 ? 1 : 0;
      }
    }
  }

  // --- FP64 parameters of the window factor, reference wave and basis, per
  // wavelength (digHolo.cpp:10386-10420: k0 of each wavelength) ---
  d.tabLam.assign(L, dh::TableArgs{});
  const double pxd = (double)px;
  for (int l = 0; l < L; ++l) {
  dh::TableArgs& t = d.tabLam[l];
  t.g = g;
  const double lamd = (double)d.lam[l];
  const double k0d = 2.0 * PI_D / lamd;
  for (int p = 0; p < P; ++p) {
    dh::PolParams& q = t.pp[p];
    q.dkx = 2.0 * PI_D / (pxd * W);
    q.dky = 2.0 * PI_D / (pxd * H);
    const int cx0 = (cx[p] + W / 2) - c.frameWidth / 2;
    const int cy0 = (cy[p] + H / 2) - c.frameHeight / 2;
    q.bcx = (double)c.beamCX[p] - cx0 * pxd;
    q.bcy = (double)c.beamCY[p] - cy0 * pxd;
    q.px = pxd;
    q.norm = 1.0 / std::sqrt((double)W * H * g.fw * g.fh);   // digHolo.cpp:1813-1822
    q.fill = c.fillFactor ? 1 : 0;
    // reference wave: pol-lock index rules of digHolo.cpp:10289-10325
    const int tp = (c.polLockTilt || c.polLockDefocus) ? 0 : p;
    const int ip = c.polLockDefocus ? 0 : p;
    const double thx = (double)c.tiltX[tp] * PI_D / 180.0;
    const double thy = (double)c.tiltY[tp] * PI_D / 180.0;
    // window-offset tilt only in low-res mode (digHolo.cpp:10326-10334)
    const double thoffx = lowRes ? (d.SX[ip] - W / 2.0) * q.dkx : 0.0;
    const double thoffy = lowRes ? (d.SY[ip] - H / 2.0) * q.dky : 0.0;
    q.kxr = -k0d * thx + thoffx;
    q.kyr = -k0d * thy + thoffy;
    q.pistx = (double)c.beamCX[p] * (-k0d * thx);
    q.pisty = (double)c.beamCY[p] * (-k0d * thy);
    q.focal = -PI_D * (double)c.defocus[ip] / lamd;
    q.pxx = pxd * W / g.fw;
    q.pxy = pxd * H / g.fh;
    float wst = c.waist[c.polLockWaist ? 0 : p];
    if (wst == 0.f) wst = 400e-6f;
    q.waist = (double)wst;
  }
  }
  d.tab = d.tabLam[0];
  d.pxx = (px * W) / g.fw;              // digHolo.cpp:8184-8185 (float)
  d.pxy = (px * H) / g.fh;
  derive_cal(c, d, kfx, kfy);
  return true;
}

bool same_tables(const Derived& a, const Derived& b) {
  const dh::Geometry &x = a.g, &y = b.g;
  if (x.P != y.P || x.W != y.W || x.H != y.H || x.w != y.w || x.h != y.h || x.G != y.G || x.fw != y.fw ||
      x.fh != y.fh)
    return false;
  for (int p = 0; p < x.P; ++p)
    if (x.c0[p] != y.c0[p] || x.SY[p] != y.SY[p]) return false;
  if (x.L != y.L || x.ordIn != y.ordIn || x.ordOut != y.ordOut || x.subB != y.subB) return false;
  if (a.tabLam.size() != b.tabLam.size()) return false;
  for (size_t l = 0; l < a.tabLam.size(); ++l)
    if (std::memcmp(a.tabLam[l].pp, b.tabLam[l].pp, sizeof(a.tabLam[l].pp)) != 0) return false;
  const Derived::Cal &ca = a.cal, &cb = b.cal;
  if (ca.active != cb.active) return false;
  if (ca.active) {
    if (ca.version != cb.version || ca.Lc != cb.Lc || ca.SX != cb.SX || ca.SY != cb.SY || ca.wt != cb.wt ||
        ca.outPos != cb.outPos)
      return false;
    for (int p = 0; p < x.P; ++p)
      if (x.cx[p] != y.cx[p] || x.cy[p] != y.cy[p]) return false;   // the calibration frame's crop
  }
  return a.mask == b.mask;
}

// ---------------------------------------------------------------------------
// FFT plans
// ---------------------------------------------------------------------------
bool factorise(int n, dh::FFTPlan& pl) {
  pl.n = n;
  pl.nstages = 0;
  int m = n;
  const int pref[] = {8, 4, 2, 3, 5};
  for (int r : pref) {
    while (m % r == 0) {
      if (pl.nstages >= dh::kMaxStages) return false;
      pl.radix[pl.nstages++] = r;
      m /= r;
    }
  }
  for (int r = 7; m > 1; r += 2) {
    while (m % r == 0) {
      if (pl.nstages >= dh::kMaxStages) return false;
      pl.radix[pl.nstages++] = r;
      m /= r;
    }
  }
  int Ns = 1;
  for (int s = 0; s < pl.nstages; ++s) {
    pl.magNbf[s] = dh::magic_div((unsigned)(n / pl.radix[s]));
    pl.magNs[s] = dh::magic_div((unsigned)Ns);
    Ns *= pl.radix[s];
  }
  return true;
}

const dh::FFTPlan* fft_plan(Handle& h, int n) {
  auto it = h.ffts.find(n);
  if (it != h.ffts.end()) return &it->second->plan;
  auto e = std::make_unique<FFTEntry>();
  if (!factorise(n, e->plan)) return nullptr;
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * PI_D * (double)k / n;
    tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  if (e->tw.ensure(n) != hipSuccess) return nullptr;
  if (hipMemcpy(e->tw.p, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  e->plan.tw = e->tw.p;
  const dh::FFTPlan* r = &e->plan;
  h.ffts[n] = std::move(e);
  return r;
}

// Picks the number of row residues R for the fused window kernel: the
// smallest divisor of H whose H/R-point column transforms fit its budget.
bool plan_fused(Handle& h, const dh::Geometry& g, dh::FusedArgs& fa) {
  const dh::FFTPlan* row = fft_plan(h, g.W / 2);
  const dh::FFTPlan* ix = fft_plan(h, g.w);
  const dh::FFTPlan* iy = fft_plan(h, g.h);
  if (!row || !ix || !iy) return false;
  for (int R = 1; R <= g.H; ++R) {
    if (g.H % R) continue;
    const int Hc = g.H / R;
    if (Hc < 8) break;
    const dh::FFTPlan* col = fft_plan(h, Hc);
    if (!col) return false;
    fa = dh::FusedArgs{};
    fa.g = g;
    fa.row = *row;
    fa.col = *col;
    fa.ix = *ix;
    fa.iy = *iy;
    fa.R = R;
    fa.Hc = Hc;
    if (dh::fused_prepare(fa)) return true;
  }
  return false;
}

bool plan_window_mfma(Handle& h, const dh::Geometry& g, dh::FusedArgs& fa) {
  const dh::FFTPlan* row = fft_plan(h, g.W / 2);
  const dh::FFTPlan* ix = fft_plan(h, g.w);
  const dh::FFTPlan* iy = fft_plan(h, g.h);
  if (!row || !ix || !iy) return false;
  fa = dh::FusedArgs{};
  fa.g = g;
  fa.row = *row;
  fa.col = *row;
  fa.ix = *ix;
  fa.iy = *iy;
  return dh::window_mfma_prepare(fa);
}

bool plan_window_v3(Handle& h, const dh::Geometry& g, dh::FusedArgs& fa) {
  const dh::FFTPlan* ix = fft_plan(h, g.w);
  const dh::FFTPlan* iy = fft_plan(h, g.h);
  if (!ix || !iy) return false;
  fa = dh::FusedArgs{};
  fa.g = g;
  fa.ix = *ix;
  fa.iy = *iy;
  return dh::window_v3_supported(fa);
}

bool is_device_pointer(const void* ptr) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.isManaged;
}

int ensure_device(Handle& h) {
  if (h.devInit) return DIGHOLO_ERROR_SUCCESS;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    fprintf(stderr, "digHolo: no HIP device available; this library has no CPU path\n");
    return DIGHOLO_ERROR_ERROR;
  }
  if (hipStreamCreateWithFlags(&h.ownStream, hipStreamNonBlocking) != hipSuccess) return DIGHOLO_ERROR_ERROR;
  if (!h.stream) h.stream = h.ownStream;
  h.devInit = true;
  return DIGHOLO_ERROR_SUCCESS;
}

#define DH_TRY(expr, what)                 \
  do {                                     \
    hipError_t _e = (expr);                \
    if (_e != hipSuccess) {                \
      log_err(&h, what, _e);               \
      return DIGHOLO_ERROR_ERROR;          \
    }                                      \
  } while (0)

// Device copy of the caller's frames (nFrames of frameWidth x frameHeight):
// device pointers are used in place; host buffers are uploaded (once for the
// whole search inside AutoAlign); transposed uint16 frames are transposed into
// a staging buffer (digHolo.cpp:1206-1229, once per call like the reference's
// conversion).
int device_frames(Handle& h, size_t nFrames, const void*& dframes) {
  const Config& c = h.cfg;
  const size_t frameElems = nFrames * (size_t)c.frameWidth * c.frameHeight;
  const size_t elemBytes = h.frameU16 ? 2 : 4;
  dframes = h.frames;
  if (!is_device_pointer(h.frames) || (reinterpret_cast<uintptr_t>(h.frames) % 8) != 0) {
    const bool staged = h.holdStaging && h.stagedFrames == h.frames && h.stagedBytes == frameElems * elemBytes;
    if (!staged) {
      DH_TRY(h.staging.ensure(frameElems * elemBytes), "frame staging alloc");
      DH_TRY(hipMemcpyAsync(h.staging.p, h.frames, frameElems * elemBytes, hipMemcpyDefault, h.stream),
             "frame upload");
      h.stagedFrames = h.frames;
      h.stagedBytes = frameElems * elemBytes;
    }
    dframes = h.staging.p;
  }
  if (h.transpose && h.frameU16) {
    DH_TRY(h.tstage.ensure(frameElems), "alloc");
    DH_TRY(dh::launch_u16_transpose(dframes, h.tstage.p, c.frameWidth, c.frameHeight, nFrames, h.stream),
           "transpose");
    dframes = h.tstage.p;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// Twiddles exp(-2 pi i k / W), k <= W/2, of the r2c row unpack for window width W
int ensure_row_twiddles(Handle& h, int W) {
  if (h.twWn == W) return DIGHOLO_ERROR_SUCCESS;
  std::vector<float2> tw(W / 2 + 1);
  for (int k = 0; k <= W / 2; ++k) {
    const double a = -2.0 * PI_D * (double)k / W;
    tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  DH_TRY(h.twW.ensure(tw.size()), "alloc");
  DH_TRY(hipMemcpy(h.twW.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice), "twiddles");
  h.twWn = W;
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Autotuner (no counterpart in the reference).  For each batch shape -- frame
// and window sizes, pols, groups, precision, input type, batch size rounded to
// a power of two -- the applicable launch variants are timed on the batch
// itself with HIP events: window path (size-specialised / MFMA columns / FFT
// columns / staged), the size-specialised kernel's row transform where it has
// two (MFMA or registers, 256-wide windows), the split overlap fused into the
// window kernel or as its own kernel, and the separate overlap kernel's
// stage-2 tiles per wave (the fewest that cover the triangle, or twice that).
// The fastest is remembered in memory and in a text cache file
// ($DIGHOLO_TUNE_CACHE, else $HOME/.cache/digholo_amd_tune.txt), one
// "key path fused tpw rows ms" line per shape.
// ---------------------------------------------------------------------------
struct TuneChoice {
  int windowMode = 3, fused = 1, tpw = 0, rows = 0;
  float ms = 0.f;
};
std::mutex g_tuneMu;
std::map<std::string, TuneChoice> g_tune;
bool g_tuneLoaded = false;

std::string tune_path() {
  if (const char* e = std::getenv("DIGHOLO_TUNE_CACHE")) return e;
  if (const char* home = std::getenv("HOME")) return std::string(home) + "/.cache/digholo_amd_tune.txt";
  return "digholo_amd_tune.txt";
}

// Tune-cache lines carry this tag: choices written by a build whose kernels
// differ (other tile counts, other paths) are not reused.
constexpr const char* kTuneVersion = "v3";

// A cached choice is applied only if the library could have made it itself.
bool tune_choice_valid(const TuneChoice& t) {
  return t.windowMode >= 0 && t.windowMode <= 3 && (t.fused == 0 || t.fused == 1) && t.tpw >= 0 && t.tpw <= 32 &&
         (t.rows == 0 || t.rows == 1) && std::isfinite(t.ms) && t.ms >= 0.f;
}

void tune_load_locked() {
  if (g_tuneLoaded) return;
  g_tuneLoaded = true;
  FILE* f = std::fopen(tune_path().c_str(), "r");
  if (!f) return;
  char key[512];
  TuneChoice t;
  while (std::fscanf(f, "%511s %d %d %d %d %f", key, &t.windowMode, &t.fused, &t.tpw, &t.rows, &t.ms) == 6)
    if (tune_choice_valid(t) && std::strncmp(key, kTuneVersion, std::strlen(kTuneVersion)) == 0) g_tune[key] = t;
  std::fclose(f);
}

// mkdir -p without a shell: every missing directory of the path, in order
void make_dirs(const std::string& dir) {
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)::mkdir(dir.substr(0, i).c_str(), 0755);
}

void tune_save_locked() {
  const std::string path = tune_path();
  const size_t slash = path.find_last_of('/');
  if (slash != std::string::npos && slash > 0) make_dirs(path.substr(0, slash));
  // a temporary name of this process's own, renamed over the file in one step,
  // so concurrent writers never interleave their lines
  const std::string tmp = path + ".tmp" + std::to_string((long)::getpid());
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  for (const auto& kv : g_tune)
    std::fprintf(f, "%s %d %d %d %d %.6f\n", kv.first.c_str(), kv.second.windowMode, kv.second.fused,
                 kv.second.tpw, kv.second.rows, kv.second.ms);
  if (std::fclose(f) == 0) {
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
  } else {
    std::remove(tmp.c_str());
  }
}

std::string tune_key(const Handle& h, const Derived& d) {
  const Config& c = h.cfg;
  const dh::Geometry& g = d.g;
  int bb = 1;
  while (bb < g.B) bb <<= 1;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s_f%dx%d_p%d_W%dx%d_w%dx%d_f%dx%d_G%d_pr%d_u%d_t%d_a%d_c%d_B%d", kTuneVersion,
                g.frameW, g.frameH, g.P, g.W, g.H, g.w, g.h, g.fw, g.fh, g.G, c.precision, h.frameU16, h.transpose, c.avgCount,
                c.calEnabled, bb);
  return buf;
}

int enqueue_batch(Handle& h, int step);

int autotune(Handle& h) {
  Config& c = h.cfg;
  Derived d;
  int err;
  if (!derive(c, c.batchCount, d, err, h.holdCrop ? &h.heldCrop : nullptr, h.holdWin ? &h.heldWin : nullptr))
    return DIGHOLO_ERROR_SUCCESS;   // enqueue_batch reports the error
  const std::string key = tune_key(h, d);
  {
    std::lock_guard<std::mutex> lk(g_tuneMu);
    tune_load_locked();
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      h.windowMode = it->second.windowMode;
      h.fusedOverlap = it->second.fused;
      h.ovTpw = it->second.tpw;
      h.rowMode = it->second.rows;
      return DIGHOLO_ERROR_SUCCESS;
    }
  }
  if (h.autoTune != 1) return DIGHOLO_ERROR_SUCCESS;
  if (int e = ensure_device(h)) return e;
  h.tuning = true;
  const int minTpw =
      d.g.G > 0 ? dh::overlap_min_tiles_per_wave(d.g, c.precision == DIGHOLO_PRECISION_SPLIT16) : 0;
  hipEvent_t e0, e1;
  DH_TRY(hipEventCreate(&e0), "event create");
  DH_TRY(hipEventCreate(&e1), "event create");
  TuneChoice best;
  best.ms = std::numeric_limits<float>::max();
  std::vector<std::array<int, 4>> seen;   // (actual path, fused, tpw, rows) already timed
  const int rowChoices = dh::window_v3_row_choice(d.g) ? 2 : 1;
  int rc = DIGHOLO_ERROR_SUCCESS;
  for (int wm = 3; wm >= 0 && !rc; --wm)
    for (int rm = 0; rm < (wm == 3 ? rowChoices : 1) && !rc; ++rm)
    for (int fu = 1; fu >= 0 && !rc; --fu)
      for (int tp : {0, 2 * minTpw}) {
        if (tp > 32 || (tp && minTpw == 0)) continue;
        h.windowMode = wm;
        h.fusedOverlap = fu;
        h.ovTpw = tp;
        h.rowMode = rm;
        if ((rc = enqueue_batch(h, 0))) break;
        const std::array<int, 4> sig{h.lastPath, h.lastFusedOverlap ? 1 : 0, h.lastFusedOverlap ? 0 : tp,
                                     h.lastPath == 3 ? rm : 0};
        if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
        seen.push_back(sig);
        constexpr int reps = 3;
        (void)hipEventRecord(e0, h.stream);
        for (int r = 0; r < reps && !rc; ++r) rc = enqueue_batch(h, 0);
        (void)hipEventRecord(e1, h.stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (c.verbosity >= DIGHOLO_VERBOSITY_DEBUG)
          fprintf(stdout, "autotune %s: path %d fused %d tpw %d rows %d  %.4f ms\n", key.c_str(), sig[0], sig[1], tp,
                  sig[3], ms);
        if (ms < best.ms) best = TuneChoice{wm, fu, tp, rm, ms};
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h.tuning = false;
  if (rc) return rc;
  h.windowMode = best.windowMode;
  h.fusedOverlap = best.fused;
  h.ovTpw = best.tpw;
  h.rowMode = best.rows;
  std::lock_guard<std::mutex> lk(g_tuneMu);
  g_tune[key] = best;
  tune_save_locked();
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Basis transforms (digHolo.cpp:9469-9605).  LG: per HG group g the unitary
// U_g with LG[k] = sum_c U_g[k][c] HG[c] (digHoloLGtoHGcoef :9819-9906): LG
// mode k of the group has the index pair of HG mode k (m = g-k, n = k),
// l = m - n, p = (g - |l|)/2, (a, b) = (p, p - l) for l < 0 and (l + p, p)
// otherwise; for HG (hm = g - c, hn = c)
//   U_g[k][c] = i^(g + hm) (-1)^[g odd, l > 0] 2^(-g/2) sqrt(hm! hn! / (a! b!))
//               sum_t (-1)^t C(b, t) C(a, hm - t),
// the sum in exact 128-bit integers (every term is at most C(g, hm)).  Stored
// row-major per block, blocks one after another.  Custom: the caller's
// column-major matrix, the min(M, in) x min(M, out) block used.
// ---------------------------------------------------------------------------
std::vector<float2> lg_table(int G) {
  using i128 = __int128;
  std::vector<std::vector<i128>> C(G + 1, std::vector<i128>(G + 1, 0));
  for (int n = 0; n <= G; ++n) {
    C[n][0] = 1;
    for (int k = 1; k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k <= n - 1 ? C[n - 1][k] : 0);
  }
  std::vector<long double> lf(G + 1, 0.0L);
  for (int n = 1; n <= G; ++n) lf[n] = lf[n - 1] + std::log((long double)n);
  std::vector<float2> t;
  for (int g = 0; g < G; ++g)
    for (int k = 0; k <= g; ++k) {
      const int l = (g - k) - k;
      const int p = (g - std::abs(l)) / 2;
      const int a = l < 0 ? p : l + p, b = l < 0 ? -l + p : p;
      for (int c = 0; c <= g; ++c) {
        const int hm = g - c, hn = c;
        i128 sum = 0;
        for (int tt = 0; tt <= std::min(hm, b); ++tt)
          if (hm - tt <= a) sum += ((tt & 1) ? -1 : 1) * C[b][tt] * C[a][hm - tt];
        long double v = (long double)sum * std::exp(0.5L * (lf[hm] + lf[hn] - lf[a] - lf[b]) - 0.5L * g * std::log(2.0L));
        if ((g & 1) && l > 0) v = -v;
        const int ph = (g + hm) & 3;   // i^ph
        const float re = ph == 0 ? (float)v : ph == 2 ? (float)-v : 0.f;
        const float im = ph == 1 ? (float)v : ph == 3 ? (float)-v : 0.f;
        t.push_back(make_float2(re, im));
      }
    }
  return t;
}

// After the overlap: the coefficients in the configured basis (an unsupported
// type, or a custom basis without a matrix, falls back to HG, :9474-9481).
int apply_basis(Handle& h) {
  const Config& c = h.cfg;
  const dh::Geometry& g = h.cur.g;
  h.outT = false;
  h.Mout = g.M;
  if (g.G <= 0 || g.M <= 0) return DIGHOLO_ERROR_SUCCESS;
  const size_t rows = (size_t)g.B * g.P;
  if (c.basisType == DIGHOLO_BASISTYPE_LG) {
    if (h.lgTabG != g.G) {
      const std::vector<float2> t = lg_table(g.G);
      DH_TRY(h.lgTab.ensure(t.size()), "alloc");
      DH_TRY(hipMemcpy(h.lgTab.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice), "LG table");
      h.lgTabG = g.G;
    }
    DH_TRY(h.coefsT.ensure(rows * g.M), "alloc");
    DH_TRY(dh::launch_lg_apply(h.coefs.p, h.coefsT.p, rows, g.M, h.lgTab.p, h.stream), "LG transform");
    h.outT = true;
  } else if (c.basisType == DIGHOLO_BASISTYPE_CUSTOM && !c.basisFull.empty() && c.basisIn > 0 && c.basisOut > 0) {
    if (h.customTVersion != c.basisVersion) {
      DH_TRY(h.customT.ensure(c.basisFull.size()), "alloc");
      DH_TRY(hipMemcpy(h.customT.p, c.basisFull.data(), c.basisFull.size() * sizeof(float2), hipMemcpyHostToDevice),
             "custom basis");
      h.customTVersion = c.basisVersion;
    }
    const int nIn = std::min(g.M, c.basisIn), nOut = std::min(g.M, c.basisOut);
    DH_TRY(h.coefsT.ensure(rows * nOut), "alloc");
    DH_TRY(dh::launch_custom_apply(h.coefs.p, g.M, (int)rows, nIn, h.customT.p, c.basisOut, nOut, h.coefsT.p,
                                   h.stream), "custom transform");
    h.outT = true;
    h.Mout = nOut;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// Calibration plane of derivation d on the device (digHolo.cpp:6729-6850,
// :8253-8292), then folded into the reference wave: the reference multiplies
// the inverse-transformed field by the plane of its wavelength before the tilt
// is removed (:8736-8758, :8770-8785), and V = -conj(E C R) = -conj(E (C R)).
int build_ref_calibration(Handle& h, const Derived& d) {
  const Config& c = h.cfg;
  const dh::Geometry& g = d.g;
  const Derived::Cal& k = d.cal;
  const int P = g.P, W = g.W, H = g.H, w = g.w, hh = g.h, fw = g.fw, fh = g.fh, Lc = k.Lc;
  const size_t frame = (size_t)c.frameWidth * c.frameHeight;
  if (h.refCalSrcVersion != c.refCalVersion) {
    DH_TRY(h.refCalSrc.ensure(c.refCal.size()), "alloc");
    DH_TRY(hipMemcpyAsync(h.refCalSrc.p, c.refCal.data(), c.refCal.size() * sizeof(float2), hipMemcpyHostToDevice,
                          h.stream), "calibration upload");
    DH_TRY(hipStreamSynchronize(h.stream), "calibration upload");
    h.refCalSrcVersion = c.refCalVersion;
  }
  // column lists [P][w], row lists [P][h]
  h.calIdxHost.assign((size_t)P * (w + hh), 0);
  for (int p = 0; p < P; ++p) {
    std::copy(k.col[p].begin(), k.col[p].end(), h.calIdxHost.begin() + (size_t)p * w);
    std::copy(k.row[p].begin(), k.row[p].end(), h.calIdxHost.begin() + (size_t)P * w + (size_t)p * hh);
  }
  DH_TRY(h.calIdx.ensure(h.calIdxHost.size()), "alloc");
  DH_TRY(h.calWt.ensure(k.wt.size()), "alloc");
  DH_TRY(hipMemcpyAsync(h.calIdx.p, h.calIdxHost.data(), h.calIdxHost.size() * sizeof(int), hipMemcpyHostToDevice,
                        h.stream), "calibration upload");
  DH_TRY(hipMemcpyAsync(h.calWt.p, k.wt.data(), k.wt.size() * sizeof(float), hipMemcpyHostToDevice, h.stream),
         "calibration upload");
  // twiddles of the four transform lengths W, H, fw, fh
  const int Ns[4] = {W, H, fw, fh};
  size_t off[4], tot = 0;
  for (int q = 0; q < 4; ++q) {
    off[q] = tot;
    tot += Ns[q];
  }
  DH_TRY(h.calTw.ensure(tot), "alloc");
  for (int q = 0; q < 4; ++q) DH_TRY(dh::launch_twiddle_f64(Ns[q], h.calTw.p + off[q], h.stream), "twiddles");
  const size_t nPlane = (size_t)fh * fw;
  DH_TRY(h.calA.ensure((size_t)H * w), "alloc");
  DH_TRY(h.calY.ensure((size_t)hh * w), "alloc");
  DH_TRY(h.calT.ensure(nPlane), "alloc");
  DH_TRY(h.calT2.ensure(nPlane), "alloc");
  DH_TRY(h.calPlane.ensure((size_t)Lc * P * nPlane), "alloc");
  DH_TRY(dh::launch_zero_complex(h.calPlane.p, (size_t)Lc * P * nPlane, h.stream), "calibration plane");
  const bool lowRes = c.resolutionIdx == 1;
  for (int cc = 0; cc < Lc; ++cc)
    for (int p = 0; p < P; ++p) {
      const float2* src = h.refCalSrc.p + cc * frame + (size_t)g.cy[p] * c.frameWidth + g.cx[p];
      const int* cols = h.calIdx.p + (size_t)p * w;
      const int* rows = h.calIdx.p + (size_t)P * w + (size_t)p * hh;
      DH_TRY(dh::launch_cdft_axis(src, c.frameWidth, 1, h.calA.p, w, 1, H, W, w, cols, -1, h.calTw.p + off[0],
                                  h.stream), "calibration transform");
      DH_TRY(dh::launch_cdft_axis(h.calA.p, 1, w, h.calY.p, 1, w, w, H, hh, rows, -1, h.calTw.p + off[1], h.stream),
             "calibration transform");
      DH_TRY(dh::launch_cal_place(h.calY.p, h.calWt.p + ((size_t)cc * P + p) * hh * w, w, hh, fw, fh, lowRes ? 1 : 0,
                                  k.SX, k.SY, W, H, h.calT.p, h.stream), "calibration window");
      DH_TRY(dh::launch_cdft_axis(h.calT.p, fw, 1, h.calT2.p, fw, 1, fh, fw, fw, nullptr, +1, h.calTw.p + off[2],
                                  h.stream), "calibration inverse transform");
      float2* dst = h.calPlane.p + ((size_t)k.outPos[cc] * P + p) * nPlane;
      DH_TRY(dh::launch_cdft_axis(h.calT2.p, 1, fw, dst, 1, fw, fw, fh, fh, nullptr, +1, h.calTw.p + off[3], h.stream),
             "calibration inverse transform");
    }
  DH_TRY(dh::launch_cmul_planes(h.ref.p, h.calPlane.p, nPlane, g.L * P, Lc * P, h.stream), "calibrated reference");
  h.calPlaneValid = true;
  h.calPlaneL = Lc;
  return DIGHOLO_ERROR_SUCCESS;
}

// Queues one full batch on the handle's stream.  On return the coefficients and
// fields are (or will be, once the stream drains) in device memory.
void publish_batch(Handle& h, const Derived& d, const void* dframes, bool u16);

// step: 0 the whole batch; 1 the FFT step only (frames -> column-transformed
// band of the window columns, [B][P][H][w] in rowspec).
int enqueue_batch(Handle& h, int step = 0) {
  Config& c = h.cfg;
  if (!h.frames) return DIGHOLO_ERROR_NULLPOINTER;
  if (c.batchCount < 1) c.batchCount = 1;
  // pol locks write pol 0's values into pol 1 (digHolo.cpp:10296-10318, :10527-10531)
  if (c.polLockTilt) {
    c.tiltX[1] = c.tiltX[0];
    c.tiltY[1] = c.tiltY[0];
  }
  if (c.polLockDefocus) c.defocus[1] = c.defocus[0];
  if (c.polLockWaist) c.waist[1] = c.waist[0];
  if (h.autoTune && step == 0 && !h.tuning && !h.capturing) {
    if (int et = autotune(h)) return et;
  }

  Derived d;
  int err;
  if (!derive(c, c.batchCount, d, err, h.holdCrop ? &h.heldCrop : nullptr, h.holdWin ? &h.heldWin : nullptr)) {
    fprintf(stderr, "digHolo: window configuration is invalid for this frame (error %d)\n", err);
    return err;
  }
  const dh::Geometry& g = d.g;
  int e0 = ensure_device(h);
  if (e0) return e0;

  const size_t B = g.B, P = g.P;
  const int A = std::max(1, c.avgCount);      // frames averaged into each batch element
  const size_t nFrames = B * (size_t)A;
  if (nFrames > (size_t)dh::kMaxBatchFrames) {
    fprintf(stderr, "digHolo: at most %d frames per batch (got %zu); split the batch\n", dh::kMaxBatchFrames,
            nFrames);
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  if (A > 1 && g.L > 1) {
    fprintf(stderr, "digHolo: frame averaging is not available together with a wavelength sweep\n");
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  const void* dframes = nullptr;
  if (const int ef = device_frames(h, nFrames, dframes)) return ef;
  dh::BatchCal cal;
  if (c.calEnabled && c.calP > 0 && c.calB > 0) {
    if (h.calDevVersion != c.calVersion) {
      DH_TRY(h.calDev.ensure(c.cal.size()), "alloc");
      DH_TRY(hipMemcpyAsync(h.calDev.p, c.cal.data(), c.cal.size() * sizeof(float2), hipMemcpyHostToDevice,
                            h.stream), "calibration upload");
      DH_TRY(hipStreamSynchronize(h.stream), "calibration upload");
      h.calDevVersion = c.calVersion;
    }
    cal.p = h.calDev.p;
    cal.P = c.calP;
    cal.B = c.calB;
  }
  dh::Geometry gA = g;   // every frame of the buffer (averaging reads A per batch element)
  gA.B = (int)nFrames;
  DH_TRY(h.fields.ensure(B * P * g.fh * g.fw), "alloc");
  DH_TRY(h.coefs.ensure(std::max<size_t>(B * P * g.M, 1)), "alloc");
  dh::Geometry gF = g;   // the field plane, for the inverse transforms, tilt and overlap
  gF.w = g.fw;
  gF.h = g.fh;
  const bool fullRes = g.fw != g.w || g.fh != g.h;

  const dh::FFTPlan* pRow = fft_plan(h, g.W / 2);
  const dh::FFTPlan* pCol = fft_plan(h, g.H);
  const dh::FFTPlan* pIx = fft_plan(h, g.fw);
  const dh::FFTPlan* pIy = fft_plan(h, g.fh);
  if (!pRow || !pCol || !pIx || !pIy) {
    fprintf(stderr, "digHolo: FFT plan setup failed\n");
    return DIGHOLO_ERROR_MEMORYALLOCATION;
  }
  if (const int et = ensure_row_twiddles(h, g.W)) return et;

  const bool timing = h.stageTiming && !h.capturing;
  hipEvent_t* ev = nullptr;
  if (timing) {
    if (h.evUsed == h.evPool.size()) {
      std::array<hipEvent_t, 6> set{};
      for (auto& e : set) DH_TRY(hipEventCreate(&e), "event create");
      h.evPool.push_back(set);
    }
    ev = h.evPool[h.evUsed++].data();
    DH_TRY(hipEventRecord(ev[0], h.stream), "event");
  }
  if (!h.tablesValid || !same_tables(d, h.tabKey)) {
    // window factor [L][P][h][w] and reference wave [L][P][fh][fw]; the HG
    // basis does not depend on the wavelength
    const size_t T = P * g.h * g.w, TR = P * (size_t)g.fh * g.fw, L = d.tabLam.size();
    DH_TRY(h.mask.ensure(T * L), "alloc");
    DH_TRY(h.factor.ensure(T * L), "alloc");
    DH_TRY(h.ref.ensure(TR * L), "alloc");
    DH_TRY(h.hgx.ensure(std::max<size_t>(P * g.G * g.fw, 1)), "alloc");
    DH_TRY(h.hgy.ensure(std::max<size_t>(P * g.G * g.fh, 1)), "alloc");
    DH_TRY(hipMemcpyAsync(h.mask.p, d.mask.data(), T * L, hipMemcpyHostToDevice, h.stream), "mask upload");
    for (size_t l = 0; l < L; ++l)
      DH_TRY(dh::launch_tables(d.tabLam[l], h.mask.p + T * l, h.factor.p + T * l, h.ref.p + TR * l,
                               l ? nullptr : h.hgx.p, l ? nullptr : h.hgy.p, h.stream), "tables");
    h.calPlaneValid = false;
    if (d.cal.active)
      if (const int ec = build_ref_calibration(h, d)) return ec;
    DH_TRY(h.colTab.ensure(dh::coldft_table_elems(g)), "alloc");
    DH_TRY(dh::launch_coldft_table(g, h.colTab.p, h.stream), "column table");
    if (const size_t n3 = dh::window_v3_table_elems(g)) {
      DH_TRY(h.w3Tab.ensure(n3), "alloc");
      DH_TRY(dh::launch_window_v3_tables(g, h.w3Tab.p, h.stream), "v3 tables");
    }
    if (g.G > 0) {
      DH_TRY(h.hgx16.ensure(dh::hg_split_elems(g, g.fw)), "alloc");
      DH_TRY(h.hgy16.ensure(dh::hg_split_elems(g, g.fh)), "alloc");
      DH_TRY(dh::launch_hg_split(g, h.hgx.p, h.hgy.p, h.hgx16.p, h.hgy16.p, h.stream), "split basis");
    }
    DH_TRY(hipStreamSynchronize(h.stream), "tables sync");   // the host mask vector goes out of scope
    h.tabKey = d;
    h.tablesValid = true;
    ++h.tabVersion;
  }
  if (timing) DH_TRY(hipEventRecord(ev[1], h.stream), "event");
  if (step == 1) {
    DH_TRY(h.rowspec.ensure(nFrames * P * g.H * g.w), "alloc");
    DH_TRY(dh::launch_row_fft(gA, *pRow, h.twW.p, dframes, h.frameU16, h.rowspec.p, h.stream), "row FFT");
    DH_TRY(dh::launch_col_fft_full(gA, *pCol, h.rowspec.p, h.stream), "column FFT");
    h.stepA = A;
    h.stepMode = c.avgMode;
    h.stepCal = cal;
    if (timing)
      for (int k = 2; k < 6; ++k) DH_TRY(hipEventRecord(ev[k], h.stream), "event");
    h.stepState = 1;
    h.stepD = d;
    h.stepFrames = dframes;
    h.stepU16 = h.frameU16;
    return DIGHOLO_ERROR_SUCCESS;
  }
  h.stepState = 0;
  dh::FusedArgs fa;
  const unsigned* peak = nullptr;   // field peaks left by the staged kernels for the split overlap
  int path = 0;
  if (A > 1) path = 0;   // averaging runs on the staged kernels
  else if ((h.windowMode >= 3 || (h.preferV3 && !h.tuning)) && h.frameU16 && dh::window_v3_table_elems(g) &&
           plan_window_v3(h, g, fa))
    path = 3;
  else if (cal.p) path = 0;   // only the specialised and the staged kernels apply a calibration
  else if (h.windowMode >= 2 && h.frameU16 && plan_window_mfma(h, g, fa)) path = 2;
  else if (h.windowMode >= 1 && plan_fused(h, g, fa)) path = 1;
  if (path) {
    // one kernel from camera window to field
    // the overlap runs inside the specialised window kernel in either precision
    const bool split = c.precision == DIGHOLO_PRECISION_SPLIT16;
    const bool fuseOv = path == 3 && (h.fusedOverlap > 0 || (h.fusedOverlap < 0 && split)) &&
                        dh::window_v3_overlap_fits(g);
    h.lastFusedOverlap = fuseOv;
    if (path == 3) {
      const dh::W3Launch wl{fa,
                            dframes,
                            h.w3Tab.p,
                            h.factor.p,
                            h.ref.p,
                            h.fields.p,
                            fuseOv && split ? h.hgx16.p : nullptr,
                            fuseOv && split ? h.hgy16.p : nullptr,
                            fuseOv ? h.coefs.p : nullptr,
                            cal,
                            fuseOv && !split ? h.hgx.p : nullptr,
                            fuseOv && !split ? h.hgy.p : nullptr,
                            h.rowMode};
      if (h.collect && c.basisType == DIGHOLO_BASISTYPE_HG && dh::window_v3_multi_supported(wl)) {
        h.pend = wl;
        h.pendValid = true;
      } else {
        DH_TRY(dh::launch_window_v3(wl.fa, wl.frames, wl.tab, wl.factor, wl.ref, wl.fields, wl.hgx16, wl.hgy16,
                                    wl.coefs, wl.cal, h.stream, wl.hgx32, wl.hgy32, wl.rowMode),
               "fused window (register rows)");
      }
    } else if (path == 2)
      DH_TRY(dh::launch_window_mfma(fa, h.twW.p, dframes, h.colTab.p, h.factor.p, h.ref.p, h.fields.p, h.stream),
             "fused window (MFMA columns)");
    else
      DH_TRY(dh::launch_fused_window(fa, h.twW.p, pCol->tw, dframes, h.frameU16, h.factor.p, h.ref.p, h.fields.p,
                                     h.stream), "fused window");
    h.lastPath = path;
    if (timing) {
      DH_TRY(hipEventRecord(ev[2], h.stream), "event");
      DH_TRY(hipEventRecord(ev[3], h.stream), "event");
      DH_TRY(hipEventRecord(ev[4], h.stream), "event");
    }
  } else if (A == 1 && !fullRes && dh::staged_ip_supported(g, *pRow, *pCol, *pIx, *pIy)) {
    // staged kernels on in-place transforms: rows, then columns + window + inverse along y,
    // then inverse along x + reference wave
    DH_TRY(h.rowspec.ensure(nFrames * P * g.H * g.w), "alloc");
    DH_TRY(h.win.ensure(B * P * g.h * g.w), "alloc");
    DH_TRY(dh::launch_row_fft_ip(gA, *pRow, h.twW.p, dframes, h.frameU16, h.rowspec.p, h.stream), "row FFT");
    if (timing) DH_TRY(hipEventRecord(ev[2], h.stream), "event");
    // the split overlap's per-field scale comes from the field kernel
    if (g.G > 0 && c.precision == DIGHOLO_PRECISION_SPLIT16) {
      DH_TRY(h.fieldPeak.ensure(B * P), "alloc");
      peak = h.fieldPeak.p;
    }
    unsigned* peakW = peak ? h.fieldPeak.p : nullptr;
    DH_TRY(dh::launch_col_window_iy(g, *pCol, *pIy, h.rowspec.p, h.factor.p, h.win.p, cal, h.stream, peakW),
           "column FFT");
    if (timing) DH_TRY(hipEventRecord(ev[3], h.stream), "event");
    DH_TRY(dh::launch_ifft_rows_tilt(gF, *pIx, h.win.p, h.ref.p, h.fields.p, h.stream, peakW), "IFFT rows");
    if (timing) DH_TRY(hipEventRecord(ev[4], h.stream), "event");
    h.lastPath = 0;
    h.lastFusedOverlap = false;
  } else {
    // staged kernels through HBM intermediates (any window size; averaging)
    DH_TRY(h.rowspec.ensure(nFrames * P * g.H * g.w), "alloc");
    DH_TRY(h.win.ensure(B * P * g.h * g.w), "alloc");
    DH_TRY(dh::launch_row_fft(gA, *pRow, h.twW.p, dframes, h.frameU16, h.rowspec.p, h.stream), "row FFT");
    if (timing) DH_TRY(hipEventRecord(ev[2], h.stream), "event");
    if (A > 1) {
      DH_TRY(h.winAll.ensure(nFrames * P * g.h * g.w), "alloc");
      DH_TRY(dh::launch_col_fft_window(gA, *pCol, h.rowspec.p, h.factor.p, h.winAll.p, dh::BatchCal{}, h.stream),
             "column FFT");
      DH_TRY(dh::launch_average_windows(g, A, c.avgMode, (int)nFrames, h.winAll.p, h.win.p, cal, h.stream),
             "averaging");
    } else {
      DH_TRY(dh::launch_col_fft_window(g, *pCol, h.rowspec.p, h.factor.p, h.win.p, cal, h.stream), "column FFT");
    }
    if (timing) DH_TRY(hipEventRecord(ev[3], h.stream), "event");
    float2* ifftIn = h.win.p;
    if (fullRes) {   // the window at its own frequencies in a zero W x H plane
      DH_TRY(h.plane.ensure(B * P * g.fh * g.fw), "alloc");
      DH_TRY(dh::launch_scatter_full(g, h.win.p, h.plane.p, h.stream), "full-resolution plane");
      ifftIn = h.plane.p;
    }
    DH_TRY(dh::launch_ifft_rows(gF, *pIx, ifftIn, h.stream), "IFFT rows");
    DH_TRY(dh::launch_ifft_cols_tilt(gF, *pIy, ifftIn, h.ref.p, h.fields.p, h.stream), "IFFT columns");
    if (timing) DH_TRY(hipEventRecord(ev[4], h.stream), "event");
    h.lastPath = 0;
    h.lastFusedOverlap = false;
  }
  if (g.G > 0 && !h.lastFusedOverlap) {
    if (h.pendValid) {   // follows the collected window kernel
      h.pendOverlap = true;
      h.pendGF = gF;
      h.pendSplit = c.precision == DIGHOLO_PRECISION_SPLIT16;
    } else {
      DH_TRY(dh::launch_overlap_mfma(gF, h.fields.p, h.hgx.p, h.hgy.p, h.coefs.p,
                                     c.precision == DIGHOLO_PRECISION_SPLIT16, h.stream, h.ovTpw, peak, h.hgx16.p,
                                     h.hgy16.p),
             "overlap");
    }
  }
  if (timing) DH_TRY(hipEventRecord(ev[5], h.stream), "event");

  publish_batch(h, d, dframes, h.frameU16);
  if (int eb = apply_basis(h)) return eb;
  h.coefsDevValid = true;
  return DIGHOLO_ERROR_SUCCESS;
}

// The fields in h.fields now belong to configuration d and frames dframes.
void publish_batch(Handle& h, const Derived& d, const void* dframes, bool u16) {
  const dh::Geometry& g = d.g;
  h.cur = d;
  h.haveBatch = true;
  h.coefsHostValid = false;
  h.fieldsHostValid = false;
  h.lastFrames = dframes;
  h.lastFrameU16 = u16;
  h.analysisValid[0] = h.analysisValid[1] = false;
  h.f16Valid = false;
  // reconstructed-plane axes, float32 like digHolo.cpp:8184-8189 / :2875-2880
  h.xAxis.resize(g.fw);
  h.yAxis.resize(g.fh);
  for (int i = 0; i < g.fw; ++i) h.xAxis[i] = (float)(i - g.fw / 2) * d.pxx;
  for (int j = 0; j < g.fh; ++j) h.yAxis[j] = (float)(j - g.fh / 2) * d.pxy;
}

// IFFT step: window rows of the band times the window factor, 2-D inverse FFT,
// field before tilt removal into h.fields.
int step_ifft(Handle& h) {
  if (h.stepState < 1) return DIGHOLO_ERROR_NULLPOINTER;
  const dh::Geometry& g = h.stepD.g;
  dh::Geometry gF = g;
  gF.w = g.fw;
  gF.h = g.fh;
  const dh::FFTPlan* pIx = fft_plan(h, g.fw);
  const dh::FFTPlan* pIy = fft_plan(h, g.fh);
  if (!pIx || !pIy) return DIGHOLO_ERROR_MEMORYALLOCATION;
  const size_t n = (size_t)g.B * g.P * g.h * g.w;
  const size_t nF = (size_t)g.B * g.P * g.fh * g.fw;
  DH_TRY(h.win.ensure(n), "alloc");
  DH_TRY(h.fields.ensure(nF), "alloc");
  if (h.stepA > 1) {
    dh::Geometry gA = g;
    gA.B = g.B * h.stepA;
    DH_TRY(h.winAll.ensure(n * h.stepA), "alloc");
    DH_TRY(dh::launch_select_window(gA, h.rowspec.p, h.factor.p, h.winAll.p, dh::BatchCal{}, h.stream),
           "window select");
    DH_TRY(dh::launch_average_windows(g, h.stepA, h.stepMode, gA.B, h.winAll.p, h.win.p, h.stepCal, h.stream),
           "averaging");
  } else {
    DH_TRY(dh::launch_select_window(g, h.rowspec.p, h.factor.p, h.win.p, h.stepCal, h.stream), "window select");
  }
  float2* ifftIn = h.win.p;
  if (nF != n || g.fw != g.w) {
    DH_TRY(h.plane.ensure(nF), "alloc");
    DH_TRY(dh::launch_scatter_full(g, h.win.p, h.plane.p, h.stream), "full-resolution plane");
    ifftIn = h.plane.p;
  }
  DH_TRY(dh::launch_ifft_rows(gF, *pIx, ifftIn, h.stream), "IFFT rows");
  DH_TRY(dh::launch_ifft_cols_tilt(gF, *pIy, ifftIn, nullptr, h.fields.p, h.stream), "IFFT columns");
  h.stepState = 2;
  h.haveBatch = false;       // h.fields holds a field that still carries the reference tilt
  h.fieldsHostValid = false;
  h.coefsDevValid = false;
  h.lastPath = 0;
  return DIGHOLO_ERROR_SUCCESS;
}

// RemoveTilt step: fields = -conj(fields * ref); publishes the batch.
int step_remove_tilt(Handle& h) {
  if (h.stepState < 2) return DIGHOLO_ERROR_NULLPOINTER;
  dh::Geometry gF = h.stepD.g;
  gF.w = gF.fw;
  gF.h = gF.fh;
  DH_TRY(dh::launch_apply_tilt(gF, h.ref.p, h.fields.p, h.stream), "tilt");
  h.stepState = 3;
  publish_batch(h, h.stepD, h.stepFrames, h.stepU16);
  h.coefsDevValid = false;
  return DIGHOLO_ERROR_SUCCESS;
}

// BasisExtractCoefs step: the overlap of the published fields.
int step_extract_coefs(Handle& h) {
  if (!h.haveBatch) return DIGHOLO_ERROR_NULLPOINTER;
  dh::Geometry g = h.cur.g;
  g.w = g.fw;
  g.h = g.fh;
  if (g.G <= 0) return DIGHOLO_ERROR_SUCCESS;
  DH_TRY(h.coefs.ensure((size_t)g.B * g.P * g.M), "alloc");
  DH_TRY(dh::launch_overlap_mfma(g, h.fields.p, h.hgx.p, h.hgy.p, h.coefs.p,
                                 h.cfg.precision == DIGHOLO_PRECISION_SPLIT16, h.stream, 0, nullptr, h.hgx16.p,
                                 h.hgy16.p),
         "overlap");
  if (int eb = apply_basis(h)) return eb;
  h.coefsDevValid = true;
  h.coefsHostValid = false;
  return DIGHOLO_ERROR_SUCCESS;
}

// Averages the stage times of every batch enqueued with timing on since the
// last collection (waits for the last one to finish).
void collect_stage_times(Handle& h) {
  if (h.evUsed == 0) return;
  (void)hipEventSynchronize(h.evPool[h.evUsed - 1][5]);
  double acc[5] = {0, 0, 0, 0, 0};
  // stage k spans events (from[k], to[k]): row FFT, column FFT + window,
  // IFFT + tilt, tables (0 when cached), overlap
  const int from[5] = {1, 2, 3, 0, 4}, to[5] = {2, 3, 4, 1, 5};
  for (size_t i = 0; i < h.evUsed; ++i) {
    for (int k = 0; k < 5; ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, h.evPool[i][from[k]], h.evPool[i][to[k]]);
      acc[k] += ms;
    }
  }
  for (int k = 0; k < 5; ++k) h.stageMs[k] = (float)(acc[k] / h.evUsed);
  h.stageCount = (int)h.evUsed;
  h.evUsed = 0;
}

int download_coefs(Handle& h) {
  if (!h.coefsDevValid) return DIGHOLO_ERROR_NULLPOINTER;
  if (h.coefsHostValid) return DIGHOLO_ERROR_SUCCESS;
  const dh::Geometry& g = h.cur.g;
  const size_t n = (size_t)g.B * g.P * h.Mout;
  DH_TRY(h.hCoefs.ensure(std::max<size_t>(n, 1)), "pinned alloc");
  if (n) DH_TRY(hipMemcpyAsync(h.hCoefs.p, h.outCoefs(), n * sizeof(float2), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  h.coefsHostValid = true;
  return DIGHOLO_ERROR_SUCCESS;
}

int download_fields(Handle& h) {
  if (h.fieldsHostValid) return DIGHOLO_ERROR_SUCCESS;
  const dh::Geometry& g = h.cur.g;
  const size_t n = (size_t)g.B * g.P * g.fh * g.fw;
  DH_TRY(h.hFields.ensure(n), "pinned alloc");
  DH_TRY(hipMemcpyAsync(h.hFields.p, h.fields.p, n * sizeof(float2), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  h.fieldsHostValid = true;
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Plane analysis (digHolo.cpp:6931-6984 Fourier plane, :8840-8870 field plane,
// statistics :2354-2654 finished as :9112-9140).  The reference runs both inside
// every batch; here they run when BatchGetSummary/GetFields16 ask, from the
// device frames and fields the last batch left (the staged copy of host frames,
// or the caller's device frames, which must then be unchanged).
// ---------------------------------------------------------------------------
int compute_analysis(Handle& h, int plane) {
  if (!h.haveBatch || !h.devInit) return DIGHOLO_ERROR_NULLPOINTER;
  if (h.analysisValid[plane]) return DIGHOLO_ERROR_SUCCESS;
  const Derived& d = h.cur;
  const dh::Geometry& g = d.g;
  const int B = g.B, P = g.P;
  const Config& c = h.cfg;
  dh::PlaneArgs a{};
  std::vector<float> xs, ys;
  std::vector<uint8_t> inc;
  const float2* src = nullptr;
  size_t planeStride = 0;
  if (plane == 0) {
    const int Nx = g.W / 2 + 1, Ny = g.H;
    dh::Geometry g2 = g;          // the whole R2C plane: every column from kx = 0
    g2.w = Nx;
    for (int p = 0; p < P; ++p) g2.c0[p] = 0;
    const dh::FFTPlan* pRow = fft_plan(h, g.W / 2);
    const dh::FFTPlan* pCol = fft_plan(h, g.H);
    if (!pRow || !pCol || !h.lastFrames) return DIGHOLO_ERROR_NULLPOINTER;
    DH_TRY(h.spec0.ensure((size_t)B * P * Ny * Nx), "alloc");
    DH_TRY(dh::launch_row_fft(g2, *pRow, h.twW.p, h.lastFrames, h.lastFrameU16, h.spec0.p, h.stream), "row FFT");
    DH_TRY(dh::launch_col_fft_full(g2, *pCol, h.spec0.p, h.stream), "column FFT");
    const float kfx = k_factor(g.W, c.pixelSize), kfy = k_factor(g.H, c.pixelSize);
    xs.resize(Nx);
    ys.resize(Ny);
    for (int i = 0; i < Nx; ++i) xs[i] = k_axis(i, g.W, kfx);
    for (int j = 0; j < Ny; ++j) ys[j] = k_axis(j, g.H, kfy);
    // zero order left out: k^2 >= (2 k0 sin r)^2 (digHolo.cpp:6936-6941, :2385-2397)
    const float k0 = 2 * PI_F / c.wavelengthCentre;
    const float R = 2 * k0 * std::sin(c.apertureSize * UNIT_ANGLE_F);
    const float R2 = R * R;
    inc.resize((size_t)Nx * Ny);
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        const float xx = xs[i] * xs[i], yy = ys[j] * ys[j];
        inc[(size_t)j * Nx + i] = (xx + yy) >= R2 ? 1 : 0;
      }
    a.Nx = Nx;
    a.Ny = Ny;
    a.x0 = 1;                      // kx = 0 column out; columns past 8 * floor(Nx / 8) not visited
    a.x1 = 8 * (Nx / 8);
    a.skip0 = 0;                   // ky = 0 and Nyquist rows out
    a.skip1 = Ny / 2;
    src = h.spec0.p;
    planeStride = (size_t)Ny * Nx;
  } else {
    a.Nx = g.fw;
    a.Ny = g.fh;
    a.x0 = 0;
    a.x1 = 8 * (g.fw / 8);
    a.skip0 = a.skip1 = -1;
    xs = h.xAxis;
    ys = h.yAxis;
    src = h.fields.p;
    planeStride = (size_t)g.fh * g.fw;
  }
  const int Nx = a.Nx, Ny = a.Ny, N = Nx * Ny;
  a.inStride = Nx;
  a.norm = (float)(1.0 / ((size_t)Nx * Ny));
  const float yShift = std::fabs((ys[1] - ys[0]) * Ny);
  std::vector<float> axes(Nx + 2 * Ny);
  std::copy(xs.begin(), xs.end(), axes.begin());
  std::copy(ys.begin(), ys.end(), axes.begin() + Nx);
  for (int j = 0; j < Ny; ++j) axes[Nx + Ny + j] = ys[j] + (ys[j] < 0 ? yShift : 0.f);
  DH_TRY(h.anaAxes.ensure(axes.size()), "alloc");
  DH_TRY(hipMemcpyAsync(h.anaAxes.p, axes.data(), axes.size() * sizeof(float), hipMemcpyHostToDevice, h.stream),
         "axes upload");
  a.x = h.anaAxes.p;
  a.y = h.anaAxes.p + Nx;
  a.yw = h.anaAxes.p + Nx + Ny;
  if (!inc.empty()) {
    DH_TRY(h.anaInc.ensure(inc.size()), "alloc");
    DH_TRY(hipMemcpyAsync(h.anaInc.p, inc.data(), inc.size(), hipMemcpyHostToDevice, h.stream), "mask upload");
    a.include = h.anaInc.p;
  }
  const size_t polStride = (size_t)g.W * g.H;     // pol stride of the reference's summed maps
  DH_TRY(h.anaPw.ensure((size_t)B * P * N), "alloc");
  DH_TRY(h.anaSumDev.ensure((size_t)P * polStride), "alloc");
  DH_TRY(h.anaStats.ensure((size_t)(B + 1) * P * 8), "alloc");
  DH_TRY(hipMemsetAsync(h.anaSumDev.p, 0, (size_t)P * polStride * sizeof(float), h.stream), "memset");
  DH_TRY(dh::launch_plane_stats(a, src, nullptr, B * P, planeStride, h.anaPw.p, h.anaStats.p, h.stream), "stats");
  DH_TRY(dh::launch_batch_sum(h.anaPw.p, B, P, N, h.anaSumDev.p, polStride, h.stream), "batch sum");
  DH_TRY(dh::launch_plane_stats(a, nullptr, h.anaSumDev.p, P, polStride, nullptr, h.anaStats.p + (size_t)B * P * 8,
                                h.stream), "stats");
  std::vector<double> st((size_t)(B + 1) * P * 8);
  h.anaSum[plane].assign((size_t)P * polStride, 0.f);
  DH_TRY(hipMemcpyAsync(st.data(), h.anaStats.p, st.size() * sizeof(double), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipMemcpyAsync(h.anaSum[plane].data(), h.anaSumDev.p, h.anaSum[plane].size() * sizeof(float),
                        hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");

  const float dA = std::fabs(xs[2] - xs[1]) * std::fabs(ys[2] - ys[1]);
  std::vector<float>& out = h.analysis[plane];
  out.assign((size_t)7 * P * (B + 1), 0.f);
  auto at = [&](int par, int p, int b) -> float& { return out[((size_t)par * P + p) * (B + 1) + b]; };
  for (int p = 0; p < P; ++p)
    for (int b = 0; b <= B; ++b) {
      const double* r = &st[((b < B ? (size_t)b * P : (size_t)B * P) + p) * 8];
      const float tot = (float)r[0];
      at(0, p, b) = tot;                                        // TOTALPOWER
      at(1, p, b) = (float)r[1] / tot;                          // COMX
      at(2, p, b) = (float)r[2] / tot;                          // COMY
      at(3, p, b) = (float)(32767.0 / std::sqrt((float)r[5]));  // MAXABS (scale to int16 full range)
      const int idx = (int)r[6];
      std::memcpy(&at(4, p, b), &idx, sizeof(int));             // MAXABSIDX: int32 bits
      at(5, p, b) = dA * (tot * tot / (float)r[4]);             // AEFF (Petermann II)
      at(6, p, b) = (float)r[3] / tot;                          // COMYWRAP
    }
  h.anaX[plane] = xs;
  h.anaY[plane] = ys;
  h.analysisValid[plane] = true;
  return DIGHOLO_ERROR_SUCCESS;
}

int compute_fields16(Handle& h) {
  if (h.f16Valid) return DIGHOLO_ERROR_SUCCESS;
  const int e = compute_analysis(h, 1);
  if (e) return e;
  const dh::Geometry& g = h.cur.g;
  const int B = g.B, P = g.P, hw = g.fh * g.fw;
  const size_t n = (size_t)B * P * hw;
  const float* scale = &h.analysis[1][(size_t)3 * P * (B + 1)];
  DH_TRY(h.anaScale.ensure((size_t)P * (B + 1)), "alloc");
  DH_TRY(hipMemcpyAsync(h.anaScale.p, scale, (size_t)P * (B + 1) * sizeof(float), hipMemcpyHostToDevice, h.stream),
         "scale upload");
  DH_TRY(h.f16.ensure(2 * n), "alloc");
  DH_TRY(dh::launch_fields16(h.fields.p, B, P, hw, h.anaScale.p, h.f16.p, h.f16.p + n, h.stream), "fields16");
  DH_TRY(h.hF16.ensure(2 * n), "pinned alloc");
  DH_TRY(hipMemcpyAsync(h.hF16.p, h.f16.p, 2 * n * sizeof(short), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  h.f16Valid = true;
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// power meter (digHolo.cpp:10714-11129): linear metrics of the coefficient
// matrix of the last batch.  pols: -1 both polarisations (the only form the
// C API's AutoAlignCalcMetrics uses).
// ---------------------------------------------------------------------------

// polIdx < 0: both polarisations of the last batch.  polIdx = p: the metric
// seen while optimising pol p alone (digHolo.cpp:10770-10860): its own
// coefficients only when the pols are independent, else its coefficients
// beside the other pol's best-so-far ones from h.coefsBackup.
// One wavelength's metrics over its B / L batch elements: output positions
// lam, lam + L, ... (wavelength the fast axis) or lam B/L ... (slow),
// digHolo.cpp:10770-10790.
int power_meter_lambda(Handle& h, int lam, int polIdx, bool independent, bool doIL, bool doMDL, bool doXtalk,
                       bool mulConj, PowerMeterOut& o) {
  const dh::Geometry& g = h.cur.g;
  const int B = g.B, P = g.P, Mpol = h.Mout;                // the coefficients of the configured basis
  const int L = std::max(1, g.L);
  const float FMAX = std::numeric_limits<float>::max();
  if (B > 1) {
    const bool indep = independent || P == 1;
    int S = (P == 2 && indep) ? 2 : 1;
    int M = (P == 2 && !indep) ? 2 * Mpol : Mpol;          // svdM
    const int N = B / L;                                   // svdN
    const size_t first = g.ordOut ? (size_t)lam * N : (size_t)lam;   // first batch element of this wavelength
    const int step = g.ordOut ? 1 : L;                     // batch elements between its rows
    int rowStride = P * Mpol;
    const float2* A = h.outCoefs();
    if (polIdx >= 0 && P == 2) {
      S = 1;
      if (indep) {
        A = h.outCoefs() + (size_t)polIdx * Mpol;
      } else {
        DH_TRY(h.coefsMix.ensure((size_t)B * 2 * Mpol), "alloc");
        DH_TRY(hipMemcpy2DAsync(h.coefsMix.p, 2 * Mpol * sizeof(float2), h.outCoefs() + (size_t)polIdx * Mpol,
                                rowStride * sizeof(float2), Mpol * sizeof(float2), B, hipMemcpyDeviceToDevice,
                                h.stream), "mix");
        DH_TRY(hipMemcpy2DAsync(h.coefsMix.p + Mpol, 2 * Mpol * sizeof(float2),
                                h.coefsBackup.p + (size_t)(1 - polIdx) * Mpol, rowStride * sizeof(float2),
                                Mpol * sizeof(float2), B, hipMemcpyDeviceToDevice, h.stream), "mix");
        A = h.coefsMix.p;
        rowStride = 2 * Mpol;
      }
    }
    A += first * rowStride;
    rowStride *= step;
    DH_TRY(h.metricRows.ensure((size_t)S * N * 5), "alloc");
    DH_TRY(dh::launch_metric_rows(A, N, M, Mpol, rowStride, S, h.metricRows.p, h.stream), "metric rows");
    std::vector<double> rows((size_t)S * N * 5);
    DH_TRY(hipMemcpyAsync(rows.data(), h.metricRows.p, rows.size() * sizeof(double), hipMemcpyDeviceToHost,
                          h.stream), "D2H");
    DH_TRY(hipStreamSynchronize(h.stream), "sync");

    // Gram matrices (FP64) when the spectrum or A^H A is needed
    const bool rowsVec = N <= M;
    const int dim = rowsVec ? N : M;
    std::vector<std::vector<std::complex<double>>> grams;
    if (doMDL || (doXtalk && mulConj)) {
      DH_TRY(h.gram.ensure((size_t)dim * dim), "alloc");
      for (int s = 0; s < S; ++s) {
        DH_TRY(dh::launch_gram_f64(A, N, M, rowStride, s, rowsVec ? 1 : 0, h.gram.p, h.stream), "gram");
        std::vector<std::complex<double>> G((size_t)dim * dim);
        DH_TRY(hipMemcpyAsync(G.data(), h.gram.p, G.size() * sizeof(double2), hipMemcpyDeviceToHost, h.stream),
               "D2H");
        DH_TRY(hipStreamSynchronize(h.stream), "sync");
        grams.push_back(std::move(G));
      }
    }

    double totalSignal = 0, totalNoise = 0, totalSNR = 0, sigMG = 0, noiMG = 0;
    double minSignal = FMAX, maxSignal = -FMAX, minSNR = FMAX, maxSNR = -FMAX;
    double signalCount = 1;
    if (doXtalk) {
      for (int s = 0; s < S; ++s) {
        if (mulConj && rowsVec) {
          // crosstalk of A^H A (N x N), digHolo.cpp:10952-10967
          const std::vector<std::complex<double>>& G = grams[s];
          signalCount = (double)N * N * S;
          for (int i = 0; i < N && i < Mpol; ++i) {
            double sig = 0, noi = 0;
            int gi = 0;          // HG group of mode i
            while ((gi + 1) * (gi + 2) / 2 <= i) ++gi;
            for (int j = 0; j < N; ++j) {
              const double pw = std::norm(G[(size_t)i * N + j]);
              if (i == j || i == j - Mpol) sig += pw;
              else noi += pw;
              int gj = 0;
              const int jm = j % Mpol;
              while ((gj + 1) * (gj + 2) / 2 <= jm) ++gj;
              if (gi == gj) sigMG += pw / (gi + 1);
              else noiMG += pw / (gi + 1);
            }
            const double snr = sig / noi;
            totalSignal += sig;
            totalNoise += noi;
            maxSignal = std::max(maxSignal, sig);
            minSignal = std::min(minSignal, sig);
            maxSNR = std::max(maxSNR, snr);
            minSNR = std::min(minSNR, snr);
          }
        } else {
          signalCount = (double)M * N * S;
          for (int i = 0; i < N && i < Mpol; ++i) {
            const double* r = &rows[((size_t)s * N + i) * 5];
            const double sig = r[0], noi = r[1];
            sigMG += r[2];
            noiMG += r[3];
            const double snr = sig / noi;
            totalSignal += sig;
            totalNoise += noi;
            maxSignal = std::max(maxSignal, sig);
            minSignal = std::min(minSignal, sig);
            maxSNR = std::max(maxSNR, snr);
            minSNR = std::min(minSNR, snr);
          }
        }
      }
      totalSNR = totalSignal / totalNoise;
    }
    double il = 0, mdl = 0;
    if (doMDL) {
      double smin = FMAX, smax = -FMAX, sum = 0;
      for (int s = 0; s < S; ++s) {
        const dh::HermitianSpectrum sp = dh::hermitian_extremes(dim, grams[s]);
        smin = std::min(smin, std::max(sp.min, 0.0));
        smax = std::max(smax, sp.max);
        sum += sp.trace;
      }
      il = sum / ((double)dim * S);                 // mean of s^2 over min(M,N) * S values
      mdl = smin / smax;                            // (s_min / s_max)^2
    } else if (doIL) {
      for (int s = 0; s < S; ++s) {
        double tot = 0;
        for (int i = 0; i < N; ++i) tot += rows[((size_t)s * N + i) * 5 + 4];
        il += tot / N;                              // insertionLoss256, digHolo.cpp:3105-3143
      }
    }
    o.v[DIGHOLO_METRIC_IL] = il;
    o.v[DIGHOLO_METRIC_MDL] = mdl;
    o.v[DIGHOLO_METRIC_DIAG] = totalSignal / signalCount;
    o.v[DIGHOLO_METRIC_SNRAVG] = totalSNR;
    o.v[DIGHOLO_METRIC_DIAGBEST] = maxSignal;
    o.v[DIGHOLO_METRIC_DIAGWORST] = minSignal;
    o.v[DIGHOLO_METRIC_SNRBEST] = maxSNR;
    o.v[DIGHOLO_METRIC_SNRWORST] = minSNR;
    o.v[DIGHOLO_METRIC_SNRMG] = sigMG / noiMG;
  } else {
    // one frame: total power and the power of mode 0 (digHolo.cpp:11093-11124)
    std::vector<float2> c((size_t)P * Mpol);
    DH_TRY(hipMemcpyAsync(c.data(), h.outCoefs(), c.size() * sizeof(float2), hipMemcpyDeviceToHost, h.stream),
           "D2H");
    DH_TRY(hipStreamSynchronize(h.stream), "sync");
    double tot = 0, diag = 0;
    for (int p = 0; p < P; ++p)
      for (int m = 0; m < Mpol; ++m) {
        const float2 v = c[(size_t)p * Mpol + m];
        const double pw = (double)v.x * v.x + (double)v.y * v.y;
        tot += pw;
        if (m == 0) diag += pw;
      }
    o.v[DIGHOLO_METRIC_IL] = tot;
    o.v[DIGHOLO_METRIC_MDL] = 0;
    o.v[DIGHOLO_METRIC_DIAG] = diag;
    o.v[DIGHOLO_METRIC_SNRAVG] = FMAX;
    o.v[DIGHOLO_METRIC_DIAGBEST] = diag;
    o.v[DIGHOLO_METRIC_DIAGWORST] = diag;
    o.v[DIGHOLO_METRIC_SNRBEST] = FMAX;
    o.v[DIGHOLO_METRIC_SNRWORST] = FMAX;
    o.v[DIGHOLO_METRIC_SNRMG] = FMAX;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// Metrics of every wavelength into h.pmLam (float, as the reference keeps
// them) and their linear average into o (digHolo.cpp:12172-12191).
int power_meter(Handle& h, int polIdx, bool independent, bool doIL, bool doMDL, bool doXtalk, bool mulConj,
                PowerMeterOut& o) {
  const dh::Geometry& g = h.cur.g;
  const int L = std::max(1, g.L);
  if (h.Mout == 0) return DIGHOLO_ERROR_INVALIDARGUMENT;   // needs the Fourier-plane power (not built)
  if (g.B < L) return DIGHOLO_ERROR_INVALIDDIMENSION;      // no wavelength has a batch element
  h.pmLam.assign(L, PowerMeterOut{});
  float sum[DIGHOLO_METRIC_COUNT] = {};
  for (int l = 0; l < L; ++l) {
    PowerMeterOut& ol = h.pmLam[l];
    if (int e = power_meter_lambda(h, l, polIdx, independent, doIL, doMDL, doXtalk, mulConj, ol)) return e;
    for (int m = 0; m < DIGHOLO_METRIC_COUNT; ++m) {
      ol.v[m] = (float)ol.v[m];
      sum[m] += (float)ol.v[m];
    }
  }
  for (int m = 0; m < DIGHOLO_METRIC_COUNT; ++m) o.v[m] = L == 1 ? h.pmLam[0].v[m] : (double)(sum[m] / L);
  return DIGHOLO_ERROR_SUCCESS;
}

// metrics in dB: out[m][l] per wavelength and the wavelength average
// out[m][L] (digHolo.cpp:12183-12191); h.pmLam holds the last power_meter call
void store_metrics_db(Handle& h, const PowerMeterOut& o) {
  const int L = std::max<int>(1, (int)h.pmLam.size());
  h.aaLambdaCount = L;
  h.aaMetrics.assign((size_t)DIGHOLO_METRIC_COUNT * (L + 1), 0.f);
  for (int m = 0; m < DIGHOLO_METRIC_COUNT; ++m) {
    float sum = 0.f;
    for (int l = 0; l < L; ++l) {
      const float lin = h.pmLam.empty() ? (float)o.v[m] : (float)h.pmLam[l].v[m];
      sum += lin;
      h.aaMetrics[(size_t)m * (L + 1) + l] = 10.f * std::log10(lin);
    }
    h.aaMetrics[(size_t)m * (L + 1) + L] = 10.f * std::log10(sum / L);
  }
}

// ---------------------------------------------------------------------------
// AutoAlign, tweak stage (digHolo.cpp:12315-12885).  One parameter at a time,
// in the order waist, defocus, tilt x, centre x, tilt y, centre y: three
// samples at -step/2, 0, +step/2 about the current value, then three samples
// at the vertex of the least-squares quadratic through every sample so far
// (held within 10 steps of the best value).  Every sample reprocesses the
// whole batch on the GPU and evaluates the goal metric there; unlocked
// parameters of the two polarisations are searched side by side, each judged
// with the other pol's best-so-far coefficients.  Attempts repeat until the
// best goal value moves by at most the tolerance (or 20 attempts).
// ---------------------------------------------------------------------------
enum { AA_WAIST = 0, AA_DEFOCUS = 1, AA_TILTX = 2, AA_CX = 3, AA_TILTY = 4, AA_CY = 5, AA_PARS = 6 };

float& aa_param(Config& c, int p, int par) {               // digHolo.cpp:13021-13110
  switch (par) {
    case AA_WAIST: return c.waist[p];
    case AA_DEFOCUS: return c.defocus[p];
    case AA_TILTX: return c.tiltX[p];
    case AA_CX: return c.beamCX[p];
    case AA_TILTY: return c.tiltY[p];
    default: return c.beamCY[p];
  }
}

// Largest window radius any tilt can use without aliasing (digHolo.cpp:11964-11971)
float window_radius_max(float px, float lambda) {
  const float wc = (float)((-2.0 + std::sqrt(68.0)) / 16.0);
  const float wmax = (float)(lambda / (2.0 * px));
  return (wc * wmax) / UNIT_ANGLE_F;
}
// A third of the smallest tilt magnitude over the pols (digHolo.cpp:11978-11993)
float window_radius_rec(const Config& c) {
  float m = std::numeric_limits<float>::max();
  for (int p = 0; p < c.polCount; ++p) m = std::min(m, std::sqrt(c.tiltX[p] * c.tiltX[p] + c.tiltY[p] * c.tiltY[p]));
  return m / 3.0f;
}

// Vertex -b/(2a) of the least-squares quadratic a x^2 + b x + c through the
// samples, or NaN when the fit is degenerate.  The reference iterates
// Gauss-Newton steps on this model (digHolo.cpp:3146-3257); the model is
// linear in a, b, c, so its fixed point is this closed-form solution, solved
// here in double on centred, scaled abscissae.
double quadratic_vertex(const double* x, const double* y, int n) {
  double xm = 0, s = 0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(x[i]) || !std::isfinite(y[i])) return NAN;
    xm += x[i];
  }
  xm /= n;
  for (int i = 0; i < n; ++i) s = std::max(s, std::fabs(x[i] - xm));
  if (!(s > 0)) return NAN;
  double S0 = 0, S1 = 0, S2 = 0, S3 = 0, S4 = 0, T0 = 0, T1 = 0, T2 = 0;
  for (int i = 0; i < n; ++i) {
    const double u = (x[i] - xm) / s, u2 = u * u;
    S0 += 1; S1 += u; S2 += u2; S3 += u2 * u; S4 += u2 * u2;
    T0 += y[i]; T1 += y[i] * u; T2 += y[i] * u2;
  }
  // normal equations [S4 S3 S2; S3 S2 S1; S2 S1 S0] [a b c]' = [T2 T1 T0]', Cramer's rule
  auto det3 = [](double a, double b, double c, double d, double e, double f, double g, double hh, double i) {
    return a * (e * i - f * hh) - b * (d * i - f * g) + c * (d * hh - e * g);
  };
  const double D = det3(S4, S3, S2, S3, S2, S1, S2, S1, S0);
  if (D == 0 || !std::isfinite(D)) return NAN;
  const double a = det3(T2, S3, S2, T1, S2, S1, T0, S1, S0) / D;
  const double b = det3(S4, T2, S2, S3, T1, S1, S2, T0, S0) / D;
  if (a == 0) return NAN;
  return xm + s * (-b / (2 * a));
}

void to_db(const PowerMeterOut& o, float* db) {
  for (int m = 0; m < DIGHOLO_METRIC_COUNT; ++m) db[m] = 10.f * std::log10((float)o.v[m]);
}

struct HoldGuard {     // frames upload once for the whole search; holds end with it
  Handle& h;
  explicit HoldGuard(Handle& hh) : h(hh) { h.holdStaging = true; }
  ~HoldGuard() { h.holdStaging = h.holdCrop = h.holdWin = false; }
};

// snapSteps: the step sizes the parameter estimate chose (FULL mode); null in
// TWEAK mode, where they follow from the current values (digHolo.cpp:12303-12320).
int autoalign_tweak(Handle& h, float& goalDb, const float* snapSteps) {
  Config& c = h.cfg;
  const int P = c.polCount;
  const int goal0 = c.aaGoalIdx;
  const bool indep = c.aaPolIndependence != 0, mulConj = c.aaMulConjTrans != 0;
  const float tol = c.aaTol;
  const float FMAX = std::numeric_limits<float>::max();
  HoldGuard hold(h);
  bool processed = false;   // the first evaluation fixes the FFT crops (digHolo.cpp:12620-12625)

  float baseStep[AA_PARS];
  if (snapSteps) {
    std::copy(snapSteps, snapSteps + AA_PARS, baseStep);
  } else {
    baseStep[AA_TILTX] = baseStep[AA_TILTY] = c.apertureSize * 0.1f;
    baseStep[AA_CX] = baseStep[AA_CY] = c.waist[0] * 0.1f;
    baseStep[AA_WAIST] = c.waist[0] * 0.1f;
    baseStep[AA_DEFOCUS] = 0.4f;
    for (int p = 0; p < P; ++p) baseStep[AA_DEFOCUS] = std::max(baseStep[AA_DEFOCUS], std::fabs(c.defocus[p]) * 0.1f);
  }

  // bounds on what may be tried (digHolo.cpp:12336-12359)
  float lim[AA_PARS][2];
  const float tiltMaxX = (c.wavelengthCentre / (2 * c.pixelSize)) / UNIT_ANGLE_F;
  lim[AA_TILTX][0] = 0;
  lim[AA_TILTX][1] = tiltMaxX;
  lim[AA_TILTY][0] = -(c.apertureSize + tiltMaxX);
  lim[AA_TILTY][1] = c.apertureSize + tiltMaxX;
  lim[AA_DEFOCUS][0] = -100;
  lim[AA_DEFOCUS][1] = 100;
  lim[AA_WAIST][0] = c.pixelSize / 3.0f;
  lim[AA_WAIST][1] = c.pixelSize * (float)std::max(c.frameHeight, c.frameWidth);
  lim[AA_CX][0] = -(c.frameWidth / 2.0f) * c.pixelSize;
  lim[AA_CX][1] = (c.frameWidth / 2.0f) * c.pixelSize;
  lim[AA_CY][0] = -(c.frameHeight / 2.0f) * c.pixelSize;
  lim[AA_CY][1] = (c.frameHeight / 2.0f) * c.pixelSize;

  float stepBound[AA_PARS];
  for (int i = 0; i < AA_PARS; ++i) stepBound[i] = baseStep[i] * 10;

  float best[2][AA_PARS];
  auto save = [&](int p) { for (int i = 0; i < AA_PARS; ++i) best[p][i] = aa_param(c, p, i); };
  auto load = [&](int p) { for (int i = 0; i < AA_PARS; ++i) aa_param(c, p, i) = best[p][i]; };
  save(0);
  save(1);

  constexpr int sampleCount = 3, stepCount = 5;
  float outBest[2][DIGHOLO_METRIC_COUNT];
  for (auto& r : outBest) std::fill(r, r + DIGHOLO_METRIC_COUNT, -FMAX);
  PowerMeterOut outf{};
  float stepSize[AA_PARS];
  std::copy(baseStep, baseStep + AA_PARS, stepSize);
  const int goalCount = goal0 ? 2 : 1;
  float lastErrTol = goal0 ? 0.1f + tol * 10.0f : tol;

  // device copies of coefficients (digHolo.cpp:12632-12637, :12692-12699)
  auto backup_all = [&]() -> int {
    const dh::Geometry& g = h.cur.g;
    const size_t n = (size_t)g.B * g.P * h.Mout;
    DH_TRY(h.coefsBackup.ensure(std::max<size_t>(n, 1)), "alloc");
    DH_TRY(hipMemcpyAsync(h.coefsBackup.p, h.outCoefs(), n * sizeof(float2), hipMemcpyDeviceToDevice, h.stream),
           "copy");
    return 0;
  };
  auto process = [&]() -> int {
    const int e = enqueue_batch(h);
    if (e) return e;
    if (!processed) {
      h.heldCrop = h.cur;
      h.holdCrop = true;
      processed = true;
    }
    return 0;
  };
  auto backup_pol = [&](int p) -> int {
    const dh::Geometry& g = h.cur.g;
    const size_t pitch = (size_t)g.P * h.Mout * sizeof(float2);
    DH_TRY(hipMemcpy2DAsync(h.coefsBackup.p + (size_t)p * h.Mout, pitch, h.outCoefs() + (size_t)p * h.Mout, pitch,
                            (size_t)h.Mout * sizeof(float2), g.B, hipMemcpyDeviceToDevice, h.stream), "copy");
    return 0;
  };

  for (int gp = 0; gp < goalCount; ++gp) {
    const int goal = gp ? goal0 : 0;
    const bool doIL = goal == 0, doMDL = goal == 1, doXtalk = goal >= 2;
    bool converged = false, firstRun = true;
    float bestPwr[2] = {-FMAX, -FMAX};
    double lastErr[2] = {FMAX, FMAX};
    outBest[0][goal] = outBest[1][goal] = -FMAX;
    int attempts = 0;
    if (gp) {    // a second pass for a goal other than IL starts near the answer: finer steps
      for (int i = 0; i < AA_PARS; ++i) stepSize[i] = baseStep[i] / 10.0f;
      lastErrTol = tol;
    }
    while (!converged) {
      bool enabled[AA_PARS];
      enabled[AA_TILTX] = enabled[AA_TILTY] = c.aaTilt != 0;
      enabled[AA_CX] = enabled[AA_CY] = c.aaCentre != 0;
      enabled[AA_WAIST] = c.aaWaist != 0;
      enabled[AA_DEFOCUS] = c.aaDefocus != 0;
      for (int par = 0; par < AA_PARS; ++par) {
        if (!enabled[par]) continue;
        const bool lock = ((par == AA_TILTX || par == AA_TILTY) && c.polLockTilt) ||
                          (par == AA_DEFOCUS && c.polLockDefocus) || (par == AA_WAIST && c.polLockWaist);
        float var0[2];
        var0[0] = aa_param(c, 0, par);
        var0[1] = lock ? var0[0] : aa_param(c, 1, par);
        float maxPwr[2] = {-FMAX, -FMAX};
        // a tilt sweep only re-applies the reference wave: the Fourier window of
        // the last full evaluation stays (digHolo.cpp:12626-12645)
        const bool tiltPar = par == AA_TILTX || par == AA_TILTY;
        if (tiltPar && processed) {
          h.heldWin = h.cur;
          h.holdWin = true;
        }
        float dTxs[stepCount + 2][2] = {};
        double xs[2][stepCount + 1], ys[2][stepCount + 1];
        for (int step = 0; step <= stepCount; ++step) {
          float* dTx = dTxs[step];
          bool valid[2] = {true, true};
          for (int p = 0; p < P; ++p) {
            if (step < sampleCount)
              dTx[p] = (float)(var0[p] + ((step - (sampleCount - 1) / 2.0) * stepSize[par] / (sampleCount - 1)));
            if (lock) dTx[1] = dTx[0];
            dTx[p] = std::min(std::max(dTx[p], lim[par][0]), lim[par][1]);
            valid[p] = std::isfinite(dTx[p]);
            if (valid[p]) aa_param(c, p, par) = dTx[p];
          }
          if (c.aaFourierRadius) c.apertureSize = std::min(c.apertureSize, window_radius_rec(c));
          int e = process();
          if (e) return e;
          if (tiltPar && !h.holdWin) {
            h.heldWin = h.cur;
            h.holdWin = true;
          }
          if (firstRun) {
            if ((e = backup_all())) return e;
            firstRun = false;
          }
          float out[2][DIGHOLO_METRIC_COUNT];
          PowerMeterOut o;
          if ((!lock && P > 1) || indep) {
            for (int p = 0; p < P; ++p) {
              if ((e = power_meter(h, p, indep, doIL, doMDL, doXtalk, mulConj, o))) return e;
              to_db(o, out[p]);
            }
          } else {
            if ((e = power_meter(h, -1, indep, doIL, doMDL, doXtalk, mulConj, o))) return e;
            to_db(o, out[0]);
            std::copy(out[0], out[0] + DIGHOLO_METRIC_COUNT, out[1]);
          }
          double vx0 = NAN;
          for (int p = 0; p < P; ++p) {
            const float pwr = out[p][goal];
            if (c.verbosity >= DIGHOLO_VERBOSITY_COOKED)
              fprintf(stdout, "AutoAlign sample  parameter %d  step %d  pol %d  value %.9g  goal %.7f\n", par, step, p,
                      (double)dTx[p], (double)pwr);
            xs[p][step] = dTx[p];
            ys[p][step] = pwr;
            if (pwr > maxPwr[p] && valid[p]) {
              maxPwr[p] = pwr;
              if (pwr > bestPwr[p]) {
                save(p);
                std::copy(out[p], out[p] + DIGHOLO_METRIC_COUNT, outBest[p]);
                if ((e = backup_pol(p))) return e;
                bestPwr[p] = pwr;
              }
            }
            if (step >= sampleCount - 1) {
              const double vx = quadratic_vertex(xs[p], ys[p], step + 1);
              if (p == 0) vx0 = vx;
              double next;
              if (std::isfinite(vx0)) {   // pol 0's fit decides for every pol, as the reference's test (:12752)
                const double d = vx - best[p][par];
                next = std::fabs(d) > stepBound[par] ? best[p][par] + (d > 0 ? stepBound[par] : -stepBound[par]) : vx;
              } else {   // no usable fit: a point near the best so far
                next = best[p][par] + ((step - (stepCount - 1) / 2.0) * stepSize[par] / (stepCount - 1));
              }
              dTxs[step + 1][p] = (float)next;
            }
          }
        }
        h.holdWin = false;
        for (int p = 0; p < P; ++p) load(p);
        int e = process();
        if (e) return e;
      }
      if (firstRun) {    // every parameter switched off
        int e = process();
        if (e || (e = backup_all())) return e;
        firstRun = false;
      }
      int e = power_meter(h, -1, indep, true, true, true, mulConj, outf);
      if (e || (e = backup_all())) return e;
      double dErr = 0;
      for (int p = 0; p < P; ++p) {
        dErr += std::fabs((double)outBest[p][goal] - lastErr[p]);
        lastErr[p] = outBest[p][goal];
      }
      dErr /= P;
      ++attempts;
      if (c.verbosity >= DIGHOLO_VERBOSITY_DEBUG) {
        float db[DIGHOLO_METRIC_COUNT];
        to_db(outf, db);
        fprintf(stdout, "AutoAlign attempt %d  goal %f  IL %f  MDL %f  dErr %f\n", attempts, db[goal0],
                db[DIGHOLO_METRIC_IL], db[DIGHOLO_METRIC_MDL], dErr);
        for (int p = 0; p < P; ++p)
          fprintf(stdout, "  pol %d  tilt %.4f %.4f  defocus %.4f  waist %.3f um  centre %.3f %.3f um\n", p,
                  best[p][AA_TILTX], best[p][AA_TILTY], best[p][AA_DEFOCUS], best[p][AA_WAIST] * 1e6,
                  best[p][AA_CX] * 1e6, best[p][AA_CY] * 1e6);
        fflush(stdout);
      }
      if (dErr <= lastErrTol || attempts == 20) converged = true;
    }
  }
  // leave the handle as a plain ProcessBatch at the final parameters would:
  // FFT crops follow the final beam centres, and the metrics are theirs
  h.holdCrop = h.holdWin = false;
  int e = enqueue_batch(h);
  if (e || (e = power_meter(h, -1, indep, true, true, true, mulConj, outf))) return e;
  store_metrics_db(h, outf);
  goalDb = h.aaMetrics[(size_t)goal0 * (h.aaLambdaCount + 1) + h.aaLambdaCount];
  if (c.verbosity >= DIGHOLO_VERBOSITY_BASIC) {
    fprintf(stdout, "AutoAlign goal %f  IL %f\n", goalDb, h.aaMetrics[h.aaLambdaCount]);
    fflush(stdout);
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// AutoAlign parameter estimate (digHolo.cpp:11517-11930) and focus search
// (:11181-11515), the first stage of the FULL and ESTIMATE modes.  From a clean
// slate (centres at the middle of each pol's half of the frame, tilt and
// defocus zero, FFT window the whole half frame):
//   1. the tilt of each pol is the peak of its batch-summed Fourier intensity
//      correlated with a Gaussian of the window's size (DCT-domain filter on
//      the GPU, dh_align.hip), outside twice the window radius about the zero
//      order; with the window radius aligned too, it shrinks to a third of the
//      tilt and the search runs a second pass;
//   2. the effective area of the Fourier power inside the window gives the
//      beam's far-field size, the centre of mass and effective area of the
//      reconstructed fields give its centre and waist (fitted relation
//      AeffToMFD), and the ratio of the two sizes the scale of the defocus;
//   3. the defocus minimising the effective area of the far field of the
//      tilt-removed fields is searched with a quadratic fit (AutoFocus).
// Every quantity is computed on the GPU; the host holds the scalars.
// ---------------------------------------------------------------------------

// Mode-field diameter from the effective area of a beam filling `groups` HG
// groups: the reference's fitted relation (digHolo.cpp:2887-2900)
float aeff_to_mfd(float Aeff, int groups) {
  if (!groups) return 0.f;
  const float A = (float)(Aeff / (15.1059697707643 * groups + 6.65360122093789));
  return (float)((double)(2.0f * std::sqrt(A / PI_F)) / 0.223008930683135);
}

// Defocus magnitude from the beam size in the camera and Fourier planes
// (digHolo.cpp:12040-12075); 0 when the ratio is not physical.
float focus_estimate(float wCamera, float wFourier) {
  if (!(wCamera > 0 && wFourier > 0)) return 0.f;
  const float ratio = wFourier / wCamera;
  const float a = 0.25f, b = PI_F / 2;
  float r = 1.0f * 2 * (b * std::sqrt((1 - a) / (ratio - a) - 1));
  if (!(r > 0) || std::isnan(r)) r = 0.f;
  return r;
}

// Batch-summed Fourier-plane intensity of pol p into h.snapMap (the analysis of
// digHolo.cpp:2354-2654 on the R2C plane in h.spec0, [nB][P][Ny][Nx]): pixels
// with |k - wc|^2 >= R^2 (R > 0), <= R^2 (R < 0) or all (R == 0); the kx = 0
// column, the ky = 0 and Nyquist rows and the columns past 8 floor(Nx / 8) are
// not visited.  aeff = the effective area of the summed map.
int fourier_pol_map(Handle& h, const dh::Geometry& g, int nB, int p, float wcx, float wcy, float R, float& aeff) {
  const Config& c = h.cfg;
  const int Nx = g.W / 2 + 1, Ny = g.H, N = Nx * Ny, P = g.P;
  const float kfx = k_factor(g.W, c.pixelSize), kfy = k_factor(g.H, c.pixelSize);
  std::vector<float> xs(Nx), ys(Ny);
  for (int i = 0; i < Nx; ++i) xs[i] = k_axis(i, g.W, kfx);
  for (int j = 0; j < Ny; ++j) ys[j] = k_axis(j, g.H, kfy);
  const float sign = R >= 0 ? 1.f : -1.f;
  const float w2 = R * R;
  std::vector<uint8_t> inc((size_t)N);
  for (int j = 0; j < Ny; ++j) {
    const float dy = ys[j] + (-wcy);
    const float dy2 = dy * dy;
    for (int i = 0; i < Nx; ++i) {
      const float dx = xs[i] + (-wcx);
      const float r2 = sign * (dx * dx + dy2);
      inc[(size_t)j * Nx + i] = r2 >= sign * w2 ? 1 : 0;
    }
  }
  std::vector<float> axes(Nx + 2 * Ny);
  std::copy(xs.begin(), xs.end(), axes.begin());
  std::copy(ys.begin(), ys.end(), axes.begin() + Nx);
  std::copy(ys.begin(), ys.end(), axes.begin() + Nx + Ny);   // COM y not used here
  DH_TRY(h.anaAxes.ensure(axes.size()), "alloc");
  DH_TRY(h.anaInc.ensure(inc.size()), "alloc");
  DH_TRY(hipMemcpyAsync(h.anaAxes.p, axes.data(), axes.size() * sizeof(float), hipMemcpyHostToDevice, h.stream),
         "axes upload");
  DH_TRY(hipMemcpyAsync(h.anaInc.p, inc.data(), inc.size(), hipMemcpyHostToDevice, h.stream), "mask upload");
  dh::PlaneArgs a{};
  a.Nx = Nx;
  a.Ny = Ny;
  a.inStride = Nx;
  a.x0 = 1;
  a.x1 = 8 * (Nx / 8);
  a.skip0 = 0;
  a.skip1 = Ny / 2;
  a.norm = (float)(1.0 / ((size_t)Nx * Ny));
  a.x = h.anaAxes.p;
  a.y = h.anaAxes.p + Nx;
  a.yw = h.anaAxes.p + Nx + Ny;
  a.include = h.anaInc.p;
  DH_TRY(h.anaPw.ensure((size_t)nB * N), "alloc");
  DH_TRY(h.snapMap.ensure((size_t)P * N), "alloc");
  DH_TRY(h.snapStats.ensure((size_t)nB * 8 + 8), "alloc");
  DH_TRY(dh::launch_plane_stats(a, h.spec0.p + (size_t)p * N, nullptr, nB, (size_t)P * N, h.anaPw.p, h.snapStats.p,
                                h.stream), "stats");
  DH_TRY(dh::launch_batch_sum(h.anaPw.p, nB, 1, N, h.snapMap.p + (size_t)p * N, (size_t)N, h.stream), "batch sum");
  DH_TRY(dh::launch_plane_stats(a, nullptr, h.snapMap.p + (size_t)p * N, 1, (size_t)N, nullptr,
                                h.snapStats.p + (size_t)nB * 8, h.stream), "stats");
  double st[8];
  DH_TRY(hipMemcpyAsync(st, h.snapStats.p + (size_t)nB * 8, sizeof(st), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  const float dA = std::fabs(xs[2] - xs[1]) * std::fabs(ys[2] - ys[1]);
  const float tot = (float)st[0];
  aeff = dA * (tot * tot / (float)st[4]);
  return DIGHOLO_ERROR_SUCCESS;
}

// Tilt (degrees) of pol p from its map in h.snapMap (digHolo.cpp:11612-11708)
int snap_tilt(Handle& h, const dh::Geometry& g, int p, float apRad, float& tx, float& ty) {
  const Config& c = h.cfg;
  const int Nx = g.W / 2 + 1, Ny = g.H, N = Nx * Ny;
  const float k0 = (2 * PI_F) / c.wavelengthCentre;
  const float aperturek = k0 * apRad;
  const float wx = (float)((1.0 * PI_F / aperturek) * std::sqrt(2.0));
  const float wy = (float)((2.0 * PI_F / aperturek) * std::sqrt(2.0));
  const float wx2 = (float)(-1.0 / (wx * wx));
  const float wy2 = (float)(-1.0 / (wy * wy));
  float ad = aperturek * 2;
  ad = ad * ad;
  const float kfx = k_factor(g.W, c.pixelSize), kfy = k_factor(g.H, c.pixelSize);
  std::vector<float> axes(Nx + Ny);
  for (int i = 0; i < Nx; ++i) axes[i] = k_axis(i, g.W, kfx);
  for (int j = 0; j < Ny; ++j) axes[Nx + j] = k_axis(j, g.H, kfy);
  DH_TRY(h.snapAxes.ensure(axes.size()), "alloc");
  DH_TRY(h.snapWork.ensure(dh::snap_work_elems(Ny, Nx)), "alloc");
  DH_TRY(h.snapIdx.ensure(1), "alloc");
  DH_TRY(hipMemcpyAsync(h.snapAxes.p, axes.data(), axes.size() * sizeof(float), hipMemcpyHostToDevice, h.stream),
         "axes upload");
  DH_TRY(dh::launch_snap_tilt_search(h.snapMap.p + (size_t)p * N, Ny, Nx, c.pixelSize, wx2, wy2, h.snapAxes.p,
                                     h.snapAxes.p + Nx, ad, h.snapWork.p, h.snapIdx.p, h.stream), "tilt search");
  int idx = -1;
  DH_TRY(hipMemcpyAsync(&idx, h.snapIdx.p, sizeof(int), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  float kx = 0.f, ky = 0.f;
  if (idx >= 0) {
    kx = axes[idx % Nx];
    ky = axes[Nx + idx / Nx];
  }
  tx = std::asin(kx / k0) / UNIT_ANGLE_F;
  ty = std::asin(ky / k0) / UNIT_ANGLE_F;
  return DIGHOLO_ERROR_SUCCESS;
}

// Effective area of the batch-summed far-field intensity of each pol's fields
// (the unnormalised inverse 2-D DFT of the tilt-removed field,
// digHolo.cpp:11280-11380): dA (sum |F|^2)^2 / sum_pixels (sum_batch |F|^2)^2.
int farfield_aeff(Handle& h, float* aeff) {
  const Derived& d = h.cur;
  const dh::Geometry& g = d.g;
  const int B = g.B, P = g.P, N = g.fw * g.fh;
  dh::Geometry gF = g;
  gF.w = g.fw;
  gF.h = g.fh;
  const dh::FFTPlan* pIx = fft_plan(h, g.fw);
  const dh::FFTPlan* pIy = fft_plan(h, g.fh);
  if (!pIx || !pIy) return DIGHOLO_ERROR_MEMORYALLOCATION;
  const size_t n = (size_t)B * P * N;
  DH_TRY(h.ffBuf.ensure(2 * n), "alloc");
  DH_TRY(h.ffMap.ensure((size_t)P * N + (size_t)std::max(g.fw, g.fh)), "alloc");
  DH_TRY(h.anaPw.ensure(n), "alloc");
  DH_TRY(h.snapStats.ensure((size_t)B * P * 8 + (size_t)P * 8), "alloc");
  float* zeroAxis = h.ffMap.p + (size_t)P * N;
  DH_TRY(hipMemsetAsync(zeroAxis, 0, (size_t)std::max(g.fw, g.fh) * sizeof(float), h.stream), "memset");
  DH_TRY(hipMemcpyAsync(h.ffBuf.p, h.fields.p, n * sizeof(float2), hipMemcpyDeviceToDevice, h.stream), "copy");
  DH_TRY(dh::launch_ifft_rows(gF, *pIx, h.ffBuf.p, h.stream), "far-field rows");
  DH_TRY(dh::launch_ifft_cols_tilt(gF, *pIy, h.ffBuf.p, nullptr, h.ffBuf.p + n, h.stream), "far-field columns");
  dh::PlaneArgs a{};
  a.Nx = g.fw;
  a.Ny = g.fh;
  a.inStride = g.fw;
  a.x0 = 0;
  a.x1 = g.fw;
  a.skip0 = a.skip1 = -1;
  a.norm = 1.f;
  a.x = a.y = a.yw = zeroAxis;
  DH_TRY(dh::launch_plane_stats(a, h.ffBuf.p + n, nullptr, B * P, (size_t)N, h.anaPw.p, h.snapStats.p, h.stream),
         "stats");
  DH_TRY(dh::launch_batch_sum(h.anaPw.p, B, P, N, h.ffMap.p, (size_t)N, h.stream), "batch sum");
  DH_TRY(dh::launch_plane_stats(a, nullptr, h.ffMap.p, P, (size_t)N, nullptr, h.snapStats.p + (size_t)B * P * 8,
                                h.stream), "stats");
  std::vector<double> st((size_t)P * 8);
  DH_TRY(hipMemcpyAsync(st.data(), h.snapStats.p + (size_t)B * P * 8, st.size() * sizeof(double),
                        hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  const float kfx = k_factor(g.W, h.cfg.pixelSize), kfy = k_factor(g.H, h.cfg.pixelSize);
  const float dA = std::fabs(k_axis(2, g.W, kfx) - k_axis(1, g.W, kfx)) *
                   std::fabs(k_axis(2, g.H, kfy) - k_axis(1, g.H, kfy));
  for (int p = 0; p < P; ++p) {
    const float v2 = (float)st[(size_t)p * 8 + 0];
    aeff[p] = dA * ((v2 * v2) / (float)st[(size_t)p * 8 + 4]);
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// Defocus of least far-field effective area (digHoloAutoFocus,
// digHolo.cpp:11181-11515).  Trials +s, -s, then along the better sign in
// quarter steps of s (s = the estimate, at least 1 dioptre), and from the fifth
// trial on the vertex of the least-squares quadratic through the four trials
// of smallest area, until the vertex is within 0.01 of the best trial (25 at
// most).  The fields are those of the beam centres cx/cy, which the
// reference's last inverse transform used.
int autofocus(Handle& h, float* scale, float* best, const float* cx, const float* cy) {
  Config& c = h.cfg;
  const int P = c.polCount;
  constexpr int stepCount = 25, stepFitCount = 5;
  const float tol = 0.01f;
  float aeffs[2][stepCount + 3] = {}, defs[2][stepCount + 3] = {};
  float bestAeff[2] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
  bool converged[2] = {false, false};
  for (int p = 0; p < P; ++p) {
    best[p] = 0.f;
    if (scale[p] < 1.0f) scale[p] = 1.0f;
  }
  float s0[2] = {scale[0], P > 1 ? scale[1] : 0.f};
  for (int st = 0; st < stepCount; st += 2)
    for (int p = 0; p < P; ++p) {
      defs[p][st] = s0[p];
      defs[p][st + 1] = -s0[p];
      s0[p] += scale[p] * 0.25f;
    }
  float keepCX[2], keepCY[2];
  for (int p = 0; p < 2; ++p) {
    keepCX[p] = c.beamCX[p];
    keepCY[p] = c.beamCY[p];
    c.beamCX[p] = cx[p];
    c.beamCY[p] = cy[p];
  }
  // (with the defocus pol-locked, pol 1's fields take pol 0's trial value, as
  // the reference's reference wave does)
  bool done = false;
  int e = DIGHOLO_ERROR_SUCCESS;
  for (int st = 0; st < stepCount && !done; ++st) {
    done = true;
    for (int p = 0; p < P; ++p) c.defocus[p] = defs[p][st];
    float aeff[2] = {0.f, 0.f};
    if ((e = enqueue_batch(h)) || (e = farfield_aeff(h, aeff))) break;
    for (int p = 0; p < P; ++p) {
      aeffs[p][st] = aeff[p];
      if (aeff[p] < bestAeff[p]) {
        bestAeff[p] = aeff[p];
        best[p] = defs[p][st];
      }
      if (st == 1) {   // continue on the side of the better of +s and -s
        float stepScale, base;
        if (aeffs[p][1] < aeffs[p][0]) {
          stepScale = defs[p][1] - defs[p][0];
          base = defs[p][1];
        } else {
          stepScale = defs[p][0] - defs[p][1];
          base = defs[p][0];
        }
        int k = 1;
        for (int i = st + 1; i < stepCount; i += 2) {
          defs[p][i] = base + k * 0.25f * stepScale;
          defs[p][i + 1] = base - k * 0.25f * stepScale;
          ++k;
        }
      }
      if (st >= stepFitCount - 1) {
        // the stepFitCount - 1 trials of smallest area (stable order, as the
        // reference's bubble sort)
        std::vector<int> order(st + 1);
        for (int i = 0; i <= st; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return aeffs[p][a] < aeffs[p][b]; });
        double xs[stepFitCount - 1], ys[stepFitCount - 1];
        for (int i = 0; i < stepFitCount - 1; ++i) {
          xs[i] = defs[p][order[i]];
          ys[i] = aeffs[p][order[i]];
        }
        const float vx = (float)quadratic_vertex(xs, ys, stepFitCount - 1);
        if (st + 1 < stepCount - 1 && !std::isnan(vx)) defs[p][st + 1] = vx;
        if (std::fabs(vx - best[p]) < tol) converged[p] = true;
      }
      done = done && converged[p];
    }
  }
  for (int p = 0; p < 2; ++p) {
    c.beamCX[p] = keepCX[p];
    c.beamCY[p] = keepCY[p];
  }
  return e;
}

// The estimate stage.  doFinal: process once more and store the metrics
// (ESTIMATE mode, or no basis).  steps: base step sizes for the tweak stage.
int autoalign_snap(Handle& h, bool doFinal, float* steps) {
  Config& c = h.cfg;
  const int P = c.polCount;
  const float px = c.pixelSize;
  for (int p = 0; p < P; ++p) {
    if (c.aaCentre) {
      c.beamCX[p] = (float)((-(P - 1) * (c.frameWidth / 4.0) + (p * c.frameWidth / 2.0)) * px);
      c.beamCY[p] = 0.f;
    }
    if (c.aaTilt) c.tiltX[p] = c.tiltY[p] = 0.f;
    if (c.aaDefocus) c.defocus[p] = 0.f;
  }
  float waists[2][2] = {};
  c.fftWindowSizeX = c.frameWidth / P;
  c.fftWindowSizeY = c.frameHeight;
  if (c.batchCount < 1) c.batchCount = 1;

  // the R2C Fourier plane of every frame under the whole-half-frame window
  int W, H, cxw[2] = {0, 0}, cyw[2] = {0, 0}, err;
  if (!fft_window(c, W, H, cxw, cyw, err)) return err;
  if (int e = ensure_device(h)) return e;
  const int nB = c.batchCount * std::max(1, c.avgCount);
  const void* dframes = nullptr;
  if (int e = device_frames(h, (size_t)nB, dframes)) return e;
  dh::Geometry g{};
  g.B = nB;
  g.P = P;
  g.frameW = c.frameWidth;
  g.frameH = c.frameHeight;
  g.W = W;
  g.H = H;
  g.w = W / 2 + 1;
  g.h = H;
  g.fw = g.w;
  g.fh = g.h;
  for (int p = 0; p < P; ++p) {
    g.cx[p] = cxw[p];
    g.cy[p] = cyw[p];
  }
  const dh::FFTPlan* pRow = fft_plan(h, W / 2);
  const dh::FFTPlan* pCol = fft_plan(h, H);
  if (!pRow || !pCol) return DIGHOLO_ERROR_MEMORYALLOCATION;
  if (int e = ensure_row_twiddles(h, W)) return e;
  const int Nx = W / 2 + 1;
  DH_TRY(h.spec0.ensure((size_t)nB * P * H * Nx), "alloc");
  DH_TRY(dh::launch_row_fft(g, *pRow, h.twW.p, dframes, h.frameU16, h.spec0.p, h.stream), "row FFT");
  DH_TRY(dh::launch_col_fft_full(g, *pCol, h.spec0.p, h.stream), "column FFT");
  h.analysisValid[0] = false;   // spec0 no longer holds the last batch's plane

  const float k0 = (2 * PI_F) / c.wavelengthCentre;
  float apRad = c.apertureSize * UNIT_ANGLE_F;
  auto zero_order_maps = [&]() -> int {   // digHoloFFTAnalysis (digHolo.cpp:6931-6984)
    const float R = 2 * k0 * std::sin(apRad);
    float unused;
    for (int p = 0; p < P; ++p)
      if (int e = fourier_pol_map(h, g, nB, p, 0.f, 0.f, R, unused)) return e;
    return 0;
  };
  if (int e = zero_order_maps()) return e;
  const int passCount = 1 + (c.aaFourierRadius ? 1 : 0);
  for (int pass = 0; pass < passCount; ++pass) {
    for (int p = 0; p < P; ++p) {
      float tx, ty;
      if (int e = snap_tilt(h, g, p, apRad, tx, ty)) return e;
      if (c.aaFourierRadius) {
        const float apMax = std::sqrt(tx * tx + ty * ty) / 3.0f;
        if (apMax < c.apertureSize) {
          c.apertureSize = apMax;
          apRad = apMax * UNIT_ANGLE_F;
          if (int e = zero_order_maps()) return e;
        }
      }
      if (c.aaTilt) {
        c.tiltX[p] = tx;
        c.tiltY[p] = ty;
      }
      const float wcx = k0 * (c.tiltX[p] * UNIT_ANGLE_F);
      const float wcy = k0 * (c.tiltY[p] * UNIT_ANGLE_F);
      const float wr = -k0 * std::sin(apRad);
      float aeff;
      if (int e = fourier_pol_map(h, g, nB, p, wcx, wcy, wr, aeff)) return e;
      const float wEst = aeff_to_mfd(aeff, c.maxMG);
      waists[0][p] = (float)((1.0 / wEst) * PI_F);
      if (c.verbosity >= DIGHOLO_VERBOSITY_DEBUG) {
        fprintf(stdout, "Tilt\t%d\t%f,%f\n", p, tx, ty);
        fflush(stdout);
      }
    }
  }

  // reconstructed plane: centre of mass and effective area (digHolo.cpp:11800-11905)
  constexpr int maxAttempts = 10;
  bool valid = true;
  int attempts = 0;
  float maxFocus = 0.25f;
  float defEst[2] = {0.f, 0.f};
  float usedCX[2] = {c.beamCX[0], c.beamCX[1]}, usedCY[2] = {c.beamCY[0], c.beamCY[1]};
  while ((!valid || attempts == 0) && attempts < maxAttempts) {
    valid = true;
    for (int p = 0; p < 2; ++p) {
      usedCX[p] = c.beamCX[p];
      usedCY[p] = c.beamCY[p];
    }
    if (int e = enqueue_batch(h)) return e;
    if (int e = compute_analysis(h, 1)) return e;
    const int B = h.cur.g.B;
    const std::vector<float>& an = h.analysis[1];
    auto at = [&](int par, int p) { return an[((size_t)par * P + p) * (B + 1) + B]; };
    for (int p = 0; p < P; ++p) {
      const float polOffsetX = (float)(((2 * p - 1) * c.pixelSize * c.frameWidth / 4.0) * (P - 1));
      const float comx = at(1, p) / 1 + polOffsetX;
      const float comy = at(2, p) / 1;
      const float Aeff = at(5, p);
      const float wEst = (float)(aeff_to_mfd(Aeff, c.maxMG) * 0.5);
      waists[1][p] = wEst;
      defEst[p] = focus_estimate(waists[1][p], waists[0][p]);
      if (c.verbosity >= DIGHOLO_VERBOSITY_DEBUG) {
        fprintf(stdout, "%d\tBeamCentre\t(%e\t%e)\tMFD:%e\tDefocus:%f\tAeff:%e\n", p, comx, comy, wEst, defEst[p],
                Aeff);
        fflush(stdout);
      }
      if (std::isnan(comx) || std::isnan(comy) || std::isnan(Aeff)) {
        valid = false;   // probably the zero order in the window: shrink it
        c.apertureSize *= 0.9f;
      }
      if (valid) {
        if (c.aaCentre) {
          c.beamCX[p] = comx;
          c.beamCY[p] = comy;
        }
        if (c.aaWaist) c.waist[p] = wEst / 1;
        if (c.aaDefocus) c.defocus[p] = 0.f;
        if (std::fabs(defEst[p]) > maxFocus) maxFocus = std::fabs(defEst[p]);
      }
      ++attempts;
    }
  }

  if (c.aaDefocus) {
    float best[2] = {0.f, 0.f};
    if (int e = autofocus(h, defEst, best, usedCX, usedCY)) return e;
    for (int p = 0; p < P; ++p) {
      c.defocus[p] = best[p];
      if (std::fabs(best[p]) > maxFocus) maxFocus = std::fabs(best[p]);
    }
  }

  if (doFinal) {
    if (int e = enqueue_batch(h)) return e;
    if (h.cur.g.M > 0) {
      PowerMeterOut o;
      if (int e = power_meter(h, -1, c.aaPolIndependence != 0, true, true, true, c.aaMulConjTrans != 0, o)) return e;
      store_metrics_db(h, o);
    }
  }
  steps[AA_TILTX] = steps[AA_TILTY] = c.apertureSize * 0.1f;
  steps[AA_CX] = steps[AA_CY] = c.waist[0] * 0.1f;
  steps[AA_DEFOCUS] = 0.1f * maxFocus;
  steps[AA_WAIST] = c.waist[0] * 0.1f;
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Alignment-loss gradient (no counterpart in the reference; dh_grad.hip holds
// the adjoint).  L = -(1/B) sum |c|^2 (minus the linear insertion loss without
// the SVD); gradient with respect to [pol][AA_WAIST..AA_CY] in API units (m,
// dioptre, degree, m, degree, m).  The window placement is a step function of
// centre and tilt: the gradient is that of the smooth piece the current
// placement belongs to; hold keeps the previous call's placement.
// ---------------------------------------------------------------------------
// Mode-dependent-loss term mu (ln s_max^2 - ln s_min^2) of the coefficient
// matrix the power meter takes the MDL of (digHolo.cpp:10796-11089: frames as
// rows, both pols side by side unless independent), and its seed
// dterm/dc (Wirtinger) into h.gExtra: with (u, s, v) a singular triplet,
// d ln s^2 = 2 Re(u^H dA v) / s, i.e. conj(u_b) v_j / s.  The extreme
// eigenpairs of the smaller FP64 Gram matrix come from the GPU Gram and the
// host eigen-solver; the other singular vector is A v / s or A^H u / s.
int mdl_seed(Handle& h, double mu, double& term) {
  const Config& c = h.cfg;
  const dh::Geometry& g = h.cur.g;
  const int B = g.B, P = g.P, Mpol = g.M;
  const bool indep = c.aaPolIndependence != 0 || P == 1;
  const int S = (P == 2 && indep) ? 2 : 1;
  const int Mm = (P == 2 && !indep) ? 2 * Mpol : Mpol;
  const int rowStride = P * Mpol;
  term = 0.0;
  const size_t total = (size_t)B * P * Mpol;
  DH_TRY(h.gExtra.ensure(total), "alloc");
  DH_TRY(dh::launch_zero_complex(h.gExtra.p, total, h.stream), "clear");
  const bool rowsVec = B <= Mm;
  const int dim = rowsVec ? B : Mm;
  struct Pick {
    int s = -1;
    double l = 0;
    std::vector<std::complex<double>> x;
  } lo, hi;
  // the Gram of each matrix on the device, its extreme eigenpairs on the host
  DH_TRY(h.gram.ensure((size_t)dim * dim), "alloc");
  std::vector<std::complex<double>> G((size_t)dim * dim);
  for (int s = 0; s < S; ++s) {
    DH_TRY(dh::launch_gram_f64(h.coefs.p, B, Mm, rowStride, s, rowsVec ? 1 : 0, h.gram.p, h.stream), "gram");
    DH_TRY(hipMemcpyAsync(G.data(), h.gram.p, G.size() * sizeof(double2), hipMemcpyDeviceToHost, h.stream), "D2H");
    DH_TRY(hipStreamSynchronize(h.stream), "sync");
    dh::HermitianExtremePairs r = dh::hermitian_extreme_pairs(dim, G);
    if (lo.s < 0 || r.lmin < lo.l) lo = Pick{s, r.lmin, r.vmin};
    if (hi.s < 0 || r.lmax > hi.l) hi = Pick{s, r.lmax, r.vmax};
  }
  if (dim >= 2 && mu != 0.0) {
    if (!(lo.l > 0) || !(hi.l > 0)) return DIGHOLO_ERROR_INVALIDARGUMENT;   // MDL undefined
    term = mu * (std::log(hi.l) - std::log(lo.l));
    // the projections and the rank-one seeds run on the device over the
    // resident coefficients; only the eigenvectors travel
    DH_TRY(h.mdlX.ensure(2 * (size_t)dim), "alloc");
    DH_TRY(h.mdlY.ensure(2 * (size_t)(rowsVec ? Mm : B)), "alloc");
    std::vector<std::complex<double>> xs(2 * (size_t)dim);
    std::copy(hi.x.begin(), hi.x.end(), xs.begin());
    std::copy(lo.x.begin(), lo.x.end(), xs.begin() + dim);
    DH_TRY(hipMemcpyAsync(h.mdlX.p, xs.data(), xs.size() * sizeof(double2), hipMemcpyHostToDevice, h.stream),
           "upload");
    const size_t ny = rowsVec ? Mm : B;
    int k = 0;
    for (const Pick* pk : {&hi, &lo}) {
      const double sig = std::sqrt(pk->l);
      const double coef = (pk == &hi ? mu : -mu) / sig;
      DH_TRY(dh::launch_mdl_seed_pair(h.coefs.p, B, Mm, rowStride, pk->s, rowsVec ? 1 : 0, h.mdlX.p + k * dim, sig,
                                      coef, h.mdlY.p + k * ny, h.gExtra.p, h.stream), "MDL seed");
      ++k;
    }
    DH_TRY(hipStreamSynchronize(h.stream), "sync");   // xs leaves scope
  }
  return DIGHOLO_ERROR_SUCCESS;
}

int alignment_gradient(Handle& h, bool hold, double mdlWeight, double& loss, double grad[2][AA_PARS]) {
  Config& c = h.cfg;
  if (c.maxMG <= 0) return DIGHOLO_ERROR_INVALIDARGUMENT;
  if (c.wavelengths.size() > 1) {   // the adjoint is written for one reference wave per pol
    fprintf(stderr, "digHolo: the alignment gradient covers single-wavelength batches\n");
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  const bool useHold = hold && h.gradHeldValid;
  const bool hc = h.holdCrop, hw = h.holdWin;
  if (useHold) {
    h.heldCrop = h.gradHeld;
    h.heldWin = h.gradHeld;
    h.holdCrop = h.holdWin = true;
  }
  int e = enqueue_batch(h);
  h.holdCrop = hc;
  h.holdWin = hw;
  if (e) return e;
  h.gradHeld = h.cur;
  h.gradHeldValid = true;
  const Derived& d = h.cur;
  const dh::Geometry& g = d.g;
  const int B = g.B, P = g.P, fw = g.fw, fh = g.fh, N = fw * fh;
  dh::Geometry gF = g;
  gF.w = fw;
  gF.h = fh;
  const dh::FFTPlan* pIx = fft_plan(h, fw);
  const dh::FFTPlan* pIy = fft_plan(h, fh);
  if (!pIx || !pIy) return DIGHOLO_ERROR_MEMORYALLOCATION;
  const size_t n = (size_t)B * P * N;
  DH_TRY(h.dhgx.ensure((size_t)P * g.G * fw), "alloc");
  DH_TRY(h.dhgy.ensure((size_t)P * g.G * fh), "alloc");
  DH_TRY(h.gq1.ensure(n), "alloc");
  DH_TRY(h.gq2.ensure(n), "alloc");
  DH_TRY(h.gx1.ensure(n), "alloc");
  DH_TRY(h.gx2.ensure(n), "alloc");
  DH_TRY(h.gPart.ensure((size_t)B * P * 8), "alloc");
  DH_TRY(h.gSum.ensure((size_t)P * 8), "alloc");
  // per pol: {pxx, pxy}, then the absolute kx of the field plane's columns and
  // ky of its rows (window order in low-res mode, FFT order in full resolution)
  const bool lowRes = fw == g.w && fh == g.h && c.resolutionIdx == 1;
  const double px = (double)c.pixelSize;
  const double dkx = 2.0 * PI_D / (px * g.W), dky = 2.0 * PI_D / (px * g.H);
  std::vector<double> ax(2 * P + (size_t)P * (fw + fh));
  double* kxT = ax.data() + 2 * P;
  double* kyT = kxT + (size_t)P * fw;
  for (int p = 0; p < P; ++p) {
    ax[2 * p] = d.tab.pp[p].pxx;
    ax[2 * p + 1] = d.tab.pp[p].pxy;
    for (int i = 0; i < fw; ++i) {
      const int col = lowRes ? g.c0[p] + i : (i <= g.W / 2 ? i : i - g.W);
      kxT[(size_t)p * fw + i] = col * dkx;
    }
    for (int j = 0; j < fh; ++j) {
      int row = lowRes ? imod(j + g.SY[p] + g.H / 2, g.H) : j;
      if (row > g.H / 2) row -= g.H;
      kyT[(size_t)p * fh + j] = row * dky;
    }
  }
  DH_TRY(h.gAxes.ensure(ax.size()), "alloc");
  DH_TRY(hipMemcpyAsync(h.gAxes.p, ax.data(), ax.size() * sizeof(double), hipMemcpyHostToDevice, h.stream),
         "axes upload");
  double mdlTerm = 0.0;
  const float2* gExtra = nullptr;
  if (mdlWeight != 0.0) {
    if (int em = mdl_seed(h, mdlWeight, mdlTerm)) return em;
    gExtra = h.gExtra.p;
  }
  DH_TRY(dh::launch_hg_dwaist(d.tab, h.dhgx.p, h.dhgy.p, h.stream), "waist derivative basis");
  DH_TRY(dh::launch_grad_backproject(gF, 1.0 / B, h.coefs.p, gExtra, h.hgx.p, h.hgy.p, h.dhgx.p, h.dhgy.p, h.fields.p,
                                     h.ref.p, h.gAxes.p, h.gq1.p, h.gq2.p, h.gPart.p, h.stream), "back-projection");
  DH_TRY(dh::launch_ifft_rows(gF, *pIx, h.gq1.p, h.stream), "adjoint IFFT rows");
  DH_TRY(dh::launch_ifft_cols_tilt(gF, *pIy, h.gq1.p, nullptr, h.gx1.p, h.stream), "adjoint IFFT columns");
  DH_TRY(dh::launch_ifft_rows(gF, *pIx, h.gq2.p, h.stream), "adjoint IFFT rows");
  DH_TRY(dh::launch_ifft_cols_tilt(gF, *pIy, h.gq2.p, nullptr, h.gx2.p, h.stream), "adjoint IFFT columns");
  DH_TRY(dh::launch_grad_kspace(gF, h.gx1.p, h.gx2.p, kxT - ax.data() + h.gAxes.p, kyT - ax.data() + h.gAxes.p,
                                h.gPart.p, h.stream), "k-space sums");
  DH_TRY(dh::launch_grad_reduce(B, P, h.gPart.p, h.gSum.p, h.stream), "batch reduction");
  double S[2][8] = {};
  DH_TRY(hipMemcpyAsync(S, h.gSum.p, (size_t)P * 8 * sizeof(double), hipMemcpyDeviceToHost, h.stream), "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");

  const double lam = (double)c.wavelengthCentre, k0 = 2.0 * PI_D / lam, deg = PI_D / 180.0;
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < AA_PARS; ++k) grad[p][k] = 0.0;
  loss = mdlTerm;
  for (int p = 0; p < P; ++p) {
    const int tp = (c.polLockTilt || c.polLockDefocus) ? 0 : p;
    const int ip = c.polLockDefocus ? 0 : p;
    const int wp = c.polLockWaist ? 0 : p;
    const double S0 = S[p][0], Sx = S[p][1], Sy = S[p][2], Sr2 = S[p][3], Sw = S[p][4];
    const double Tbx = S[p][6] / N, Tby = S[p][7] / N;
    const double thx = (double)c.tiltX[tp] * deg, thy = (double)c.tiltY[tp] * deg;
    const double bcx = (double)c.beamCX[p], bcy = (double)c.beamCY[p];
    grad[tp][AA_TILTX] += -2.0 * k0 * (Sx + bcx * S0) * deg;
    grad[tp][AA_TILTY] += -2.0 * k0 * (Sy + bcy * S0) * deg;
    grad[ip][AA_DEFOCUS] += -2.0 * (PI_D / lam) * Sr2;
    grad[p][AA_CX] += -2.0 * Tbx - 2.0 * k0 * thx * S0;
    grad[p][AA_CY] += -2.0 * Tby - 2.0 * k0 * thy * S0;
    grad[wp][AA_WAIST] += 2.0 * Sw;
    loss -= S[p][5] / B;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

// Gradient-based alignment (no counterpart in the reference): BFGS on the
// enabled parameters (the AutoAlign enable flags and pol locks), each scaled by
// the step the tweak search would start with, Armijo backtracking, placement
// re-derived at every trial point.  Ends at the best point with the metrics of
// a final evaluation stored; returns the goal metric in dB.
int autoalign_gradient_search(Handle& h, int maxIter, double mdlWeight, float& goalDb) {
  Config& c = h.cfg;
  const int P = c.polCount;
  HoldGuard guard(h);
  if (c.aaFourierRadius) c.apertureSize = window_radius_rec(c);
  struct Var {
    int p, par;
    double scale;
  };
  std::vector<Var> vars;
  const float w0 = c.waist[0];
  for (int par = 0; par < AA_PARS; ++par) {
    const bool on = (par == AA_TILTX || par == AA_TILTY) ? c.aaTilt != 0
                    : (par == AA_CX || par == AA_CY)     ? c.aaCentre != 0
                    : par == AA_WAIST                    ? c.aaWaist != 0
                                                         : c.aaDefocus != 0;
    if (!on) continue;
    const bool lock = ((par == AA_TILTX || par == AA_TILTY) && c.polLockTilt) ||
                      (par == AA_DEFOCUS && c.polLockDefocus) || (par == AA_WAIST && c.polLockWaist);
    double sc = (par == AA_TILTX || par == AA_TILTY) ? 0.1 * c.apertureSize
                : par == AA_DEFOCUS ? std::max(0.4, 0.1 * std::fabs((double)c.defocus[0]))
                                    : 0.1 * w0;
    for (int p = 0; p < (lock ? 1 : P); ++p) vars.push_back({p, par, sc});
  }
  const int nv = (int)vars.size();
  std::vector<double> x0(nv);
  for (int i = 0; i < nv; ++i) x0[i] = aa_param(c, vars[i].p, vars[i].par);
  double L0 = 0.0, G0[2][AA_PARS];
  auto apply = [&](const std::vector<double>& z) {
    for (int i = 0; i < nv; ++i) aa_param(c, vars[i].p, vars[i].par) = (float)(x0[i] + z[i] * vars[i].scale);
  };
  double norm = 1.0;
  auto eval = [&](const std::vector<double>& z, double& f, std::vector<double>& gz) -> int {
    apply(z);
    double L, Gr[2][AA_PARS];
    if (int e = alignment_gradient(h, false, mdlWeight, L, Gr)) return e;
    f = L / norm;
    gz.resize(nv);
    for (int i = 0; i < nv; ++i) gz[i] = Gr[vars[i].p][vars[i].par] * vars[i].scale / norm;
    return 0;
  };
  if (int e = alignment_gradient(h, false, mdlWeight, L0, G0)) return e;
  norm = std::max(std::fabs(L0), 1e-30);
  std::vector<double> z(nv, 0.0), gz, zBest = z;
  double f = 0, fBest;
  if (nv > 0) {
    if (int e = eval(z, f, gz)) return e;
  }
  fBest = f;
  std::vector<double> Hi((size_t)nv * nv, 0.0);
  for (int i = 0; i < nv; ++i) Hi[(size_t)i * nv + i] = 1.0;
  int quiet = 0;
  for (int it = 0; it < maxIter && nv > 0; ++it) {
    std::vector<double> dir(nv, 0.0);
    double gd = 0.0, dn = 0.0;
    for (int i = 0; i < nv; ++i) {
      for (int j = 0; j < nv; ++j) dir[i] -= Hi[(size_t)i * nv + j] * gz[j];
      gd += gz[i] * dir[i];
      dn += dir[i] * dir[i];
    }
    if (gd >= 0) {   // not a descent direction: restart from steepest descent
      std::fill(Hi.begin(), Hi.end(), 0.0);
      for (int i = 0; i < nv; ++i) Hi[(size_t)i * nv + i] = 1.0;
      gd = dn = 0.0;
      for (int i = 0; i < nv; ++i) {
        dir[i] = -gz[i];
        gd += gz[i] * dir[i];
        dn += dir[i] * dir[i];
      }
    }
    dn = std::sqrt(dn);
    double t = dn > 2.0 ? 2.0 / dn : 1.0;   // at most two start steps per iteration
    std::vector<double> zn(nv), gn;
    double fn = 0;
    bool ok = false;
    for (int ls = 0; ls < 12; ++ls) {
      for (int i = 0; i < nv; ++i) zn[i] = z[i] + t * dir[i];
      if (int e = eval(zn, fn, gn)) return e;
      if (fn <= f + 1e-4 * t * gd) {
        ok = true;
        break;
      }
      t *= 0.5;
    }
    if (!ok) break;
    std::vector<double> sv(nv), yv(nv);
    double sy = 0.0;
    for (int i = 0; i < nv; ++i) {
      sv[i] = zn[i] - z[i];
      yv[i] = gn[i] - gz[i];
      sy += sv[i] * yv[i];
    }
    if (sy > 1e-300) {   // BFGS update of the inverse Hessian
      std::vector<double> Hy(nv, 0.0);
      double yHy = 0.0;
      for (int i = 0; i < nv; ++i) {
        for (int j = 0; j < nv; ++j) Hy[i] += Hi[(size_t)i * nv + j] * yv[j];
        yHy += yv[i] * Hy[i];
      }
      const double r = 1.0 / sy;
      for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j)
          Hi[(size_t)i * nv + j] += (1.0 + yHy * r) * r * sv[i] * sv[j] - r * (Hy[i] * sv[j] + sv[i] * Hy[j]);
    }
    const double df = f - fn;
    z = zn;
    gz = gn;
    f = fn;
    if (f < fBest) {
      fBest = f;
      zBest = z;
    }
    if (c.verbosity >= DIGHOLO_VERBOSITY_DEBUG) {
      fprintf(stdout, "Gradient align iteration %d  loss %.9f  step %.3g\n", it, f, t * dn);
      fflush(stdout);
    }
    quiet = df < 1e-9 ? quiet + 1 : 0;
    if (quiet >= 2) break;
  }
  apply(zBest);
  // the window radius follows the final tilt, as the tweak search leaves it
  // (a third of the smallest tilt, digHolo.cpp:12560-12575)
  if (c.aaFourierRadius) c.apertureSize = window_radius_rec(c);
  if (int e = enqueue_batch(h)) return e;
  PowerMeterOut o;
  if (int e = power_meter(h, -1, c.aaPolIndependence != 0, true, true, true, c.aaMulConjTrans != 0, o)) return e;
  store_metrics_db(h, o);
  goalDb = h.aaMetrics[(size_t)c.aaGoalIdx * (h.aaLambdaCount + 1) + h.aaLambdaCount];
  return DIGHOLO_ERROR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Several handles' batches in one launch sequence (no counterpart in the
// reference).  The first call for a handle list captures it into a graph:
// each handle's batch is prepared on its own stream, a branch forked from and
// joined to the first handle's stream; batches on the size-specialised window
// kernel (HG basis) only record that launch, and the recorded launches are
// issued on the first handle's stream after the join -- as one launch of the
// multi-batch window kernel while the batches are small, else one after the
// other -- followed by the separate overlaps of those batches.  When the graph
// holds nothing but these launches, later calls issue them directly (a graph
// launch costs more host time than a few kernel launches); otherwise they
// replay the graph.  Either is kept per handle list while every handle's
// configuration and launch state are unchanged.
// ---------------------------------------------------------------------------

// Every Config field, compared bitwise; the batch-calibration, custom-basis
// and reference-calibration arrays by their version counters.  The wavelength
// list is compared by value.
#define DH_CONFIG_SCALARS(X)                                                                                      \
  X(frameWidth) X(frameHeight) X(pixelSize) X(apertureSize) X(resolutionIdx) X(waist) X(tiltX) X(tiltY) X(defocus) \
  X(beamCX) X(beamCY) X(batchCount) X(avgCount) X(avgMode) X(wavelengthCentre) X(wavelengthOrdering)              \
  X(polLockTilt) X(polLockWaist) X(polLockDefocus) X(polCount) X(aaTilt) X(aaDefocus) X(aaCentre) X(aaWaist)      \
  X(aaFourierRadius) X(aaPolIndependence) X(aaMulConjTrans) X(aaTol) X(aaMode) X(aaGoalIdx) X(fillFactor)         \
  X(fftWindowSizeX) X(fftWindowSizeY) X(maxMG) X(basisType) X(verbosity) X(threadCount) X(precision) X(calP)       \
  X(calB) X(calEnabled) X(calVersion) X(basisIn) X(basisOut) X(basisVersion) X(refCalW) X(refCalH) X(refCalL)     \
  X(refCalEnabled) X(refCalVersion)
#define DH_FIELD_BYTES(f) +sizeof(Config::f)
// a field added to Config has to join DH_CONFIG_SCALARS (or be compared by a
// version counter like the three arrays)
constexpr size_t kConfigPad = 12;   // alignment padding ahead of the vector members (x86-64)
static_assert(sizeof(Config) == (0 DH_CONFIG_SCALARS(DH_FIELD_BYTES)) + 4 * sizeof(std::vector<float>) + kConfigPad,
              "Config changed: update DH_CONFIG_SCALARS");

bool same_config(const Config& a, const Config& b) {
#define DH_FIELD_EQ(f) if (std::memcmp(&a.f, &b.f, sizeof(a.f)) != 0) return false;
  DH_CONFIG_SCALARS(DH_FIELD_EQ)
#undef DH_FIELD_EQ
  return a.wavelengths == b.wavelengths;
}

// what a captured launch sequence depends on besides the configuration
struct LaunchState {
  const void* frames;
  int frameU16, transpose, windowMode, fusedOverlap, ovTpw, autoTune, rowMode;
  bool holdCrop, holdWin;
  unsigned tabVersion;
  hipStream_t stream;
  bool operator==(const LaunchState& o) const {
    return frames == o.frames && frameU16 == o.frameU16 && transpose == o.transpose && windowMode == o.windowMode &&
           fusedOverlap == o.fusedOverlap && ovTpw == o.ovTpw && autoTune == o.autoTune && rowMode == o.rowMode &&
           holdCrop == o.holdCrop && holdWin == o.holdWin && tabVersion == o.tabVersion && stream == o.stream;
  }
};
LaunchState launch_state(const Handle& h) {
  return LaunchState{h.frames,   h.frameU16, h.transpose, h.windowMode, h.fusedOverlap, h.ovTpw,
                     h.autoTune, h.rowMode,  h.holdCrop,  h.holdWin,    h.tabVersion,   h.stream};
}

struct MultiOp {
  int kind = 0;   // 0: multi-batch window kernel over w3[i0, i0 + n); 1: window kernel w3[i0]; 2: overlap
  int i0 = 0, n = 0;
  dh::Geometry g{};   // overlap operands
  const float2* fields = nullptr;
  const float *hgx = nullptr, *hgy = nullptr;
  const void *hgx16 = nullptr, *hgy16 = nullptr;
  float2* coefs = nullptr;
  int split = 0, tpw = 0;
};

struct MultiGraph {
  std::vector<Handle*> hs;
  std::vector<Config> cfgs;   // without the three versioned arrays
  std::vector<LaunchState> states;
  std::vector<dh::W3Launch> w3;
  std::vector<MultiOp> ops;
  bool direct = false;        // the graph is exactly `ops`: issue them without it
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;  // recorded after each launch; the other handles' streams wait on it
};
std::map<std::vector<int>, MultiGraph> g_multi;

// Called with g_multiMu held: the other handles' streams take their pending
// wait on m.done before the event goes.
void multi_release(MultiGraph& m) {
  for (Handle* h : m.hs)
    if (m.done && h->afterMulti == m.done) {
      (void)hipStreamWaitEvent(h->stream, m.done, 0);
      h->afterMulti = nullptr;
    }
  if (m.exec) (void)hipGraphExecDestroy(m.exec);
  if (m.graph) (void)hipGraphDestroy(m.graph);
  if (m.done) (void)hipEventDestroy(m.done);
  m.exec = nullptr;
  m.graph = nullptr;
  m.done = nullptr;
  m.direct = false;
  m.w3.clear();
  m.ops.clear();
}

Config config_sans_arrays(const Config& c) {
  Config k;
#define DH_FIELD_CP(f) std::memcpy(&k.f, &c.f, sizeof(c.f));
  DH_CONFIG_SCALARS(DH_FIELD_CP)
#undef DH_FIELD_CP
  k.wavelengths = c.wavelengths;
  return k;
}

hipError_t run_multi_ops(const MultiGraph& m, hipStream_t s) {
  for (const MultiOp& o : m.ops) {
    hipError_t e = hipSuccess;
    if (o.kind == 0) {
      e = dh::launch_window_v3_multi(m.w3.data() + o.i0, o.n, s);
    } else if (o.kind == 1) {
      const dh::W3Launch& w = m.w3[o.i0];
      e = dh::launch_window_v3(w.fa, w.frames, w.tab, w.factor, w.ref, w.fields, w.hgx16, w.hgy16, w.coefs, w.cal, s,
                               w.hgx32, w.hgy32, w.rowMode);
    } else {
      e = dh::launch_overlap_mfma(o.g, o.fields, o.hgx, o.hgy, o.coefs, o.split, s, o.tpw, nullptr, o.hgx16,
                                  o.hgy16);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Batches whose window kernels together hold at most this many (pol, frame)
// work items share one launch of the multi-batch window kernel: below a few
// waves of workgroups over the chip, separate launches leave it idle in their
// tails.  Larger batches keep their own kernels, which run at the occupancy of
// their own shape rather than that of the largest one in the launch.
#ifndef DH_MULTI_MERGE_ITEMS
#define DH_MULTI_MERGE_ITEMS 1024
#endif
constexpr long kMultiMergeItems = DH_MULTI_MERGE_ITEMS;

// Called with g_multiMu held.
int capture_multi(MultiGraph& m, const std::vector<Handle*>& hs) {
  Handle& h0 = *hs[0];
  // small batches take the specialised window kernel wherever it applies (a
  // tuned choice made for the batch on its own does not hold inside one launch)
  long cfgItems = 0;
  for (Handle* h : hs) cfgItems += (long)std::max(1, h->cfg.polCount) * std::max(1, h->cfg.batchCount);
  struct Prefer {
    const std::vector<Handle*>& hs;
    ~Prefer() {
      for (Handle* h : hs) h->preferV3 = false;
    }
  } prefer{hs};
  for (Handle* h : hs) h->preferV3 = cfgItems <= kMultiMergeItems;
  // one eager batch per handle: tables, twiddles, buffers and tuning settle
  // outside the capture, then the launch sequence is captured
  for (Handle* h : hs)
    if (int e = enqueue_batch(*h)) return e;
  for (Handle* h : hs) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      log_err(h, "synchronize", e);
      return DIGHOLO_ERROR_ERROR;
    }
  }
  hipStream_t s0 = h0.stream;
  std::vector<hipEvent_t> evs(hs.size(), nullptr);
  for (auto& e : evs)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
  auto cleanup = [&]() {
    for (auto& e : evs)
      if (e) (void)hipEventDestroy(e);
    for (Handle* h : hs) h->capturing = h->collect = h->pendValid = h->pendOverlap = false;
  };
  for (auto e : evs)
    if (!e) {
      cleanup();
      return DIGHOLO_ERROR_ERROR;
    }
  if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    cleanup();
    return DIGHOLO_ERROR_ERROR;
  }
  int rc = DIGHOLO_ERROR_SUCCESS;
  (void)hipEventRecord(evs[0], s0);
  for (size_t i = 1; i < hs.size(); ++i)
    if (hs[i]->stream != s0) (void)hipStreamWaitEvent(hs[i]->stream, evs[0], 0);
  long items = 0;
  for (size_t i = 0; i < hs.size() && !rc; ++i) {
    hs[i]->capturing = true;
    hs[i]->collect = true;
    hs[i]->pendValid = hs[i]->pendOverlap = false;
    rc = enqueue_batch(*hs[i]);
    if (hs[i]->pendValid) items += (long)hs[i]->pend.fa.g.P * hs[i]->pend.fa.g.B;
  }
  for (size_t i = 1; i < hs.size(); ++i)
    if (hs[i]->stream != s0) {
      (void)hipEventRecord(evs[i], hs[i]->stream);
      (void)hipStreamWaitEvent(s0, evs[i], 0);
    }
  // the recorded window kernels, then the separate overlaps that follow some of them
  m.w3.clear();
  m.ops.clear();
  for (Handle* h : hs)
    if (h->pendValid) m.w3.push_back(h->pend);
  if (items <= kMultiMergeItems) {
    for (size_t i0 = 0; i0 < m.w3.size(); i0 += dh::kW3MaxBatches) {
      MultiOp o;
      o.kind = 0;
      o.i0 = (int)i0;
      o.n = (int)std::min<size_t>(dh::kW3MaxBatches, m.w3.size() - i0);
      m.ops.push_back(o);
    }
  } else {
    for (size_t i = 0; i < m.w3.size(); ++i) {
      MultiOp o;
      o.kind = 1;
      o.i0 = (int)i;
      m.ops.push_back(o);
    }
  }
  for (Handle* h : hs)
    if (h->pendValid && h->pendOverlap) {
      MultiOp o;
      o.kind = 2;
      o.g = h->pendGF;
      o.fields = h->fields.p;
      o.hgx = h->hgx.p;
      o.hgy = h->hgy.p;
      o.hgx16 = h->hgx16.p;
      o.hgy16 = h->hgy16.p;
      o.coefs = h->coefs.p;
      o.split = h->pendSplit ? 1 : 0;
      o.tpw = h->ovTpw;
      m.ops.push_back(o);
    }
  if (!rc) {
    const hipError_t e = run_multi_ops(m, s0);
    if (e != hipSuccess) {
      log_err(&h0, "multi-batch window kernels", e);
      rc = DIGHOLO_ERROR_ERROR;
    }
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s0, &graph);
  cleanup();
  if (rc || ec != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    fprintf(stderr, "digHolo: graph capture failed (%s)\n", hipGetErrorString(ec));
    return rc ? rc : DIGHOLO_ERROR_ERROR;
  }
  size_t nodes = 0;
  if (hipGraphGetNodes(graph, nullptr, &nodes) == hipSuccess && nodes == m.ops.size()) {
    std::vector<hipGraphNode_t> list(nodes);
    bool kernels = hipGraphGetNodes(graph, list.data(), &nodes) == hipSuccess;
    for (size_t i = 0; i < nodes && kernels; ++i) {
      hipGraphNodeType t;
      kernels = hipGraphNodeGetType(list[i], &t) == hipSuccess && t == hipGraphNodeTypeKernel;
    }
    m.direct = kernels;
  }
  if (m.direct) {
    (void)hipGraphDestroy(graph);
  } else {
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return DIGHOLO_ERROR_ERROR;
    }
    m.graph = graph;
    m.exec = exec;
  }
  m.stream = s0;
  if (hipEventCreateWithFlags(&m.done, hipEventDisableTiming) != hipSuccess) {
    multi_release(m);
    return DIGHOLO_ERROR_ERROR;
  }
  m.hs = hs;
  m.cfgs.clear();
  m.states.clear();
  for (Handle* h : hs) {
    m.cfgs.push_back(config_sans_arrays(h->cfg));
    m.states.push_back(launch_state(*h));
  }
  return DIGHOLO_ERROR_SUCCESS;
}

int process_multi(const std::vector<Handle*>& hs, const std::vector<int>& ids) {
  for (Handle* h : hs) {
    if (!h->frames) return DIGHOLO_ERROR_NULLPOINTER;
    if (!is_device_pointer(h->frames)) {
      fprintf(stderr, "digHolo: ProcessBatchMultiDevice needs device-resident frames\n");
      return DIGHOLO_ERROR_INVALIDARGUMENT;
    }
    if (int e = ensure_device(*h)) return e;
  }
  std::lock_guard<std::mutex> lk(g_multiMu);
  MultiGraph& m = g_multi[ids];
  Handle& h0 = *hs[0];
  hipStream_t s0 = h0.stream;
  bool same = (m.exec || m.direct) && m.stream == s0 && m.hs == hs;
  for (size_t i = 0; i < hs.size() && same; ++i)
    same = same_config(hs[i]->cfg, m.cfgs[i]) && launch_state(*hs[i]) == m.states[i];
  if (!same) {
    multi_release(m);
    // the eager batches run on every handle's stream: each first follows the
    // multi-batch launches that wrote its outputs
    for (Handle* h : hs)
      if (h->afterMulti) {
        (void)hipStreamWaitEvent(h->stream, h->afterMulti, 0);
        h->afterMulti = nullptr;
      }
    if (int e = capture_multi(m, hs)) {
      multi_release(m);
      return e;
    }
  } else {
    // s0 follows the work queued on the other handles' streams since their
    // last multi-batch launch; one whose stream still waits on this list's
    // last launch has had none (every API call on it would have taken that wait)
    if (h0.afterMulti) {
      (void)hipStreamWaitEvent(s0, h0.afterMulti, 0);
      h0.afterMulti = nullptr;
    }
    for (size_t i = 1; i < hs.size(); ++i) {
      Handle& h = *hs[i];
      if (h.stream == s0 || (m.done && h.afterMulti == m.done)) continue;
      if (h.afterMulti) {
        (void)hipStreamWaitEvent(h.stream, h.afterMulti, 0);
        h.afterMulti = nullptr;
      }
      if (!h.fence && hipEventCreateWithFlags(&h.fence, hipEventDisableTiming) != hipSuccess) {
        h.fence = nullptr;
        return DIGHOLO_ERROR_ERROR;
      }
      if (hipEventRecord(h.fence, h.stream) != hipSuccess || hipStreamWaitEvent(s0, h.fence, 0) != hipSuccess)
        return DIGHOLO_ERROR_ERROR;
    }
  }
  const hipError_t e = m.direct ? run_multi_ops(m, s0) : hipGraphLaunch(m.exec, s0);
  if (e != hipSuccess) {
    log_err(&h0, "multi-batch launch", e);
    return DIGHOLO_ERROR_ERROR;
  }
  // later work on any other handle's own stream (result downloads, the next
  // batch) follows this launch: the handle's next API call makes its stream wait
  if (hipEventRecord(m.done, s0) != hipSuccess) return DIGHOLO_ERROR_ERROR;
  for (size_t i = 1; i < hs.size(); ++i)
    if (hs[i]->stream != s0) hs[i]->afterMulti = m.done;
  for (Handle* h : hs) {   // every handle's outputs are now those of this launch on h0's stream
    h->coefsHostValid = h->fieldsHostValid = false;
    h->analysisValid[0] = h->analysisValid[1] = false;
    h->f16Valid = false;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

void destroy_handle(Handle* h) {
  if (!h) return;
  {
    std::lock_guard<std::mutex> lk(g_multiMu);
    for (auto it = g_multi.begin(); it != g_multi.end();) {
      const bool uses = std::find(it->second.hs.begin(), it->second.hs.end(), h) != it->second.hs.end();
      if (uses) {
        multi_release(it->second);
        it = g_multi.erase(it);
      } else {
        ++it;
      }
    }
    // the stream synchronisation below then covers a launch from another stream
    if (h->afterMulti) (void)hipStreamWaitEvent(h->stream, h->afterMulti, 0);
    h->afterMulti = nullptr;
  }
  if (h->devInit) {
    (void)hipStreamSynchronize(h->stream);
    h->staging.release();
    h->rowspec.release();
    h->win.release();
    h->fields.release();
    h->coefs.release();
    h->factor.release();
    h->ref.release();
    h->twW.release();
    h->hgx.release();
    h->hgy.release();
    h->mask.release();
    h->colTab.release();
    h->w3Tab.release();
    h->plane.release();
    h->winAll.release();
    h->tstage.release();
    h->calDev.release();
    h->hgx16.release();
    h->hgy16.release();
    for (auto& kv : h->ffts) kv.second->tw.release();
    h->hCoefs.release();
    h->hFields.release();
    h->metricRows.release();
    h->gram.release();
    h->coefsBackup.release();
    h->coefsMix.release();
    h->spec0.release();
    h->anaPw.release();
    h->anaSumDev.release();
    h->anaAxes.release();
    h->anaScale.release();
    h->anaInc.release();
    h->anaStats.release();
    h->f16.release();
    h->hF16.release();
    for (auto& set : h->evPool)
      for (auto& e : set) (void)hipEventDestroy(e);
    if (h->fence) (void)hipEventDestroy(h->fence);
    if (h->ownStream) (void)hipStreamDestroy(h->ownStream);
  }
  delete h;
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
#define H_OR(ret)        \
  Handle* h = get(handleIdx); \
  if (!h) return ret;
#define CFG (h->cfg)

EXT_C int digHoloCreate(void) {
  std::lock_guard<std::mutex> lk(g_registry);
  g_handles.push_back(new Handle());
  return (int)g_handles.size() - 1;           // handles are never reused (digHolo.cpp:13434-13446)
}

EXT_C int digHoloDestroy(int handleIdx) {
  Handle* h = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_registry);
    if (handleIdx < 0 || handleIdx >= (int)g_handles.size() || !g_handles[handleIdx])
      return DIGHOLO_ERROR_INVALIDHANDLE;
    h = g_handles[handleIdx];
    g_handles[handleIdx] = nullptr;
  }
  destroy_handle(h);
  return DIGHOLO_ERROR_SUCCESS;
}

// ---- frame buffer (digHolo.cpp:4235-4350) ----
EXT_C int digHoloSetFrameBuffer(int handleIdx, float* buffer) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!buffer) return DIGHOLO_ERROR_NULLPOINTER;
  h->frames = buffer;
  h->frameU16 = 0;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float* digHoloGetFrameBuffer(int handleIdx) {
  H_OR(nullptr);
  return h->frameU16 ? nullptr : (float*)h->frames;
}
EXT_C int digHoloSetFrameBufferUint16(int handleIdx, unsigned short* buffer, int transposeMode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!buffer) return DIGHOLO_ERROR_NULLPOINTER;
  h->frames = buffer;
  h->frameU16 = 1;
  h->transpose = transposeMode ? 1 : 0;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C unsigned short* digHoloGetFrameBufferUint16(int handleIdx, int* transposeMode) {
  H_OR(nullptr);
  if (transposeMode) *transposeMode = h->transpose;
  return h->frameU16 ? (unsigned short*)h->frames : nullptr;
}

// ---- frame config ----
EXT_C int digHoloConfigSetFrameDimensions(int handleIdx, int width, int height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (width % PIXEL_QUANTA || height % PIXEL_QUANTA || width < 0 || height < 0)   // digHolo.cpp:4154-4170
    return DIGHOLO_ERROR_INVALIDDIMENSION;
  CFG.frameWidth = width;
  CFG.frameHeight = height;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetFrameDimensions(int handleIdx, int* width, int* height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  *width = CFG.frameWidth;
  *height = CFG.frameHeight;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigSetFramePixelSize(int handleIdx, float pixelSize) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!(pixelSize > 0 && std::isfinite(pixelSize))) return DIGHOLO_ERROR_INVALIDDIMENSION;
  CFG.pixelSize = pixelSize;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetFramePixelSize(int handleIdx) {
  H_OR(0.f);
  return CFG.pixelSize;
}
EXT_C int digHoloConfigSetPolCount(int handleIdx, int polCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (polCount > 0 && polCount <= POLCOUNTMAX) {
    CFG.polCount = polCount;
    return DIGHOLO_ERROR_SUCCESS;
  }
  return DIGHOLO_ERROR_INVALIDPOLARISATION;
}
EXT_C int digHoloConfigGetPolCount(int handleIdx) {
  H_OR(0);
  return CFG.polCount;
}
EXT_C int digHoloConfigSetFillFactorCorrectionEnabled(int handleIdx, int enabled) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.fillFactor = !(enabled == 0);
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetFillFactorCorrectionEnabled(int handleIdx) {
  H_OR(0);
  return CFG.fillFactor;
}

// ---- batch (digHolo.cpp:4401-4467, :15073-15121) ----
static int set_batch(Handle* h, int batchCount, const void* buf, int u16, int avgCount, int avgMode, int transpose) {
  if (!buf) return DIGHOLO_ERROR_NULLPOINTER;
  h->frames = buf;
  h->frameU16 = u16;
  h->transpose = transpose ? 1 : 0;
  if (batchCount < 1) batchCount = 1;
  if (avgCount < 1) avgCount = 1;
  if (!u16 && (avgMode < 0 || avgMode > 1)) avgMode = 0;
  h->cfg.batchCount = batchCount;
  h->cfg.avgCount = avgCount;
  h->cfg.avgMode = avgMode;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloSetBatch(int handleIdx, int batchCount, float* frameBuffer) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  return set_batch(h, batchCount, frameBuffer, 0, 1, 0, 0);
}
EXT_C int digHoloSetBatchAvg(int handleIdx, int batchCount, float* frameBuffer, int avgCount, int avgMode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  return set_batch(h, batchCount, frameBuffer, 0, avgCount, avgMode, 0);
}
EXT_C int digHoloSetBatchUint16(int handleIdx, int batchCount, unsigned short* frameBuffer, int transpose) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  return set_batch(h, batchCount, frameBuffer, 1, 1, 0, transpose);
}
EXT_C int digHoloSetBatchAvgUint16(int handleIdx, int batchCount, unsigned short* frameBuffer, int avgCount,
                                   int avgMode, int transpose) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  return set_batch(h, batchCount, frameBuffer, 1, avgCount, avgMode, transpose);
}
EXT_C int digHoloConfigSetBatchCount(int handleIdx, int batchCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (batchCount < 1) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.batchCount = batchCount;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBatchCount(int handleIdx) {
  H_OR(0);
  return CFG.batchCount;
}
EXT_C int digHoloConfigSetBatchAvgCount(int handleIdx, int avgCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (avgCount < 1) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.avgCount = avgCount;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBatchAvgCount(int handleIdx) {
  H_OR(0);
  return CFG.avgCount;
}
EXT_C int digHoloConfigSetBatchAvgMode(int handleIdx, int avgMode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (avgMode < 0 || avgMode >= DIGHOLO_AVGMODE_COUNT) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.avgMode = avgMode;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBatchAvgMode(int handleIdx) {
  H_OR(0);
  return CFG.avgMode;
}

// ---- batch calibration (digHolo.cpp:4469-4530, :14917-14985) ----
EXT_C int digHoloConfigSetBatchCalibration(int handleIdx, complex64* cal, int polCount, int batchCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (polCount > 0 && batchCount > 0 && cal) {
    CFG.cal.resize((size_t)polCount * batchCount);
    std::memcpy(CFG.cal.data(), cal, CFG.cal.size() * sizeof(float2));
    CFG.calP = polCount;
    CFG.calB = batchCount;
    CFG.calEnabled = 1;
    ++CFG.calVersion;
    return DIGHOLO_ERROR_SUCCESS;
  }
  CFG.calEnabled = 0;
  return cal ? DIGHOLO_ERROR_INVALIDARGUMENT : DIGHOLO_ERROR_NULLPOINTER;
}
EXT_C complex64* digHoloConfigGetBatchCalibration(int handleIdx, int* polCount, int* batchCount) {
  Handle* h = get(handleIdx);
  if (!h) return nullptr;
  if (CFG.cal.empty()) {
    *polCount = 0;
    *batchCount = 0;
    return nullptr;
  }
  *polCount = CFG.calP;
  *batchCount = CFG.calB;
  return reinterpret_cast<complex64*>(CFG.cal.data());
}
EXT_C int digHoloConfigSetBatchCalibrationFromFile(int handleIdx, const char* fname, int polCount, int batchCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  FILE* f = fname ? std::fopen(fname, "rb") : nullptr;
  if (!f) return DIGHOLO_ERROR_FILENOTFOUND;
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  // the file must hold exactly polCount x batchCount complex64 (digHolo.cpp:4543-4580)
  const long long expect = (long long)polCount * batchCount * (long long)sizeof(float2);
  if (polCount < 1 || batchCount < 1 || bytes != expect) {
    std::fclose(f);
    CFG.calEnabled = 0;
    return DIGHOLO_ERROR_INVALIDDIMENSION;
  }
  std::vector<float2> buf((size_t)polCount * batchCount);
  const size_t got = std::fread(buf.data(), sizeof(float2), buf.size(), f);
  std::fclose(f);
  if (got != buf.size()) {
    CFG.calEnabled = 0;
    return DIGHOLO_ERROR_INVALIDDIMENSION;
  }
  return digHoloConfigSetBatchCalibration(handleIdx, reinterpret_cast<complex64*>(buf.data()), polCount, batchCount);
}
EXT_C int digHoloConfigSetBatchCalibrationEnabled(int handleIdx, int enabled) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.calEnabled = enabled ? 1 : 0;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBatchCalibrationEnabled(int handleIdx) {
  H_OR(0);
  return CFG.calEnabled;
}

// ---- reference-wave calibration (digHolo.cpp:5309-5520, :15361-15550) ----
// Intensity: a = sqrt(1 / (I / max I)), zero where that is not finite.
EXT_C int digHoloConfigSetRefCalibrationIntensity(int handleIdx, unsigned short* cal, int wavelengthCount, int width,
                                                  int height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  Config& c = CFG;
  if (!cal) {
    c.refCalEnabled = 0;
    return DIGHOLO_ERROR_NULLPOINTER;
  }
  if (width <= 0 || height <= 0 || wavelengthCount <= 0) {
    c.refCalEnabled = 0;
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  const size_t n = (size_t)width * height * wavelengthCount;
  int maxV = 0;
  for (size_t i = 0; i < n; ++i) maxV = std::max<int>(maxV, cal[i]);
  const float maxScale = 1.0f / (float)maxV;
  c.refCal.resize(n);
  for (size_t i = 0; i < n; ++i) {
    float v = 1.0f / ((float)cal[i] * maxScale);
    if (std::isnan(v) || std::isinf(v)) v = 0;
    c.refCal[i] = make_float2(std::sqrt(v), 0.f);
  }
  c.refCalW = width;
  c.refCalH = height;
  c.refCalL = wavelengthCount;
  ++c.refCalVersion;
  c.refCalEnabled = 1;
  return DIGHOLO_ERROR_SUCCESS;
}

// Field: normalised to a peak magnitude of 1, then each part inverted as the
// reference does it, (1 / re, -1 / im), zero where either is not finite.
EXT_C int digHoloConfigSetRefCalibrationField(int handleIdx, complex64* cal, int wavelengthCount, int width,
                                              int height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  Config& c = CFG;
  if (!cal) {
    c.refCalEnabled = 0;
    return DIGHOLO_ERROR_NULLPOINTER;
  }
  if (width <= 0 || height <= 0 || wavelengthCount <= 0) {
    c.refCalEnabled = 0;
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  const size_t n = (size_t)width * height * wavelengthCount;
  const float2* z = reinterpret_cast<const float2*>(cal);
  float maxV = 0;
  for (size_t i = 0; i < n; ++i) maxV = std::max(maxV, z[i].x * z[i].x + z[i].y * z[i].y);
  maxV = std::sqrt(maxV);
  const float maxScale = 1.0f / maxV;
  c.refCal.resize(n);
  for (size_t i = 0; i < n; ++i) {
    float vR = 1.0f / (z[i].x * maxScale);
    float vI = -1.0f / (z[i].y * maxScale);
    if (std::isnan(vR) || std::isinf(vR) || std::isnan(vI) || std::isinf(vI)) vR = vI = 0;
    c.refCal[i] = make_float2(vR, vI);
  }
  c.refCalW = width;
  c.refCalH = height;
  c.refCalL = wavelengthCount;
  ++c.refCalVersion;
  c.refCalEnabled = 1;
  return DIGHOLO_ERROR_SUCCESS;
}

// The file holds the intensity (uint16) or the field (complex64) form; which
// one follows from its size (digHolo.cpp:5320-5417, including its rules for a
// file of one frame).
EXT_C int digHoloConfigSetRefCalibrationFromFile(int handleIdx, const char* fname, int wavelengthCount, int width,
                                                 int height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  FILE* f = fname ? std::fopen(fname, "rb") : nullptr;
  if (!f) {
    CFG.refCalEnabled = 0;
    return DIGHOLO_ERROR_FILENOTFOUND;
  }
  std::fseek(f, 0, SEEK_END);
  const long long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int L = wavelengthCount;
  const long long frame16 = (long long)width * height * (long long)sizeof(unsigned short);
  const long long frame32 = (long long)width * height * (long long)sizeof(float2);
  long long total16 = frame16 * L, total32 = frame32 * L;
  if (bytes == frame16) {
    L = 1;
    total16 = frame16;
  } else if (bytes == frame32 && L != 2) {
    L = 1;
    total32 = frame32;
  }
  int e = DIGHOLO_ERROR_SUCCESS;
  if (bytes == total16 && bytes > 0) {
    std::vector<unsigned short> buf((size_t)(bytes / 2));
    const size_t got = std::fread(buf.data(), 2, buf.size(), f);
    e = got == buf.size() ? digHoloConfigSetRefCalibrationIntensity(handleIdx, buf.data(), L, width, height)
                          : DIGHOLO_ERROR_INVALIDDIMENSION;
  } else if (bytes == total32 && bytes > 0) {
    std::vector<float2> buf((size_t)(bytes / 8));
    const size_t got = std::fread(buf.data(), 8, buf.size(), f);
    e = got == buf.size()
            ? digHoloConfigSetRefCalibrationField(handleIdx, reinterpret_cast<complex64*>(buf.data()), L, width, height)
            : DIGHOLO_ERROR_INVALIDDIMENSION;
  } else {
    e = DIGHOLO_ERROR_INVALIDDIMENSION;
  }
  std::fclose(f);
  if (e) CFG.refCalEnabled = 0;
  return e;
}

EXT_C int digHoloConfigSetRefCalibrationEnabled(int handleIdx, int enabled) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.refCalEnabled = enabled ? 1 : 0;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetRefCalibrationEnabled(int handleIdx) {
  H_OR(0);
  return CFG.refCalEnabled;
}

// The calibration plane the last batch applied, [wavelengthCount][pol][height][width]
// on the field plane's axes; null before a calibrated batch.
EXT_C complex64* digHoloConfigGetRefCalibrationFields(int handleIdx, int* wavelengthCount, int* polCount, float** x,
                                                      float** y, int* width, int* height) {
  Handle* h = get(handleIdx);
  auto none = [&]() -> complex64* {
    if (wavelengthCount) *wavelengthCount = 0;
    if (polCount) *polCount = 0;
    if (x) *x = nullptr;
    if (y) *y = nullptr;
    if (width) *width = 0;
    if (height) *height = 0;
    return nullptr;
  };
  if (!h || !h->calPlaneValid || !h->haveBatch || !h->cur.cal.active) return none();
  const dh::Geometry& g = h->cur.g;
  const size_t n = (size_t)h->calPlaneL * g.P * g.fh * g.fw;
  if (h->hCal.ensure(n) != hipSuccess) return none();
  if (hipMemcpyAsync(h->hCal.p, h->calPlane.p, n * sizeof(float2), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return none();
  if (wavelengthCount) *wavelengthCount = h->calPlaneL;
  if (polCount) *polCount = g.P;
  if (x) *x = h->xAxis.data();
  if (y) *y = h->yAxis.data();
  if (width) *width = g.fw;
  if (height) *height = g.fh;
  return reinterpret_cast<complex64*>(h->hCal.p);
}

// ---- windows (digHolo.cpp:14234-14330, :13518-13540, :14414-14436, :14589-14611) ----
EXT_C int digHoloConfigSetfftWindowSize(int handleIdx, int width, int height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (width < 0 || height < 0) return DIGHOLO_ERROR_INVALIDDIMENSION;
  CFG.fftWindowSizeX = PIXEL_QUANTA * (width / PIXEL_QUANTA);
  CFG.fftWindowSizeY = PIXEL_QUANTA * (height / PIXEL_QUANTA);
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetfftWindowSize(int handleIdx, int* width, int* height) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  *width = CFG.fftWindowSizeX;
  *height = CFG.fftWindowSizeY;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigSetfftWindowSizeX(int handleIdx, int width) {
  H_OR(0);
  if (width < 0) return 0;
  CFG.fftWindowSizeX = PIXEL_QUANTA * (width / PIXEL_QUANTA);
  return CFG.fftWindowSizeX;                  // returns the value set (digHolo.cpp:14234-14248)
}
EXT_C int digHoloConfigGetfftWindowSizeX(int handleIdx) {
  H_OR(0);
  return CFG.fftWindowSizeX;
}
EXT_C int digHoloConfigSetfftWindowSizeY(int handleIdx, int height) {
  H_OR(0);
  CFG.fftWindowSizeY = PIXEL_QUANTA * (height / PIXEL_QUANTA);
  return CFG.fftWindowSizeY;
}
EXT_C int digHoloConfigGetfftWindowSizeY(int handleIdx) {
  H_OR(0);
  return CFG.fftWindowSizeY;
}
EXT_C int digHoloConfigSetFourierWindowRadius(int handleIdx, float windowRadius) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (windowRadius > 0) {
    CFG.apertureSize = windowRadius;
    return DIGHOLO_ERROR_SUCCESS;
  }
  return DIGHOLO_ERROR_INVALIDDIMENSION;
}
EXT_C float digHoloConfigGetFourierWindowRadius(int handleIdx) {
  H_OR(0.f);
  return CFG.apertureSize;
}
EXT_C int digHoloConfigSetIFFTResolutionMode(int handleIdx, int resolutionModeIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (resolutionModeIdx == 0 || resolutionModeIdx == 1) {
    CFG.resolutionIdx = resolutionModeIdx;
    return DIGHOLO_ERROR_SUCCESS;
  }
  return DIGHOLO_ERROR_INVALIDARGUMENT;
}
EXT_C int digHoloConfigGetIFFTResolutionMode(int handleIdx) {
  H_OR(0);
  return CFG.resolutionIdx;
}
EXT_C int digHoloConfigSetWavelengthCentre(int handleIdx, float lambda0) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (lambda0 > 0 && std::isfinite(lambda0)) {
    CFG.wavelengthCentre = lambda0;
    return DIGHOLO_ERROR_SUCCESS;
  }
  return DIGHOLO_ERROR_INVALIDARGUMENT;
}
EXT_C float digHoloConfigGetWavelengthCentre(int handleIdx) {
  H_OR(0.f);
  return CFG.wavelengthCentre;
}

// Wavelength axis (digHolo.cpp:6578-6680, :14611-14700).  The axis applies
// from the next batch on; one entry means the centre wavelength.
EXT_C int digHoloConfigSetWavelengths(int handleIdx, float* wavelengths, int wavelengthCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (wavelengthCount <= 0) return DIGHOLO_ERROR_INVALIDARGUMENT;
  if (!wavelengths) return DIGHOLO_ERROR_NULLPOINTER;
  CFG.wavelengths.assign(wavelengths, wavelengths + wavelengthCount);
  return DIGHOLO_ERROR_SUCCESS;
}

namespace {
// Linear in optical frequency f = c0 / lambda, in float32 as the reference
// evaluates it (digHolo.cpp:6577-6650; c0 an int converted to float; a
// wavelength above 1 is taken in the unit DIGHOLO_UNIT_LAMBDA).
int set_wavelengths_linear_frequency(Config& c, float lamStart, float lamStop, int count) {
  if (count <= 0) return DIGHOLO_ERROR_INVALIDDIMENSION;
  const float c0 = (float)299792458;
  if (lamStart > 1) lamStart = lamStart * DIGHOLO_UNIT_LAMBDA;
  if (lamStop > 1) lamStop = lamStop * DIGHOLO_UNIT_LAMBDA;
  float fStart = c0 / lamStart;
  const float fStop = c0 / lamStop;
  float df = (fStop - fStart) / (float)(count - 1);
  if (count <= 1) {
    // one step: the start wavelength's double frequency when only it is valid
    // (either single-ended case, as the reference writes both), the centre
    // wavelength when neither is, else the mean of the two
    if (lamStop <= 0 && lamStart > 0) fStart = c0 / (lamStart / 2);
    else if (lamStart <= 0 && lamStop > 0) fStart = c0 / (lamStart / 2);
    else if (lamStart <= 0 && lamStop <= 0) fStart = c0 / (float)(c.wavelengthCentre * DIGHOLO_UNIT_LAMBDA);
    else fStart = c0 / ((lamStop + lamStart) / 2);
    df = 0;
  }
  c.wavelengths.resize(count);
  for (int i = 0; i < count; ++i) {
    c.wavelengths[i] = c0 / (fStart + df * (float)i);   // built without FMA contraction
  }
  return DIGHOLO_ERROR_SUCCESS;
}
}  // namespace

EXT_C int digHoloConfigSetWavelengthsLinearFrequency(int handleIdx, float wavelengthStart, float wavelengthStop,
                                                      int wavelengthCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!wavelengthCount) return DIGHOLO_ERROR_INVALIDARGUMENT;
  return set_wavelengths_linear_frequency(CFG, wavelengthStart, wavelengthStop, wavelengthCount);
}

// The axis the next batch uses (the centre wavelength alone for fewer than two entries).
EXT_C float* digHoloConfigGetWavelengths(int handleIdx, int* wavelengthCount) {
  Handle* h = get(handleIdx);
  if (!h) return nullptr;
  h->lamAxis = h->cfg.wavelengths.size() > 1 ? h->cfg.wavelengths : std::vector<float>{h->cfg.wavelengthCentre};
  if (wavelengthCount) *wavelengthCount = (int)h->lamAxis.size();
  return h->lamAxis.data();
}

EXT_C int digHoloConfigSetWavelengthOrdering(int handleIdx, int inoutIdx, int ordering) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if ((inoutIdx != DIGHOLO_WAVELENGTHORDER_INPUT && inoutIdx != DIGHOLO_WAVELENGTHORDER_OUTPUT) ||
      (ordering != DIGHOLO_WAVELENGTHORDER_FAST && ordering != DIGHOLO_WAVELENGTHORDER_SLOW))
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.wavelengthOrdering[inoutIdx] = ordering;
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C int digHoloConfigGetWavelengthOrdering(int handleIdx, int inoutIdx) {
  H_OR(0);
  if (inoutIdx != DIGHOLO_WAVELENGTHORDER_INPUT && inoutIdx != DIGHOLO_WAVELENGTHORDER_OUTPUT) return 0;
  return CFG.wavelengthOrdering[inoutIdx];
}

// ---- AutoAlign switches (digHolo.cpp:13542-13735): non-zero stored as 1 ----
#define AA_FLAG(Name, field)                                         \
  EXT_C int digHoloConfigSet##Name(int handleIdx, int v) {           \
    H_OR(DIGHOLO_ERROR_INVALIDHANDLE);                               \
    CFG.field = v ? 1 : 0;                                           \
    return DIGHOLO_ERROR_SUCCESS;                                    \
  }                                                                  \
  EXT_C int digHoloConfigGet##Name(int handleIdx) {                  \
    H_OR(0);                                                         \
    return CFG.field;                                                \
  }
AA_FLAG(AutoAlignBeamCentre, aaCentre)
AA_FLAG(AutoAlignDefocus, aaDefocus)
AA_FLAG(AutoAlignTilt, aaTilt)
AA_FLAG(AutoAlignBasisWaist, aaWaist)
AA_FLAG(AutoAlignFourierWindowRadius, aaFourierRadius)
AA_FLAG(AutoAlignPolIndependence, aaPolIndependence)
AA_FLAG(AutoAlignBasisMulConjTrans, aaMulConjTrans)
#undef AA_FLAG

EXT_C int digHoloConfigSetAutoAlignTol(int handleIdx, float tolerance) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!std::isfinite(tolerance)) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.aaTol = tolerance;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetAutoAlignTol(int handleIdx) {
  H_OR(0.f);
  return CFG.aaTol;
}
EXT_C int digHoloConfigSetAutoAlignMode(int handleIdx, int alignMode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (alignMode < 0 || alignMode >= DIGHOLO_AUTOALIGNMODE_COUNT) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.aaMode = alignMode;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetAutoAlignMode(int handleIdx) {
  H_OR(0);
  return CFG.aaMode;
}
EXT_C int digHoloConfigSetAutoAlignGoalIdx(int handleIdx, int goalIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (goalIdx < 0 || goalIdx >= DIGHOLO_METRIC_COUNT) return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.aaGoalIdx = goalIdx;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetAutoAlignGoalIdx(int handleIdx) {
  H_OR(0);
  return CFG.aaGoalIdx;
}

// ---- alignment parameters (digHolo.cpp:13736-13810, :14114-14230) ----
EXT_C int digHoloConfigSetBeamCentre(int handleIdx, int axisIdx, int polIdx, float value) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!std::isfinite(value)) return DIGHOLO_ERROR_INVALIDARGUMENT;
  if (!(polIdx < CFG.polCount && polIdx >= 0)) return DIGHOLO_ERROR_INVALIDPOLARISATION;
  if (axisIdx == 0) CFG.beamCX[polIdx] = value;
  else if (axisIdx == 1) CFG.beamCY[polIdx] = value;
  else return DIGHOLO_ERROR_INVALIDAXIS;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetBeamCentre(int handleIdx, int axisIdx, int polIdx) {
  H_OR(0.f);
  if (!(polIdx < CFG.polCount && polIdx >= 0)) return 0.f;
  if (axisIdx == 0) return CFG.beamCX[polIdx];
  if (axisIdx == 1) return CFG.beamCY[polIdx];
  return 0.f;
}
EXT_C int digHoloConfigSetTilt(int handleIdx, int axisIdx, int polIdx, float value) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!std::isfinite(value)) return DIGHOLO_ERROR_INVALIDARGUMENT;
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return DIGHOLO_ERROR_INVALIDPOLARISATION;
  if (axisIdx == 0) CFG.tiltX[polIdx] = value;
  else if (axisIdx == 1) CFG.tiltY[polIdx] = value;
  else return DIGHOLO_ERROR_INVALIDAXIS;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetTilt(int handleIdx, int axisIdx, int polIdx) {
  H_OR(0.f);
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return 0.f;
  if (axisIdx == 0) return CFG.tiltX[polIdx];
  if (axisIdx == 1) return CFG.tiltY[polIdx];
  return 0.f;
}
EXT_C int digHoloConfigSetDefocus(int handleIdx, int polIdx, float value) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return DIGHOLO_ERROR_INVALIDPOLARISATION;
  CFG.defocus[polIdx] = value;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetDefocus(int handleIdx, int polIdx) {
  H_OR(0.f);
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return 0.f;
  return CFG.defocus[polIdx];
}
#define POL_LOCK(Name, field)                                        \
  EXT_C int digHoloConfigSet##Name(int handleIdx, int v) {           \
    H_OR(DIGHOLO_ERROR_INVALIDHANDLE);                               \
    CFG.field = v ? 1 : 0;                                           \
    return DIGHOLO_ERROR_SUCCESS;                                    \
  }                                                                  \
  EXT_C int digHoloConfigGet##Name(int handleIdx) {                  \
    H_OR(0);                                                         \
    return CFG.field;                                                \
  }
POL_LOCK(PolLockTilt, polLockTilt)
POL_LOCK(PolLockDefocus, polLockDefocus)
POL_LOCK(PolLockBasisWaist, polLockWaist)
#undef POL_LOCK

// ---- basis (digHolo.cpp:14023-14071, :14717-14790) ----
EXT_C int digHoloConfigSetBasisGroupCount(int handleIdx, int groupCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.maxMG = std::abs(groupCount);
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBasisGroupCount(int handleIdx) {
  H_OR(0);
  return CFG.maxMG;
}
EXT_C int digHoloConfigSetBasisWaist(int handleIdx, int polIdx, float value) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!(value > 0 && std::isfinite(value))) return DIGHOLO_ERROR_INVALIDDIMENSION;
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return DIGHOLO_ERROR_INVALIDPOLARISATION;
  for (int p = 0; p < CFG.polCount; ++p) CFG.waist[p] = value;   // all pols, as digHolo.cpp:14058-14063
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloConfigGetBasisWaist(int handleIdx, int polIdx) {
  H_OR(0.f);
  if (!(polIdx < POLCOUNTMAX && polIdx >= 0)) return 0.f;
  return CFG.waist[polIdx];
}
EXT_C int digHoloConfigSetBasisTypeHG(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.basisType = DIGHOLO_BASISTYPE_HG;
  return DIGHOLO_ERROR_SUCCESS;
}
// digHolo.cpp:14736-14742 stores any value and returns it; processing treats
// an unsupported type as HG (:9474-9481)
EXT_C int digHoloConfigSetBasisType(int handleIdx, int basisType) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.basisType = basisType;
  return basisType;
}
EXT_C int digHoloConfigSetBasisTypeLG(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.basisType = DIGHOLO_BASISTYPE_LG;
  return DIGHOLO_ERROR_SUCCESS;
}
// digHolo.cpp:14774-14803, :4750-4765: the matrix is copied
EXT_C int digHoloConfigSetBasisTypeCustom(int handleIdx, int modeCountIn, int modeCountOut, complex64* transform) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!transform) return DIGHOLO_ERROR_NULLPOINTER;
  if (modeCountIn <= 0 || modeCountOut <= 0) return DIGHOLO_ERROR_INVALIDDIMENSION;
  CFG.basisType = DIGHOLO_BASISTYPE_CUSTOM;
  const size_t n = (size_t)modeCountIn * modeCountOut;
  CFG.basisFull.resize(n);
  std::memcpy(CFG.basisFull.data(), transform, n * sizeof(float2));
  CFG.basisIn = modeCountIn;
  CFG.basisOut = modeCountOut;
  ++CFG.basisVersion;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetBasisType(int handleIdx) {
  H_OR(0);
  return CFG.basisType;
}

// ---- processing ----
EXT_C complex64* digHoloProcessBatch(int handleIdx, int* batchCount, int* modeCount, int* polCount) {
  Handle* h = get(handleIdx);
  if (!h) {
    *batchCount = 0;
    *modeCount = 0;
    *polCount = 0;
    return nullptr;
  }
  int e = enqueue_batch(*h);
  if (!e) e = download_coefs(*h);
  collect_stage_times(*h);
  if (e) {
    *batchCount = 0;
    *modeCount = 0;
    *polCount = 0;
    return nullptr;
  }
  const dh::Geometry& g = h->cur.g;
  *batchCount = g.B;
  *polCount = g.P;
  *modeCount = h->Mout;
  return h->Mout > 0 ? reinterpret_cast<complex64*>(h->hCoefs.p) : nullptr;   // G=0: no decomposition
}

// A batch whose frames follow a wavelength axis: the axis is set first (its
// error, if any, leaves the previous axis in place), then the batch runs
// (digHolo.cpp:5046-5076, :15194-15236).
EXT_C complex64* digHoloProcessBatchFrequencySweepLinear(int handleIdx, int* batchCount, int* modeCount, int* polCount,
                                                         float lambdaStart, float lambdaStop, int lambdaCount) {
  if (Handle* h = get(handleIdx)) (void)set_wavelengths_linear_frequency(h->cfg, lambdaStart, lambdaStop, lambdaCount);
  return digHoloProcessBatch(handleIdx, batchCount, modeCount, polCount);
}

EXT_C complex64* digHoloProcessBatchWavelengthSweepArbitrary(int handleIdx, int* batchCount, int* modeCount,
                                                             int* polCount, float* wavelengths, int lambdaCount) {
  if (Handle* h = get(handleIdx))
    if (lambdaCount > 0 && wavelengths) h->cfg.wavelengths.assign(wavelengths, wavelengths + lambdaCount);
  return digHoloProcessBatch(handleIdx, batchCount, modeCount, polCount);
}

EXT_C complex64* digHoloBasisGetCoefs(int handleIdx, int* batchCount, int* modeCount, int* polCount) {
  Handle* h = get(handleIdx);
  if (!h || !h->haveBatch || h->cur.g.M == 0 || download_coefs(*h)) {
    *batchCount = 0;
    *modeCount = 0;
    *polCount = 0;
    return nullptr;
  }
  const dh::Geometry& g = h->cur.g;
  *batchCount = g.B;
  *polCount = g.P;
  *modeCount = h->Mout;
  return reinterpret_cast<complex64*>(h->hCoefs.p);
}

// Basis fields of the rows of W (modes x M HG weights) on the field plane g
// into the pinned host buffer h.hBasisFields, [modes][pol][g.h][g.w].
int synth_basis_fields(Handle& h, const dh::Geometry& g, const std::vector<float2>& W, int modes, int M) {
  const size_t nOut = (size_t)modes * g.P * g.h * g.w;
  DH_TRY(h.basisW.ensure(W.size()), "alloc");
  DH_TRY(h.basisFieldsDev.ensure(nOut), "alloc");
  DH_TRY(h.hBasisFields.ensure(nOut), "pinned alloc");
  DH_TRY(hipMemcpyAsync(h.basisW.p, W.data(), W.size() * sizeof(float2), hipMemcpyHostToDevice, h.stream), "upload");
  DH_TRY(dh::launch_basis_synth(g, h.basisW.p, modes, M, h.hgx.p, h.hgy.p, h.basisFieldsDev.p, h.stream),
         "basis fields");
  DH_TRY(hipMemcpyAsync(h.hBasisFields.p, h.basisFieldsDev.p, nOut * sizeof(float2), hipMemcpyDeviceToHost, h.stream),
         "D2H");
  DH_TRY(hipStreamSynchronize(h.stream), "sync");
  return DIGHOLO_ERROR_SUCCESS;
}

// digHolo.cpp:5233-5296, :15378-15394.  The fields of the output basis modes
// on each pol's field plane of the last batch, [modes][pol][height][width]:
// HG modes, LG modes (rows conj(U_g), the inverse transform of the identity)
// or the custom matrix's rows T(j, :) as given (unconjugated, as the reference
// builds them), synthesised on the GPU from the FP64-derived HG tables.
EXT_C complex64* digHoloBasisGetFields(int handleIdx, int* modeCount, int* polCount, float** x, float** y,
                                       int* width, int* height) {
  Handle* h = get(handleIdx);
  auto fail = [&]() -> complex64* {
    *modeCount = *polCount = *width = *height = 0;
    *x = *y = nullptr;
    return nullptr;
  };
  if (!h || !h->haveBatch || !h->tablesValid || h->cur.g.G <= 0) return fail();
  const Config& c = h->cfg;
  dh::Geometry g = h->cur.g;
  g.w = g.fw;
  g.h = g.fh;
  const int M = g.M;
  int Mo = M;
  std::vector<float2> W;
  if (c.basisType == DIGHOLO_BASISTYPE_LG) {
    W.assign((size_t)M * M, make_float2(0.f, 0.f));
    const std::vector<float2> t = lg_table(g.G);
    size_t off = 0;
    for (int gg = 0; gg < g.G; ++gg) {
      const int g0 = gg * (gg + 1) / 2;
      for (int k = 0; k <= gg; ++k)
        for (int cc = 0; cc <= gg; ++cc) {
          const float2 u = t[off + (size_t)k * (gg + 1) + cc];
          W[(size_t)(g0 + k) * M + g0 + cc] = make_float2(u.x, -u.y);
        }
      off += (size_t)(gg + 1) * (gg + 1);
    }
  } else if (c.basisType == DIGHOLO_BASISTYPE_CUSTOM && !c.basisFull.empty() && c.basisIn > 0 && c.basisOut > 0) {
    const int nIn = std::min(M, c.basisIn);
    Mo = std::min(M, c.basisOut);
    W.assign((size_t)Mo * M, make_float2(0.f, 0.f));
    for (int j = 0; j < Mo; ++j)
      for (int i = 0; i < nIn; ++i) W[(size_t)j * M + i] = c.basisFull[(size_t)i * c.basisOut + j];
  } else {
    W.assign((size_t)M * M, make_float2(0.f, 0.f));
    for (int i = 0; i < M; ++i) W[(size_t)i * M + i] = make_float2(1.f, 0.f);
  }
  if (synth_basis_fields(*h, g, W, Mo, M)) return fail();
  *modeCount = Mo;
  *polCount = g.P;
  *width = g.fw;
  *height = g.fh;
  *x = h->xAxis.data();
  *y = h->yAxis.data();
  return reinterpret_cast<complex64*>(h->hBasisFields.p);
}

EXT_C complex64* digHoloGetFields(int handleIdx, int* batchCount, int* polCount, float** x, float** y, int* width,
                                  int* height) {
  Handle* h = get(handleIdx);
  if (!h || !h->haveBatch || download_fields(*h)) {
    *batchCount = 0;
    *polCount = 0;
    *x = nullptr;
    *y = nullptr;
    *width = 0;
    *height = 0;
    return nullptr;
  }
  const dh::Geometry& g = h->cur.g;
  *batchCount = g.B;
  *polCount = g.P;
  *width = g.fw;
  *height = g.fh;
  *x = h->xAxis.data();
  *y = h->yAxis.data();
  return reinterpret_cast<complex64*>(h->hFields.p);
}

// ---- threads / console ----
EXT_C int digHoloConfigSetThreadCount(int handleIdx, int threadCount) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (threadCount < 1) threadCount = 1;      // accepted and stored; GPU work ignores it
  CFG.threadCount = threadCount;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetThreadCount(int handleIdx) {
  H_OR(0);
  return CFG.threadCount;
}
EXT_C int digHoloConfigSetVerbosity(int handleIdx, int verbosity) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  CFG.verbosity = verbosity;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetVerbosity(int handleIdx) {
  H_OR(0);
  return CFG.verbosity;
}

// digHoloBenchmark (digHolo.cpp:13241-13408): after an untimed warm-up batch,
// repeated batches for about goalDuration seconds.  Returns batches per second
// end to end (the reference's hzTotalMeas, :13345, :13396) and fills info[0..6]
// (7 floats, as the reference writes, :13374-13382): calls per second of the
// FFT, IFFT, ApplyTilt and overlap stages; info[3] the rate at which the HG
// basis is regenerated with a changing waist (:13318-13333); info[5] the rate
// the stage times add up to, 1 / sum of each kernel's time per batch
// (:13348); info[6] the measured end-to-end rate again.  Stages are HIP-event
// intervals on the handle's stream, taken on a second set of batches so the
// end-to-end time carries no event records between kernels.  Where one kernel
// runs several stages (the fused window kernel: FFT, window, IFFT and tilt,
// and with the fused overlap the coefficients too), each of those stages is
// reported at that kernel's rate and the kernel is counted once in info[5].
EXT_C float digHoloBenchmark(int handleIdx, float goalDuration, float* info) {
  Handle* h = get(handleIdx);
  if (!h || !(goalDuration > 0.f)) return 0.f;
  const bool prevTiming = h->stageTiming;
  h->stageTiming = false;
  if (enqueue_batch(*h) || hipStreamSynchronize(h->stream) != hipSuccess) {
    h->stageTiming = prevTiming;
    return 0.f;
  }
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return 0.f;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return 0.f;
  }
  // the first batch sizes the run: about goalDuration seconds in total
  float ms1 = 0.f;
  (void)hipEventRecord(a, h->stream);
  int rc = enqueue_batch(*h);
  (void)hipEventRecord(b, h->stream);
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(&ms1, a, b);
  const double est = std::max(1e-6, ms1 * 1e-3);
  const int iters = (int)std::min(1e6, std::max(3.0, std::ceil(goalDuration / est)));
  float msAll = 0.f;
  (void)hipEventRecord(a, h->stream);
  for (int i = 0; i < iters && !rc; ++i) rc = enqueue_batch(*h);
  (void)hipEventRecord(b, h->stream);
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(&msAll, a, b);
  const double hzTotal = (!rc && msAll > 0.f) ? iters / (msAll * 1e-3) : 0.0;
  // stage intervals on a smaller set of batches (at most a tenth of the run)
  double st[5] = {0, 0, 0, 0, 0};
  const int sIters = std::max(3, std::min(iters, 200));
  if (!rc) {
    h->stageTiming = true;
    collect_stage_times(*h);   // drop anything pending
    for (int i = 0; i < sIters && !rc; ++i) rc = enqueue_batch(*h);
    collect_stage_times(*h);
    for (int k = 0; k < 5; ++k) st[k] = h->stageMs[k];
  }
  h->stageTiming = prevTiming;
  // HG basis regeneration with the waist moving by 1e-5 of itself per call, as
  // the reference's loop does, into work buffers (the handle's tables stay)
  double hzBasis = 0.0;
  const dh::Geometry& g = h->cur.g;
  if (!rc && g.G > 0) {
    DevBuf<float> bx, by;
    if (bx.ensure((size_t)g.P * g.G * g.fw) == hipSuccess && by.ensure((size_t)g.P * g.G * g.fh) == hipSuccess) {
      dh::TableArgs ta = h->cur.tab;
      const int n2 = std::max(3, std::min(iters * 100, 20000));
      const double step = ta.pp[0].waist / 100000.0;
      float msB = 0.f;
      (void)hipEventRecord(a, h->stream);
      for (int i = 0; i < n2; ++i) {
        for (int p = 0; p < g.P; ++p) ta.pp[p].waist += step;
        if (dh::launch_hg_basis(ta, bx.p, by.p, h->stream) != hipSuccess) break;
      }
      (void)hipEventRecord(b, h->stream);
      (void)hipEventSynchronize(b);
      (void)hipEventElapsedTime(&msB, a, b);
      hzBasis = msB > 0.f ? n2 / (msB * 1e-3) : 0.0;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return 0.f;
  if (info) {
    // stage k of stageMs: row FFT (or the fused window kernel), column FFT +
    // window, IFFT + tilt, tables, overlap
    const bool fused = h->lastPath > 0;
    const bool fusedOv = fused && h->lastFusedOverlap;
    const double fftMs = fused ? st[0] : st[0] + st[1];
    const double ifftMs = fused ? st[0] : st[2];
    const double ovMs = fusedOv ? st[0] : st[4];
    const double sumMs = fused ? st[0] + (fusedOv ? 0.0 : st[4]) : st[0] + st[1] + st[2] + st[4];
    auto hz = [](double ms) { return ms > 0 ? (float)(1e3 / ms) : 0.f; };
    info[DIGHOLO_BENCHMARK_FFT] = hz(fftMs);
    info[DIGHOLO_BENCHMARK_IFFT] = hz(ifftMs);
    info[DIGHOLO_BENCHMARK_APPLYTILT] = hz(ifftMs);   // the tilt runs in the IFFT kernel on every path
    info[DIGHOLO_BENCHMARK_BASIS] = (float)hzBasis;
    info[DIGHOLO_BENCHMARK_OVERLAP] = g.G > 0 ? hz(ovMs) : 0.f;
    info[DIGHOLO_BENCHMARK_TOTAL] = hz(sumMs);
    info[DIGHOLO_BENCHMARK_TOTAL + 1] = (float)hzTotal;
  }
  return (float)hzTotal;
}

// ---- extensions ----
EXT_C int digHoloConfigSetOverlapPrecision(int handleIdx, int precision) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (precision != DIGHOLO_PRECISION_FP32 && precision != DIGHOLO_PRECISION_SPLIT16)
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  CFG.precision = precision;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetOverlapPrecision(int handleIdx) {
  H_OR(0);
  return CFG.precision;
}

EXT_C int digHoloConfigSetAutoTune(int handleIdx, int mode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (mode < 0 || mode > 2) return DIGHOLO_ERROR_INVALIDARGUMENT;
  h->autoTune = mode;
  if (!mode) {   // back to the default kernels
    h->windowMode = 3;
    h->fusedOverlap = -1;
    h->ovTpw = 0;
  }
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetAutoTune(int handleIdx) {
  H_OR(0);
  return h->autoTune;
}
EXT_C int digHoloProcessBatchMultiDevice(int count, const int* handleIdx) {
  if (count < 1 || !handleIdx) return DIGHOLO_ERROR_INVALIDARGUMENT;
  std::vector<Handle*> hs;
  std::vector<int> ids(handleIdx, handleIdx + count);
  for (int i = 0; i < count; ++i) {
    Handle* h = get_raw(ids[i]);   // process_multi orders the streams itself
    if (!h) return DIGHOLO_ERROR_INVALIDHANDLE;
    if (std::find(hs.begin(), hs.end(), h) != hs.end()) return DIGHOLO_ERROR_INVALIDARGUMENT;
    hs.push_back(h);
  }
  return process_multi(hs, ids);
}

// Allocation-free form of digHoloProcessBatchDevice (see the header): the
// staged kernels with their intermediates in the per-GPU scratch array, on the
// tables and plans of the handle's last processed configuration.
EXT_C int digHoloProcessBatchDeviceNoAlloc(int handleIdx, const void* frames, int batchCount, int frameIsUint16,
                                           complex64* fieldsOut, complex64* coefsOut) {
  Handle* hp = get(handleIdx);
  if (!hp) return DIGHOLO_ERROR_INVALIDHANDLE;
  Handle& h = *hp;
  if (!frames || !fieldsOut) return DIGHOLO_ERROR_NULLPOINTER;
  if (batchCount < 1 || batchCount > dh::kMaxBatchFrames) return DIGHOLO_ERROR_INVALIDARGUMENT;
  if (!h.devInit || !h.tablesValid) {
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc needs the handle's tables: process one batch with this "
                    "configuration first\n");
    return DIGHOLO_ERROR_NULLPOINTER;
  }
  const Derived& d = h.tabKey;
  dh::Geometry g = d.g;
  if (g.L != 1 || g.fw != g.w || g.fh != g.h || h.cfg.basisType != DIGHOLO_BASISTYPE_HG) {
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc covers one wavelength, IFFTResolutionMode 1 and the HG "
                    "basis\n");
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  if (g.G > 0 && !coefsOut) return DIGHOLO_ERROR_NULLPOINTER;
  if (!is_device_pointer(frames) || !is_device_pointer(fieldsOut) || (coefsOut && !is_device_pointer(coefsOut)) ||
      (reinterpret_cast<uintptr_t>(frames) % 8) != 0) {
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc takes device frames (8-byte aligned) and device outputs\n");
    return DIGHOLO_ERROR_INVALIDARGUMENT;
  }
  // the plans and twiddles the handle's last batch made (looked up, never created)
  auto plan = [&](int n) -> const dh::FFTPlan* {
    auto it = h.ffts.find(n);
    return it == h.ffts.end() ? nullptr : &it->second->plan;
  };
  const dh::FFTPlan *pRow = plan(g.W / 2), *pCol = plan(g.H), *pIx = plan(g.w), *pIy = plan(g.h);
  if (!pRow || !pCol || !pIx || !pIy || h.twWn != g.W) {
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc needs the handle's FFT plans: process one batch with this "
                    "configuration first\n");
    return DIGHOLO_ERROR_NULLPOINTER;
  }
  if (!dh::staged_ip_supported(g, *pRow, *pCol, *pIx, *pIy)) {
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc: transform lengths outside the staged kernels' radices\n");
    return DIGHOLO_ERROR_INVALIDDIMENSION;
  }
  g.B = batchCount;
  g.subB = batchCount;
  // scratch layout: row-spectrum band [B][P][H][w], window [B][P][h][w], peaks [B][P]
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t BP = (size_t)g.B * g.P;
  const size_t bandBytes = up(BP * g.H * g.w * sizeof(float2));
  const size_t winBytes = up(BP * g.h * g.w * sizeof(float2));
  const size_t peakBytes = up(BP * sizeof(unsigned));
  if (bandBytes + winBytes + peakBytes > dh::kNoAllocScratchBytes) {
    const size_t per = (size_t)g.P * ((size_t)g.H * g.w + (size_t)g.h * g.w) * sizeof(float2) + g.P * 4;
    fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc: %d frames need %zu bytes of scratch, more than the %zu "
                    "available (at most about %zu frames of this shape per call)\n",
            batchCount, bandBytes + winBytes + peakBytes, dh::kNoAllocScratchBytes, dh::kNoAllocScratchBytes / per);
    return DIGHOLO_ERROR_MEMORYALLOCATION;
  }
  void* base = nullptr;
  DH_TRY(dh::noalloc_scratch(&base), "scratch address");
  char* sb = static_cast<char*>(base);
  float2* band = reinterpret_cast<float2*>(sb);
  float2* win = reinterpret_cast<float2*>(sb + bandBytes);
  unsigned* peak = reinterpret_cast<unsigned*>(sb + bandBytes + winBytes);
  const Config& c = h.cfg;
  dh::BatchCal cal;
  if (c.calEnabled && c.calP > 0 && c.calB > 0) {
    if (h.calDevVersion != c.calVersion || !h.calDev.p) {
      fprintf(stderr, "digHolo: ProcessBatchDeviceNoAlloc: the batch calibration is not on the device yet\n");
      return DIGHOLO_ERROR_NULLPOINTER;
    }
    cal.p = h.calDev.p;
    cal.P = c.calP;
    cal.B = c.calB;
  }
  const bool split = c.precision == DIGHOLO_PRECISION_SPLIT16;
  float2* fields = reinterpret_cast<float2*>(fieldsOut);
  // the column kernel clears the peaks before the row-IFFT kernel raises them
  unsigned* peakW = (g.G > 0 && split) ? peak : nullptr;
  DH_TRY(dh::launch_row_fft_ip(g, *pRow, h.twW.p, frames, frameIsUint16 ? 1 : 0, band, h.stream), "row FFT");
  DH_TRY(dh::launch_col_window_iy(g, *pCol, *pIy, band, h.factor.p, win, cal, h.stream, peakW), "column FFT");
  DH_TRY(dh::launch_ifft_rows_tilt(g, *pIx, win, h.ref.p, fields, h.stream, peakW), "IFFT rows");
  if (g.G > 0)
    DH_TRY(dh::launch_overlap_mfma(g, fields, h.hgx.p, h.hgy.p, reinterpret_cast<float2*>(coefsOut), split ? 1 : 0,
                                   h.stream, h.ovTpw, peakW, h.hgx16.p, h.hgy16.p),
           "overlap");
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C complex64* digHoloProcessBatchDevice(int handleIdx, int* batchCount, int* modeCount, int* polCount) {
  Handle* h = get(handleIdx);
  if (!h || enqueue_batch(*h)) {
    *batchCount = 0;
    *modeCount = 0;
    *polCount = 0;
    return nullptr;
  }
  const dh::Geometry& g = h->cur.g;
  *batchCount = g.B;
  *polCount = g.P;
  *modeCount = h->Mout;
  return reinterpret_cast<complex64*>(h->outCoefs());
}

EXT_C complex64* digHoloGetFieldsDevice(int handleIdx, int* batchCount, int* polCount, int* width, int* height) {
  Handle* h = get(handleIdx);
  if (!h || !h->haveBatch) {
    *batchCount = 0;
    *polCount = 0;
    *width = 0;
    *height = 0;
    return nullptr;
  }
  const dh::Geometry& g = h->cur.g;
  *batchCount = g.B;
  *polCount = g.P;
  *width = g.fw;
  *height = g.fh;
  return reinterpret_cast<complex64*>(h->fields.p);
}

EXT_C int digHoloSynchronize(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!h->devInit) return DIGHOLO_ERROR_SUCCESS;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    log_err(h, "synchronize", e);
    return DIGHOLO_ERROR_ERROR;
  }
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C int digHoloSetStream(int handleIdx, void* stream) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (h->devInit) (void)hipStreamSynchronize(h->stream);
  h->stream = stream ? static_cast<hipStream_t>(stream) : h->ownStream;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C void* digHoloGetStream(int handleIdx) {
  H_OR(nullptr);
  return h->stream;
}

EXT_C int digHoloGetWindowGeometry(int handleIdx, int* out) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!out) return DIGHOLO_ERROR_NULLPOINTER;
  Config c = h->cfg;
  if (c.polLockTilt) {
    c.tiltX[1] = c.tiltX[0];
    c.tiltY[1] = c.tiltY[0];
  }
  Derived d;
  int err;
  if (!derive(c, std::max(c.batchCount, 1), d, err)) return err;
  const dh::Geometry& g = d.g;
  const int vals[12] = {g.W, g.H, g.w, g.h, g.cx[0], g.cy[0], g.cx[1], g.cy[1], d.SX[0], d.SY[0], d.SX[1], d.SY[1]};
  std::memcpy(out, vals, sizeof(vals));
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C int digHoloAutoAlignCalcMetrics(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!h->haveBatch || !h->devInit) return DIGHOLO_ERROR_NULLPOINTER;
  PowerMeterOut o;
  const int e = power_meter(*h, -1, CFG.aaPolIndependence != 0, true, true, true, CFG.aaMulConjTrans != 0, o);
  if (e) return e;
  store_metrics_db(*h, o);
  return DIGHOLO_ERROR_SUCCESS;
}
// digHolo.cpp:12237-12970: FULL runs the parameter estimate then the tweak
// search, ESTIMATE the estimate alone (then a final pass whose goal metric is
// returned), TWEAK the search from the current values.  The FFT window size is
// restored afterwards.
EXT_C float digHoloAutoAlign(int handleIdx) {
  H_OR(0.f);
  const float FMAX = std::numeric_limits<float>::max();
  if (!h->frames) return -FMAX;
  Config& c = CFG;
  const int mode = c.aaMode;
  const int winX = c.fftWindowSizeX, winY = c.fftWindowSizeY;
  HoldGuard hold(*h);
  if (c.aaFourierRadius)   // digHolo.cpp:12274-12289
    c.apertureSize = c.aaTilt ? window_radius_max(c.pixelSize, c.wavelengthCentre) : window_radius_rec(c);
  float steps[AA_PARS];
  bool snapped = false;
  int e = DIGHOLO_ERROR_SUCCESS;
  const bool estimateOnly = mode == DIGHOLO_AUTOALIGNMODE_ESTIMATE || c.maxMG == 0;
  if (mode == DIGHOLO_AUTOALIGNMODE_FULL || estimateOnly) {
    h->aaMetrics.clear();
    e = autoalign_snap(*h, estimateOnly, steps);
    snapped = true;
  }
  c.fftWindowSizeX = winX;
  c.fftWindowSizeY = winY;
  float goal = -FMAX;
  if (!e && (mode == DIGHOLO_AUTOALIGNMODE_FULL || mode == DIGHOLO_AUTOALIGNMODE_TWEAK) && c.maxMG > 0) {
    e = autoalign_tweak(*h, goal, snapped ? steps : nullptr);
  } else if (!e && !h->aaMetrics.empty()) {
    goal = h->aaMetrics[(size_t)c.aaGoalIdx * (h->aaLambdaCount + 1) + h->aaLambdaCount];
  }
  if (e) {
    fprintf(stderr, "digHolo: AutoAlign failed (error %d)\n", e);
    return -FMAX;
  }
  return goal;
}
EXT_C int digHoloAutoAlignGradient(int handleIdx, int holdPlacement, float mdlWeight, float* loss, float* grad) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!loss || !grad) return DIGHOLO_ERROR_NULLPOINTER;
  if (!h->frames) return DIGHOLO_ERROR_NULLPOINTER;
  double L = 0.0, gr[2][AA_PARS] = {};
  const int e = alignment_gradient(*h, holdPlacement != 0, (double)mdlWeight, L, gr);
  if (e) return e;
  *loss = (float)L;
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < AA_PARS; ++k) grad[p * AA_PARS + k] = (float)gr[p][k];
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C float digHoloAutoAlignGradientSearch(int handleIdx, int maxIterations, float mdlWeight) {
  H_OR(0.f);
  const float FMAX = std::numeric_limits<float>::max();
  if (!h->frames || CFG.maxMG <= 0) return -FMAX;
  float goal = -FMAX;
  const int e = autoalign_gradient_search(*h, maxIterations > 0 ? maxIterations : 30, (double)mdlWeight, goal);
  if (e) {
    fprintf(stderr, "digHolo: gradient alignment failed (error %d)\n", e);
    return -FMAX;
  }
  return goal;
}
EXT_C float digHoloAutoAlignGetMetric(int handleIdx, int metricIdx) {
  H_OR(0.f);
  if (metricIdx < 0 || metricIdx >= DIGHOLO_METRIC_COUNT) return 0.f;
  if (h->aaMetrics.empty()) return -std::numeric_limits<float>::max();
  return h->aaMetrics[(size_t)metricIdx * (h->aaLambdaCount + 1) + h->aaLambdaCount];
}
EXT_C float* digHoloAutoAlignGetMetrics(int handleIdx, int metricIdx) {
  H_OR(nullptr);
  if (metricIdx < 0 || metricIdx >= DIGHOLO_METRIC_COUNT || h->aaMetrics.empty()) return nullptr;
  return &h->aaMetrics[(size_t)metricIdx * (h->aaLambdaCount + 1)];
}

// Pipeline steps (digHolo.cpp:15146-15192 wrap the same four stages).  Each is
// synchronous like the reference call.
EXT_C int digHoloProcessFFT(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  int e = enqueue_batch(*h, 1);
  if (!e && hipStreamSynchronize(h->stream) != hipSuccess) e = DIGHOLO_ERROR_ERROR;
  collect_stage_times(*h);
  return e;
}
EXT_C int digHoloProcessIFFT(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  int e = step_ifft(*h);
  if (!e && hipStreamSynchronize(h->stream) != hipSuccess) e = DIGHOLO_ERROR_ERROR;
  return e;
}
EXT_C int digHoloProcessRemoveTilt(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  int e = step_remove_tilt(*h);
  if (!e && hipStreamSynchronize(h->stream) != hipSuccess) e = DIGHOLO_ERROR_ERROR;
  return e;
}
EXT_C int digHoloProcessBasisExtractCoefs(int handleIdx) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  int e = step_extract_coefs(*h);
  if (!e && hipStreamSynchronize(h->stream) != hipSuccess) e = DIGHOLO_ERROR_ERROR;
  return e;
}

EXT_C int digHoloBatchGetSummary(int handleIdx, int planeIdx, int* parameterCount, int* totalCount, int* polCount,
                                 float** parameters, int* pixelCountX, int* pixelCountY, float** xAxis, float** yAxis,
                                 float** totalIntensity) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!parameterCount || !totalCount || !polCount || !parameters || !pixelCountX || !pixelCountY || !xAxis ||
      !yAxis || !totalIntensity)
    return DIGHOLO_ERROR_NULLPOINTER;
  if (!h->haveBatch) return DIGHOLO_ERROR_NULLPOINTER;
  const dh::Geometry& g = h->cur.g;
  *totalCount = g.B + 1;
  *polCount = g.P;
  *parameterCount = DIGHOLO_ANALYSIS_COUNT;
  if (planeIdx != 0 && planeIdx != 1) return DIGHOLO_ERROR_INVALIDDIMENSION;
  const int e = compute_analysis(*h, planeIdx);
  if (e) return e;
  *pixelCountX = planeIdx == 0 ? g.W / 2 + 1 : g.fw;
  *pixelCountY = planeIdx == 0 ? g.H : g.fh;
  *xAxis = h->anaX[planeIdx].data();
  *yAxis = h->anaY[planeIdx].data();
  *parameters = h->analysis[planeIdx].data();
  *totalIntensity = h->anaSum[planeIdx].data();
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C int digHoloGetFields16(int handleIdx, int* batchCount, int* polCount, short** fieldR, short** fieldI,
                             float** fieldScale, float** x, float** y, int* width, int* height) {
  auto zero = [&]() {
    *batchCount = *polCount = *width = *height = 0;
    *fieldR = *fieldI = nullptr;
    *fieldScale = *x = *y = nullptr;
  };
  Handle* h = get(handleIdx);
  if (!h) {
    zero();
    return DIGHOLO_ERROR_INVALIDHANDLE;
  }
  if (!h->haveBatch || compute_fields16(*h) != DIGHOLO_ERROR_SUCCESS) {
    zero();
    return DIGHOLO_ERROR_NULLPOINTER;
  }
  const dh::Geometry& g = h->cur.g;
  const size_t n = (size_t)g.B * g.P * g.fh * g.fw;
  *batchCount = g.B;
  *polCount = g.P;
  *fieldR = h->hF16.p;
  *fieldI = h->hF16.p + n;
  *fieldScale = &h->analysis[1][(size_t)DIGHOLO_ANALYSIS_MAXABS * g.P * (g.B + 1)];
  *x = h->xAxis.data();
  *y = h->yAxis.data();
  *width = g.fw;
  *height = g.fh;
  return DIGHOLO_ERROR_SUCCESS;
}

EXT_C int digHoloConfigSetFusedWindow(int handleIdx, int enabled) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (enabled < 0 || enabled > 3) return DIGHOLO_ERROR_INVALIDARGUMENT;
  h->windowMode = enabled;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigSetFusedOverlap(int handleIdx, int mode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (mode < -1 || mode > 1) return DIGHOLO_ERROR_INVALIDARGUMENT;
  h->fusedOverlap = mode;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloGetWindowPath(int handleIdx) {
  H_OR(-1);
  return h->haveBatch ? h->lastPath : -1;
}
EXT_C int digHoloConfigSetWindowRowTransform(int handleIdx, int mode) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (mode < 0 || mode > 1) return DIGHOLO_ERROR_INVALIDARGUMENT;
  h->rowMode = mode;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloConfigGetWindowRowTransform(int handleIdx) {
  H_OR(0);
  return h->rowMode;
}

EXT_C int digHoloConfigSetStageTiming(int handleIdx, int enabled) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  h->stageTiming = enabled != 0;
  return DIGHOLO_ERROR_SUCCESS;
}
EXT_C int digHoloGetStageTimes(int handleIdx, float* ms) {
  H_OR(DIGHOLO_ERROR_INVALIDHANDLE);
  if (!ms) return DIGHOLO_ERROR_NULLPOINTER;
  collect_stage_times(*h);
  std::memcpy(ms, h->stageMs, sizeof(h->stageMs));
  return DIGHOLO_ERROR_SUCCESS;
}
