"""CPU restatement of digHolo's batch-processing path (FP64 oracle).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  It is never the thing measured or shipped, and the
product library has no code path that reaches it.

What it restates (all citations are digHolo.cpp unless noted):

* FFT-window placement ............................ :7267-7391
* Fourier/camera axes ............................. :2841-2883, :7726-7731
* IFFT-window bounds (low-res, r2c) ............... :7868-8048
* Fourier-window copy (mask, fill-factor, phase) .. :1808-2122, :8558-8700
* in-phase frame averaging, batch calibration ..... :1831-1840, :1885-1928, :2066-2114, :8639-8672
* transposed uint16 input ......................... :1206-1229
* inverse FFT of the window ....................... :7533-7582, :8717-8735
* reference wave (tilt, defocus, centre) .......... :10201-10504
* tilt removal (V = -conj(E * ref)) ............... :2701-2837
* HG indices and FP64 Hermite basis ............... :1245-1484, :10044-10645
* separable overlap c = sum_y HGy sum_x HGx V ...... :2149-2341, :9348-9467
* power-meter metrics (IL, MDL, diag, SNR) ........ :10714-11129
* wavelength sweeps (axis, ordering, per-lambda) .. :2656-2700, :6577-6680, :7823-7865, :8644-8670, :12172-12191
* reference-wave calibration ...................... :5309-5555, :6729-6850, :8253-8292, :8736-8785
* plane analysis (power, COM, peak, Aeff) ......... :2354-2654, :8990-9188

Numerical contract
------------------
Discrete decisions (window sizes and corners, the circular Fourier mask) are
evaluated in float32 exactly as the reference evaluates them, so that the HIP
path and this oracle pick the same pixels.  Every continuous quantity (DFTs,
phases, Hermite functions, sinc envelope, normalisation) is evaluated in
float64 from the float32 user inputs.  The reference's int16 storage of fields,
reference wave and modes (:2820-2832, :10490-10497, :1444-1471) is *not*
reproduced by default: the output is the exact value those int16 paths
approximate.  ``emulate_int16=True`` adds the int16 field quantisation of
``applyTilt16``; ``emulate_reference=True`` runs the reference's whole int16 tail
(float32 sincos with its :786 slip, int16 reference wave, int16 fields, int16 HG
modes with per-mode scales, float32 accumulation of 16-row blocks in the
overlap) for the error budget between the reference and the exact path.

Pinning: every function here is checked in ``tests/test_oracle.py`` against
closed forms or brute force (direct DFT, Hermite polynomials, discrete
orthonormality, known-answer hologram recovery, Parseval, numpy.linalg SVD).
The reference publishes no golden vectors and cannot be compiled here (its
FFTW/MKL headers are absent), so nothing is pinned against reference output.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

F32 = np.float32
PI_F32 = F32(3.14159265358979323846)          # digHolo.cpp:93 (float pi)
UNIT_ANGLE_F32 = F32(PI_F32 / F32(180.0))     # digHolo.h:192 (pi/180.0f)
PIXEL_QUANTA = 16                              # digHolo.cpp:99
ALIGN_FLOATS = 16                              # digHolo.cpp:117 (ALIGN 64 / sizeof(float))

TILTX, TILTY, DEFOCUS = 0, 1, 2


@dataclass
class Config:
    """Subset of ``digHoloConfig`` (digHolo.cpp:3323-3425) that the path reads.

    Defaults follow the reference constructor, except ``resolution_mode``:
    1 is the low-res IFFT of the window alone, 0 the reference's default
    full-resolution IFFT (the window in a zero W x H plane).
    """
    frame_width: int = 320
    frame_height: int = 256
    pixel_size: float = 20e-6
    pol_count: int = 1
    fft_window_x: int = 0
    fft_window_y: int = 0
    aperture_deg: float = 0.4
    resolution_mode: int = 1
    wavelength: float = 1565e-9
    beam_centre_x: List[float] = field(default_factory=lambda: [0.0, 0.0])
    beam_centre_y: List[float] = field(default_factory=lambda: [0.0, 0.0])
    tilt_x: List[float] = field(default_factory=lambda: [-0.75, -0.75])
    tilt_y: List[float] = field(default_factory=lambda: [0.75, 0.75])
    defocus: List[float] = field(default_factory=lambda: [0.0, 0.0])
    waist: List[float] = field(default_factory=lambda: [510.0, 510.0])
    group_count: int = 0
    fill_factor: int = 1
    pol_lock_tilt: int = 0
    pol_lock_defocus: int = 0
    pol_lock_waist: int = 1
    # wavelength sweep (digHolo.cpp:3361, :6578-6680): None or one entry means
    # the centre wavelength alone; ordering (input, output), 0 fast, 1 slow
    wavelengths: List[float] = None
    wavelength_ordering: Tuple[int, int] = (0, 0)


# ---------------------------------------------------------------------------
# Geometry (float32 mimicry of the reference's discrete decisions)
# ---------------------------------------------------------------------------

def _round_f32(v) -> int:
    """C ``round`` on a float: half away from zero."""
    v = float(v)
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def fft_window(cfg: Config) -> Tuple[int, int, List[int], List[int]]:
    """FFT window size and per-pol corner.  digHolo.cpp:7283-7377."""
    fw = cfg.fft_window_x if cfg.fft_window_x > 0 else cfg.frame_width // cfg.pol_count
    fh = cfg.fft_window_y if cfg.fft_window_y > 0 else cfg.frame_height
    W = PIXEL_QUANTA * (fw // PIXEL_QUANTA)
    H = PIXEL_QUANTA * (fh // PIXEL_QUANTA)
    W = max(W, 16)
    H = max(H, 16)
    W = min(W, cfg.frame_width)
    H = min(H, cfg.frame_height)
    px = F32(cfg.pixel_size)
    cxs, cys = [], []
    for p in range(cfg.pol_count):
        bcx = F32(cfg.beam_centre_x[p])
        bcy = F32(cfg.beam_centre_y[p])
        CX = F32(bcx / px) + F32(cfg.frame_width // 2)
        CY = F32(bcy / px) + F32(cfg.frame_height // 2)
        cx = _round_f32(F32(F32(CX - F32(W // 2)) / F32(ALIGN_FLOATS))) * ALIGN_FLOATS
        cy = _round_f32(F32(CY - F32(H // 2)))
        cx = max(cx, 0)
        cy = max(cy, 0)
        if cx + W > cfg.frame_width:
            cx = cfg.frame_width - W
        if cy + H > cfg.frame_height:
            cy = cfg.frame_height - H
        cxs.append(cx)
        cys.append(cy)
    return W, H, cxs, cys


def k_axis_f32(n: int, pixel_size: float) -> np.ndarray:
    """Fourier axis in FFT order as the reference stores it (float32).

    kFactor = (float)(2*(2*pi/(2*px))/n) (digHolo.cpp:7726-7729);
    k[i] = kFactor*i for i <= n/2 else kFactor*(i-n) (:2843-2867).
    """
    px = F32(pixel_size)
    pixel_period = F32(2.0 * float(px))
    kf = F32(2.0 * (2.0 * math.pi / float(pixel_period)) / n)
    i = np.arange(n)
    idx = np.where(i <= n // 2, i, i - n).astype(np.float32)
    return (kf * idx).astype(np.float32)


def tilt_rad_f32(deg: float) -> np.float32:
    return F32(F32(deg) * UNIT_ANGLE_F32)


def wavelength_axis(cfg: Config) -> np.ndarray:
    """The wavelengths a batch uses (float32): the sweep, or the centre
    wavelength when the sweep has fewer than two entries (digHolo.cpp:7823-7865)."""
    if cfg.wavelengths is not None and len(cfg.wavelengths) > 1:
        return np.asarray(cfg.wavelengths, dtype=np.float32)
    return np.array([cfg.wavelength], dtype=np.float32)


def wavelength_ordering_calc(idx: int, L: int, B: int, ord_in: int, ord_out: int) -> Tuple[int, int, int]:
    """(output position, wavelength index, sub-batch index) of input element idx
    (digHolo.cpp:2656-2700): wavelength the fast axis (0) or the slow one (1)."""
    sub_count = B // L
    if not ord_in:
        lam, sub = idx % L, idx // L
    elif sub_count:
        sub, lam = idx % sub_count, idx // sub_count
    else:
        lam = sub = 0
    out = sub * L + lam if not ord_out else lam * sub_count + sub
    return out, lam, sub


def frequency_sweep_linear(lam_start: float, lam_stop: float, count: int, centre: float) -> np.ndarray:
    """Wavelengths linear in optical frequency, float32 as digHolo.cpp:6577-6650
    evaluates them (c0 = 299792458 converted to float; one step takes the mean of
    valid end points, or the centre wavelength when neither is valid)."""
    c0 = F32(299792458)
    a, b = F32(lam_start), F32(lam_stop)     # DIGHOLO_UNIT_LAMBDA = 1
    f_start = F32(c0 / a) if a != 0 else F32(np.inf)
    f_stop = F32(c0 / b) if b != 0 else F32(np.inf)
    with np.errstate(all="ignore"):
        df = F32(F32(f_stop - f_start) / F32(count - 1))
    if count <= 1:
        if b <= 0 < a or a <= 0 < b:
            f_start = F32(c0 / F32(a / F32(2)))
        elif a <= 0 and b <= 0:
            f_start = F32(c0 / F32(centre))
        else:
            f_start = F32(c0 / F32(F32(b + a) / F32(2)))
        df = F32(0)
    return np.array([F32(c0 / F32(f_start + F32(df * F32(i)))) for i in range(count)], dtype=np.float32)


def _window_bounds(cfg: Config, W: int, H: int, lam_a, lam_b, tilts, r2c: bool):
    """IFFT window size and per-pol corner, digHolo.cpp:7868-8048: the outer edge
    at k0 of the longer end wavelength, the inner edge at k0 of the shorter;
    tilts [(tx_deg, ty_deg)] per pol; r2c keeps the window in the kx >= 0 half."""
    k0max = F32(F32(2.0) * PI_F32 / max(F32(lam_a), F32(lam_b)))
    k0min = F32(F32(2.0) * PI_F32 / min(F32(lam_a), F32(lam_b)))
    kx = k_axis_f32(W, cfg.pixel_size)
    ky = k_axis_f32(H, cfg.pixel_size)
    dkx = F32(kx[1] - kx[0])
    dky = F32(ky[1] - ky[0])
    r0 = F32(F32(cfg.aperture_deg) * UNIT_ANGLE_F32)
    wc, hc = [], []
    w = 0
    h = 0
    for tx_deg, ty_deg in tilts:
        tx = tilt_rad_f32(tx_deg)
        ty = tilt_rad_f32(ty_deg)
        maxEX = F32(float(F32(F32(tx + r0) * k0max) / dkx) + W / 2.0)
        maxEY = F32(float(F32(F32(ty + r0) * k0max) / dky) + H / 2.0)
        minSX = F32(float(F32(F32(tx - r0) * k0min) / dkx) + W / 2.0)
        minSY = F32(float(F32(F32(ty - r0) * k0min) / dky) + H / 2.0)
        wc.append(_round_f32(float(F32(maxEX + minSX)) / 4.0) * 2)
        hc.append(_round_f32(float(F32(maxEY + minSY)) / 4.0) * 2)
        wp = int(F32(maxEX - minSX + F32(1.0)))
        hp = int(F32(maxEY - minSY + F32(1.0)))
        wp = int(16 * math.ceil(float(F32(wp)) / 16))
        hp = int(16 * math.ceil(float(F32(hp)) / 16))
        w = max(w, wp)
        h = max(h, hp)
    w = max(w, 16)
    h = max(h, 16)
    SX, SY = [], []
    for p in range(len(tilts)):
        sx = int(wc[p] - w / 2.0)
        if r2c and sx < W // 2:
            sx = W // 2
        if sx < 0:
            sx = 0
        if sx + w > W:
            sx = W - w
        sy = int(hc[p] - h / 2.0)
        SX.append(sx)
        SY.append(sy)
    return w, h, SX, SY


def ifft_window(cfg: Config, W: int, H: int) -> Tuple[int, int, List[int], List[int]]:
    """IFFT window size and per-pol corner (low-res, r2c).  digHolo.cpp:7868-8048:
    wide enough for the longest and the shortest wavelength of the axis."""
    lam = wavelength_axis(cfg)
    tilts = [(cfg.tilt_x[0 if cfg.pol_lock_tilt else p], cfg.tilt_y[0 if cfg.pol_lock_tilt else p])
             for p in range(cfg.pol_count)]
    return _window_bounds(cfg, W, H, lam[0], lam[-1], tilts, r2c=True)


@dataclass
class Plan:
    W: int
    H: int
    cx: List[int]
    cy: List[int]
    w: int
    h: int
    SX: List[int]
    SY: List[int]


def apply_pol_locks(cfg: Config) -> Config:
    """Tilt/defocus locks copy pol 0's values into pol 1 (digHolo.cpp:10296-10318)."""
    import copy
    c = copy.deepcopy(cfg)
    if c.pol_lock_tilt:
        c.tilt_x[1] = c.tilt_x[0]
        c.tilt_y[1] = c.tilt_y[0]
    if c.pol_lock_defocus:
        c.defocus[1] = c.defocus[0]
    return c


def make_plan(cfg: Config) -> Plan:
    W, H, cx, cy = fft_window(cfg)
    w, h, SX, SY = ifft_window(cfg, W, H)
    return Plan(W, H, cx, cy, w, h, SX, SY)


# ---------------------------------------------------------------------------
# Hermite-Gaussian basis (FP64)
# ---------------------------------------------------------------------------

def hg_indices(G: int) -> Tuple[np.ndarray, np.ndarray]:
    """Mode enumeration: group g, k=0..g -> (m=g-k, n=k).  digHolo.cpp:1245-1262."""
    M, N = [], []
    for g in range(G):
        for k in range(g + 1):
            M.append(g - k)
            N.append(k)
    return np.array(M, dtype=np.int64), np.array(N, dtype=np.int64)


def hermite_functions(nmax: int, u: np.ndarray) -> np.ndarray:
    """Normalised Hermite functions psi_n(u), n < nmax (FP64, three-term recurrence).

    psi_0 = pi^-1/4 exp(-u^2/2); psi_1 = sqrt(2) u psi_0;
    psi_n = sqrt(2/n) u psi_{n-1} - sqrt((n-1)/n) psi_{n-2}   (digHolo.cpp:1327-1376).
    Mantissa/exponent are tracked separately (the reference's trick, :1378-1415)
    so that nothing under- or overflows at high order.
    """
    u = np.asarray(u, dtype=np.float64)
    out = np.zeros((nmax,) + u.shape)
    if nmax == 0:
        return out
    # log-magnitude bookkeeping: psi = mant * 2**expo * exp(-u^2/2)
    g = -0.5 * u * u
    m0 = np.full(u.shape, math.pi ** -0.25)
    e0 = np.zeros(u.shape)
    out[0] = m0 * np.exp(g)
    if nmax == 1:
        return out
    m1 = math.sqrt(2.0) * math.pi ** -0.25 * u
    e1 = np.zeros(u.shape)
    out[1] = m1 * np.exp(g)
    for n in range(2, nmax):
        a = math.sqrt(2.0 / n)
        b = math.sqrt((n - 1.0) / n)
        # bring the older term to the exponent of the newer one
        mn = a * u * m1 - b * m0 * np.exp2(e0 - e1)
        en = e1.copy()
        fm, fe = np.frexp(mn)
        mn = fm
        en = en + fe
        m0, e0 = m1, e1
        m1, e1 = mn, en
        with np.errstate(over="ignore", under="ignore"):
            out[n] = mn * np.exp(g + en * math.log(2.0))
    return out


def hg_modes_1d(nmax: int, x: np.ndarray, waist: float, dx: float) -> np.ndarray:
    """HG mode profile psi_n(sqrt(2) x / w) * sqrt(dx * sqrt(2) / w).  digHolo.cpp:1272-1451."""
    u = math.sqrt(2.0) * np.asarray(x, dtype=np.float64) / waist
    return hermite_functions(nmax, u) * math.sqrt(dx * math.sqrt(2.0) / waist)


# ---------------------------------------------------------------------------
# The processing path
# ---------------------------------------------------------------------------

@dataclass
class PolTables:
    """Per-polarisation quantities that do not depend on the frame."""
    col: np.ndarray        # [w] column of the r2c plane each window column reads
    row: np.ndarray        # [h] row of the r2c plane each window row reads
    factor: np.ndarray     # [h, w] complex mask * sinc * phase * norm
    ref: np.ndarray        # [fh, fw] complex reference wave exp(i(phix+phiy)) on the field plane
    hgx: np.ndarray        # [G, fw]
    hgy: np.ndarray        # [G, fh]
    mask: np.ndarray = None    # [h, w] bool, the circular Fourier window


def _pol_index(cfg: Config, p: int) -> Tuple[int, int]:
    """Which pol's tilt/defocus the reference wave uses (digHolo.cpp:10289-10325).

    With PolLockDefocus the reference overwrites ``polIDX`` with 0 and then
    reads tilt, tilt offset and defocus through it; with PolLockTilt only the
    tilt is copied from pol 0.
    """
    tilt_p = 0 if (cfg.pol_lock_tilt or cfg.pol_lock_defocus) else p
    idx_p = 0 if cfg.pol_lock_defocus else p
    return tilt_p, idx_p


def pol_tables(cfg: Config, plan: Plan, p: int, mask=None, wavelength=None) -> PolTables:
    """Tables of pol p at one wavelength (default the centre): the aperture and the
    reference wave follow k0 = 2 pi / lambda (digHolo.cpp:8644-8656, :10386-10420)."""
    W, H, w, h = plan.W, plan.H, plan.w, plan.h
    px32 = F32(cfg.pixel_size)
    px = float(px32)
    lam32 = F32(cfg.wavelength if wavelength is None else wavelength)
    lam = float(lam32)

    # --- Fourier-window copy, low-res r2c branch (digHolo.cpp:1845-2065) ---
    kxa = k_axis_f32(W, cfg.pixel_size)
    kya = k_axis_f32(H, cfg.pixel_size)
    SX0 = plan.SX[p] - W // 2
    col = SX0 + np.arange(w)
    j = np.arange(h)
    jSY = j + plan.SY[p]
    row = np.mod(jSY + H // 2, H)
    # float32 mask exactly as the reference evaluates it (:1950-2003, :8647-8656).
    # The mask centre uses this pol's own tilt (zernCoefs[polIdx], :8620-8621).
    k0_32 = F32(F32(2.0) * PI_F32 / lam32)
    tx32 = tilt_rad_f32(cfg.tilt_x[p])
    ty32 = tilt_rad_f32(cfg.tilt_y[p])
    r32 = F32(F32(cfg.aperture_deg) * UNIT_ANGLE_F32)
    kx0 = F32(k0_32 * F32(math.sin(float(tx32))))
    ky0 = F32(k0_32 * F32(math.sin(float(ty32))))
    kr = F32(k0_32 * F32(math.sin(float(r32))))
    r02 = F32(kr * kr)
    full_shift = F32(abs(F32(F32(kya[1] - kya[0]) * F32(H))))
    shift = np.where(jSY < 0, full_shift, np.where(jSY > H, -full_shift, F32(0))).astype(np.float32)
    kx32 = kxa[col]
    ky32 = kya[row]
    dkx = (kx32 - kx0).astype(np.float32)
    dky = (ky32 - (ky0 + shift).astype(np.float32)).astype(np.float32)
    dr2 = (dkx * dkx)[None, :] + (dky * dky)[:, None]
    inside = (dr2.astype(np.float32) <= r02)
    if mask is not None:            # a held Fourier window (see Hold)
        inside = np.asarray(mask, dtype=bool)

    # continuous parts in FP64
    dk_x = 2.0 * math.pi / (px * W)
    dk_y = 2.0 * math.pi / (px * H)
    kx = col.astype(np.float64) * dk_x                      # col in [0, W/2)
    ky = np.where(row <= H // 2, row, row - H).astype(np.float64) * dk_y
    cx0 = (plan.cx[p] + W // 2) - cfg.frame_width // 2
    cy0 = (plan.cy[p] + H // 2) - cfg.frame_height // 2
    bcx = float(F32(cfg.beam_centre_x[p])) - cx0 * px
    bcy = float(F32(cfg.beam_centre_y[p])) - cy0 * px
    if cfg.fill_factor:
        def xcsc(t):
            t = np.asarray(t, dtype=np.float64)
            out = np.ones_like(t)
            nz = t != 0
            out[nz] = t[nz] / np.sin(t[nz])
            return out
        sx = xcsc(kx * px / 2.0)
        sy = xcsc(ky * px / 2.0)
    else:
        sx = np.ones(w)
        sy = np.ones(h)
    # field plane: the window (low-res) or the whole FFT window (full resolution),
    # digHolo.cpp:8081-8082; the IFFT length in the norm follows it (:1813-1822)
    low = cfg.resolution_mode == 1
    fw, fh = (w, h) if low else (W, H)
    norm = 1.0 / math.sqrt(float(W) * H * fw * fh)
    fx = sx * np.exp(1j * kx * bcx)
    fy = sy * np.exp(1j * ky * bcy)
    factor = inside * (fy[:, None] * fx[None, :]) * norm

    # --- reference wave (digHolo.cpp:10289-10497), FP64 phase ---
    tp, ip = _pol_index(cfg, p)
    k0 = 2.0 * math.pi / lam
    thx = float(F32(cfg.tilt_x[tp])) * math.pi / 180.0
    thy = float(F32(cfg.tilt_y[tp])) * math.pi / 180.0
    # the window-offset tilt exists only in low-res mode (digHolo.cpp:10326-10334)
    thoff_x = (plan.SX[ip] - W / 2.0) * dk_x if low else 0.0
    thoff_y = (plan.SY[ip] - H / 2.0) * dk_y if low else 0.0
    D = float(F32(cfg.defocus[ip]))
    pxx = px * W / fw
    pxy = px * H / fh
    x = (np.arange(fw) - fw // 2) * pxx
    y = (np.arange(fh) - fh // 2) * pxy
    kxr = -k0 * thx + thoff_x
    kyr = -k0 * thy + thoff_y
    focal = -math.pi * D / lam
    phix = kxr * x + float(F32(cfg.beam_centre_x[p])) * (-k0 * thx) + focal * x * x
    phiy = kyr * y + float(F32(cfg.beam_centre_y[p])) * (-k0 * thy) + focal * y * y
    ref = np.exp(1j * phiy)[:, None] * np.exp(1j * phix)[None, :]

    # --- HG basis (digHolo.cpp:10506-10645) ---
    G = cfg.group_count
    wp = float(F32(cfg.waist[0 if cfg.pol_lock_waist else p]))
    if wp == 0:
        wp = float(F32(400e-6))
    hgx = hg_modes_1d(G, x, wp, pxx)
    hgy = hg_modes_1d(G, y, wp, pxy)
    return PolTables(col=col, row=row, factor=factor, ref=ref, hgx=hgx, hgy=hgy, mask=inside)


# ---------------------------------------------------------------------------
# Reference-wave calibration (digHolo.cpp:5309-5555, :6729-6850, :8253-8292,
# :8736-8785 and the complex branch of the window copy, :1808-2122 with r2c = 0)
# ---------------------------------------------------------------------------

def ref_cal_from_intensity(cal: np.ndarray) -> np.ndarray:
    """Stored form of an intensity calibration (:5423-5480), float32:
    sqrt(1 / (I / max I)), zero where that is not finite."""
    I = np.asarray(cal, dtype=np.uint16).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        scale = F32(1.0) / F32(int(I.max()))
        v = F32(1.0) / (I * scale)
    v = np.where(np.isfinite(v), v, F32(0)).astype(np.float32)
    return np.sqrt(v).astype(np.complex64)


def ref_cal_from_field(cal: np.ndarray) -> np.ndarray:
    """Stored form of a field calibration (:5484-5550), float32: normalised to a
    peak magnitude of 1, then (1 / re, -1 / im), zero where either is not finite."""
    z = np.asarray(cal, dtype=np.complex64)
    re, im = z.real.astype(np.float32), z.imag.astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        mx = F32(np.sqrt(F32((re * re + im * im).max())))
        scale = F32(1.0) / mx
        vr = F32(1.0) / (re * scale)
        vi = F32(-1.0) / (im * scale)
    bad = ~(np.isfinite(vr) & np.isfinite(vi))
    vr[bad] = 0
    vi[bad] = 0
    return (vr + 1j * vi).astype(np.complex64)


def ref_cal_planes(src: np.ndarray, cfg: Config, plan: Plan) -> np.ndarray:
    """Calibration planes [Lc, pol, fh, fw] on the field plane.  Each stored frame
    is cropped to the pol's FFT window, transformed (complex, unnormalised), its
    Fourier window -- placed for zero tilt with its own size over the
    calibration's wavelength range, copied with the field window's size -- is
    weighted with the aperture of the frame's wavelength, the fill-factor
    envelope and 1 / (W H), shifted to zero frequency (low resolution) or kept
    at its own frequencies (full resolution) and inverse transformed."""
    src = np.asarray(src)
    Lc = src.shape[0]
    W, H, w, h = plan.W, plan.H, plan.w, plan.h
    P = cfg.pol_count
    lams = wavelength_axis(cfg)
    L = len(lams)
    lam_a, lam_b = (cfg.wavelength, cfg.wavelength) if Lc == 1 else (lams[0], lams[-1])
    _, _, SXc, SYc = _window_bounds(cfg, W, H, lam_a, lam_b, [(0.0, 0.0)] * P, r2c=False)
    sx, sy = SXc[0], SYc[0]
    low = cfg.resolution_mode == 1
    fw, fh = (w, h) if low else (W, H)
    kxa = k_axis_f32(W, cfg.pixel_size)
    kya = k_axis_f32(H, cfg.pixel_size)
    col = np.mod(np.arange(w) + sx - W // 2, W)
    jSY = np.arange(h) + sy
    row = np.mod(jSY + H // 2, H)
    full_shift = F32(abs(F32(F32(kya[1] - kya[0]) * F32(H))))
    shift = np.where(jSY < 0, full_shift, np.where(jSY > H, -full_shift, F32(0))).astype(np.float32)
    kx = kxa[col]
    ky = kya[row]
    dkx = (kx - F32(0)).astype(np.float32)
    dky = (ky - (F32(0) + shift)).astype(np.float32)
    dr2 = ((dkx * dkx)[None, :] + (dky * dky)[:, None]).astype(np.float32)
    r32 = F32(F32(cfg.aperture_deg) * UNIT_ANGLE_F32)
    px = float(F32(cfg.pixel_size))
    if cfg.fill_factor:
        def xcsc(t):
            out = np.ones_like(t)
            nz = t != 0
            out[nz] = t[nz] / np.sin(t[nz])
            return out
        env = xcsc(ky.astype(np.float64) * px / 2)[:, None] * xcsc(kx.astype(np.float64) * px / 2)[None, :]
    else:
        env = np.ones((h, w))
    norm = 1.0 / (W * H)
    planes = np.zeros((Lc, P, fh, fw), dtype=np.complex128)
    for c in range(Lc):
        out, lam_i, _ = wavelength_ordering_calc(c, L, Lc, *cfg.wavelength_ordering)
        k0 = F32(F32(2.0) * PI_F32 / lams[min(lam_i, L - 1)])
        kr = F32(k0 * F32(math.sin(float(r32))))
        inside = dr2 <= F32(kr * kr)
        wt = inside * env * norm
        for p in range(P):
            crop = src[c, plan.cy[p]:plan.cy[p] + H, plan.cx[p]:plan.cx[p] + W].astype(np.complex128)
            F = np.fft.fft2(crop)
            Y = F[row[:, None], col[None, :]] * wt
            Pl = np.zeros((fh, fw), dtype=np.complex128)
            if low:
                Pl[np.mod(np.arange(h) + h // 2, h)[:, None], np.mod(np.arange(w) + w // 2, w)[None, :]] = Y
            else:
                Pl[np.mod(jSY + H // 2, H)[:, None], np.mod(np.arange(w) + sx + W // 2, W)[None, :]] = Y
            planes[out if 0 <= out < Lc else c, p] = np.fft.ifft2(Pl) * (fw * fh)
    return planes


def fft_plane(window: np.ndarray) -> np.ndarray:
    """Unnormalised forward 2-D R2C DFT (FFTW r2c convention, digHolo.cpp:7502)."""
    return np.fft.rfft2(np.asarray(window, dtype=np.float64))


def transpose_u16(frames: np.ndarray) -> np.ndarray:
    """uint16 frames given transposed (convert_int16tofloat32, digHolo.cpp:1206-1229):
    the float frame at flat index i*H + j is the input at flat index j*W + i."""
    frames = np.asarray(frames)
    N, H, W = frames.shape
    return np.ascontiguousarray(frames.reshape(N, H, W).transpose(0, 2, 1)).reshape(N, H, W)


def average_frame_index(b: int, a: int, B: int, A: int, mode: int) -> int:
    """Frame read for batch element b, average slot a (digHolo.cpp:8661-8672, one
    wavelength): sequential b*A + a, interlaced b + a*B, sequential sweep b + a."""
    if mode == 0:
        return b * A + a
    if mode == 1:
        return b + a * B
    return b + a


def average_windows(o: np.ndarray) -> np.ndarray:
    """In-phase average of the A Fourier windows o [A, h, w] of one output
    (digCopyWindow's two passes, digHolo.cpp:1885-1928, :2066-2114):
    W = o_0 / A, then for a >= 1, u = phi / |phi| with phi = sum W conj(o_a),
    W += o_a u / A."""
    A = o.shape[0]
    W = o[0] / A
    for a in range(1, A):
        phi = np.sum(W * np.conj(o[a]))
        u = phi / abs(phi) if abs(phi) > 0 else 1.0
        W = W + o[a] * u / A
    return W


@dataclass
class Hold:
    """Window placement held fixed: FFT crops, Fourier-window position and size
    (the plan) and the circular Fourier mask of each pol, as AutoAlign holds them
    between re-placements (digHolo.cpp:12620-12645).  With them held the path is
    a smooth function of centre, tilt, defocus and waist."""
    plan: Plan
    masks: List[np.ndarray]


def hold_of(cfg: Config) -> Hold:
    c = apply_pol_locks(cfg)
    plan = make_plan(c)
    return Hold(plan, [pol_tables(c, plan, p).mask for p in range(c.pol_count)])


def process_batch(frames: np.ndarray, cfg: Config, emulate_int16: bool = False, avg_count: int = 1,
                  avg_mode: int = 0, batch_cal=None, transpose: bool = False, hold: Hold = None,
                  ref_cal=None, emulate_reference: bool = False):
    """Full ``digHoloProcessBatch`` path.

    frames: [B * avg_count, frameH, frameW] (uint16 or float).  Returns (fields,
    coefs, plan) with fields [B, pol, h, w] complex128 and coefs [B, pol, M]
    complex128.  batch_cal [calP, calB] complex multiplies the Fourier window of
    batch element b, pol p by batch_cal[p % calP, b % calB] (digHolo.cpp:1831-1840,
    :8639-8642); transpose applies the uint16 transposed-input conversion.  ref_cal
    [Lc, frameH, frameW] complex is a stored reference-wave calibration
    (ref_cal_from_*): the inverse-transformed field of wavelength l is multiplied
    by its plane l mod Lc before the tilt is removed (digHolo.cpp:8736-8758).
    emulate_reference replaces the tilt removal and the overlap by the reference's
    int16 forms (ref_wave_int16, tilt_int16, hg_int16, overlap16; low-res IFFT, one
    wavelength, no averaging or calibration): fields are the int16 values times the
    frame's float32 scale as digHoloGetFields returns them (:4950-4976).
    """
    frames = np.asarray(frames)
    if transpose:
        frames = transpose_u16(frames)
    A = max(1, int(avg_count))
    B = frames.shape[0] // A
    cfg = apply_pol_locks(cfg)
    plan = hold.plan if hold is not None else make_plan(cfg)
    P = cfg.pol_count
    G = cfg.group_count
    Mi, Ni = hg_indices(G)
    lams = wavelength_axis(cfg)
    L = len(lams)
    if L > 1 and (A > 1 or hold is not None):
        raise ValueError("a wavelength sweep is restated without averaging or a held window")
    # frame b uses wavelength lam_b and lands at output position out_b
    # (digHolo.cpp:8644-8670, applyTilt :9329-9340 in output order)
    pos = [wavelength_ordering_calc(b, L, B, *cfg.wavelength_ordering) for b in range(B)]
    out_b = np.array([q[0] for q in pos])
    lam_b = np.array([q[1] for q in pos])
    tabs_l = [[pol_tables(cfg, plan, p, mask=hold.masks[p] if hold is not None else None, wavelength=lams[l])
               for p in range(P)] for l in range(L)]
    tabs = tabs_l[0]
    cal_planes = None
    if ref_cal is not None and np.asarray(ref_cal).shape[1:] == (cfg.frame_height, cfg.frame_width):
        cal_planes = ref_cal_planes(ref_cal, cfg, plan)
    low = cfg.resolution_mode == 1
    fh, fw = (plan.h, plan.w) if low else (plan.H, plan.W)
    fields = np.zeros((B, P, fh, fw), dtype=np.complex128)
    coefs = np.zeros((B, P, len(Mi)), dtype=np.complex128)
    for p in range(P):
        t = tabs[p]
        win = frames[:, plan.cy[p]:plan.cy[p] + plan.H, plan.cx[p]:plan.cx[p] + plan.W]
        F = np.fft.rfft2(win.astype(np.float64), axes=(-2, -1))
        factor = np.stack([tabs_l[l][p].factor for l in range(L)])      # the window itself is the same
        ref = np.stack([tabs_l[l][p].ref for l in range(L)])
        if cal_planes is not None:
            ref = ref * np.stack([cal_planes[l % cal_planes.shape[0], p] for l in range(L)])
        Fw = F[:B * A, t.row[:, None], t.col[None, :]]
        Fw = Fw * (factor[lam_b] if A == 1 else factor[0][None])
        if batch_cal is not None:
            cal = np.asarray(batch_cal, dtype=np.complex128)
            cal_b = lambda b: cal[p % cal.shape[0], b % cal.shape[1]]
        else:
            cal_b = lambda b: 1.0
        if A > 1:
            Fw = np.stack([average_windows(np.stack([Fw[average_frame_index(b, a, B, A, avg_mode)] * cal_b(b)
                                                     for a in range(A)])) for b in range(B)])
        elif batch_cal is not None:
            Fw = Fw * np.array([cal_b(b) for b in range(B)])[:, None, None]
        if not low:
            # full resolution: the window keeps its own frequencies in a zero
            # W x H plane (digHolo.cpp:1855-1872, :1938-1950, :2042-2065)
            full = np.zeros((B, plan.H, plan.W), dtype=np.complex128)
            full[:, t.row[:, None], t.col[None, :]] = Fw
            Fw = full
        E = np.fft.ifft2(Fw, axes=(-2, -1)) * (fw * fh)              # unnormalised e^{+i}
        if emulate_reference:
            if L > 1 or A > 1 or batch_cal is not None or cal_planes is not None or not low:
                raise ValueError("the reference emulation covers one wavelength, low-res, no averaging "
                                 "or calibration")
            rx, ry = ref_wave_int16(cfg, plan, p)
            if G > 0:
                qx, sx = hg_int16(t.hgx)
                qy, sy = hg_int16(t.hgy)
            for b in range(B):
                Eb = E[b].astype(np.complex64)
                # the frame's scale from the field analysis of E (:2649, :9174)
                p2 = (Eb.real * Eb.real + Eb.imag * Eb.imag).astype(np.float32)
                mx = np.sqrt(np.max(p2)).astype(np.float32)
                maxabs = F32(32767.0 / float(mx))
                q = tilt_int16(Eb, rx, ry, maxabs)
                fscale = F32(F32(1.0) / maxabs)
                fields[out_b[b], p] = (q.real.astype(np.float32) * fscale).astype(np.float64) + \
                    1j * (q.imag.astype(np.float32) * fscale).astype(np.float64)
                if G > 0:
                    coefs[out_b[b], p] = overlap16(q, fscale, qx, sx, qy, sy, G)
            continue
        V = -np.conj(E * ref[lam_b])
        if emulate_int16:
            V = quantise_int16(V)
        fields[out_b, p] = V
        if G > 0:
            T = np.einsum("mx,byx->bmy", t.hgx, V)       # stage 1: x contraction
            C = np.einsum("ny,bmy->bmn", t.hgy, T)       # stage 2: y contraction
            coefs[out_b, p] = C[:, Mi, Ni]
    return fields, coefs, plan


# ---------------------------------------------------------------------------
# Reference emulation: the reference's float32 sincos, its int16 reference
# wave, int16 fields and int16 HG modes with per-mode scales, and its overlap's
# integer products with float32 accumulation over 16-row blocks.  Test
# infrastructure for the error budget between the reference and the exact path
# (low-res IFFT, one wavelength).
# ---------------------------------------------------------------------------

def _fma32(a, b, c):
    """float32 fused multiply-add (the product of two float32 values is exact in float64)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def sincos_f32(x0) -> Tuple[np.ndarray, np.ndarray]:
    """``sincos_ps`` (digHolo.cpp:716-812) in float32: the quadrant y = rint(|x| 2/pi), a
    three-part Cody-Waite reduction onto [-pi/4, pi/4] with Cephes' split of pi/2,
    Taylor polynomials to x^9 (sin) and x^8 (cos) in Horner form, then the quadrant's
    swaps and signs.  The cos polynomial's second step multiplies by x where x^2 is
    meant (:786), which this restatement keeps: cos carries an error up to ~6e-6."""
    x0 = np.asarray(x0, dtype=np.float32)
    xa = np.abs(x0)
    y = np.rint((xa * F32(2.0 / math.pi)).astype(np.float32)).astype(np.float32)
    q = y.astype(np.int64)
    x = _fma32(y, F32(-0.78515625 * 2.0), xa)
    x = _fma32(y, F32(-2.4187564849853515625e-4 * 2.0), x)
    x = _fma32(y, F32(-3.77489497744594108e-8 * 2.0), x)
    x2 = (x * x).astype(np.float32)
    x3 = (x2 * x).astype(np.float32)
    fact = [math.factorial(n) for n in range(10)]
    s = (F32(1.0 / fact[9]) * x2).astype(np.float32)
    s = _fma32(s, x2, F32(-1.0 / fact[7]))
    s = _fma32(s, x2, F32(1.0 / fact[5]))
    s = _fma32(s, x2, F32(-1.0 / fact[3]))
    s = _fma32(s, x3, x)
    c = (F32(1.0 / fact[8]) * x2).astype(np.float32)
    c = _fma32(c, x, F32(-1.0 / fact[6]))          # x, as at :786
    c = _fma32(c, x2, F32(1.0 / fact[4]))
    c = _fma32(c, x2, F32(-0.5))
    c = _fma32(c, x2, F32(1.0))
    odd = (q & 1) == 1
    s, c = np.where(odd, c, s), np.where(odd, s, c)
    s = np.where(x0 < 0, -s, s)
    s = np.where(((q >> 1) & 1) == 1, -s, s)
    c = np.where((((q + 1) >> 1) & 1) == 1, -c, c)
    return s.astype(np.float32), c.astype(np.float32)


def ref_wave_int16(cfg: Config, plan: Plan, p: int):
    """The reference wave as the reference stores it (digHolo.cpp:10289-10497): separable
    x and y factors exp(i(k r + piston + focal r^2)) evaluated in float32 on the field-plane
    axes, each component times 16383 rounded to int16.  Returns (rx, ry) complex arrays
    of the integer values (real and imaginary parts are the stored int16s)."""
    if cfg.resolution_mode != 1:
        raise ValueError("the reference emulation covers the low-res IFFT")
    W, H = plan.W, plan.H
    tp, ip = _pol_index(cfg, p)
    lam = float(F32(cfg.wavelength))
    k0 = F32(2.0 * math.pi / lam)
    thx = tilt_rad_f32(cfg.tilt_x[tp])
    thy = tilt_rad_f32(cfg.tilt_y[tp])
    kxa = k_axis_f32(W, cfg.pixel_size)
    kya = k_axis_f32(H, cfg.pixel_size)
    dkx = F32(kxa[1] - kxa[0])
    dky = F32(kya[1] - kya[0])
    offx = F32((plan.SX[ip] - W / 2.0) * float(dkx))
    offy = F32((plan.SY[ip] - H / 2.0) * float(dky))
    D = F32(cfg.defocus[ip])
    kx = F32(F32(-k0 * thx) + offx)
    ky = F32(F32(-k0 * thy) + offy)
    kx0 = kx if offx == 0 else F32(k0 * F32(-thx))
    ky0 = ky if offy == 0 else F32(k0 * F32(-thy))
    focal = F32(-math.pi * float(D) / lam)
    x, y = field_axes(cfg, plan)
    out = []
    for axis, k, k00, centre in ((x, kx, kx0, cfg.beam_centre_x[p]), (y, ky, ky0, cfg.beam_centre_y[p])):
        piston = F32(F32(centre) * k00)
        r = axis.astype(np.float32)
        arg = (_fma32(k, r, piston) + (focal * (r * r).astype(np.float32)).astype(np.float32)).astype(np.float32)
        sn, cs = sincos_f32(arg)
        qr = np.clip(np.rint((F32(16383.0) * cs).astype(np.float32)), -32768, 32767)
        qi = np.clip(np.rint((F32(16383.0) * sn).astype(np.float32)), -32768, 32767)
        out.append(qr.astype(np.int64) + 1j * qi.astype(np.int64))
    return out[0], out[1]


def hg_int16(modes: np.ndarray):
    """int16 HG modes with a per-mode scale (digHolo.cpp:1444-1471): scale factor
    32767 / max|mode| in float64, the float32 mode times its float32 value rounded to
    int16; the stored scale is float32(1 / factor).  modes [G, n] -> (q [G, n] int64,
    scales [G] float32)."""
    modes = np.asarray(modes, dtype=np.float64)
    q = np.zeros(modes.shape, dtype=np.int64)
    sc = np.zeros(modes.shape[0], dtype=np.float32)
    for m in range(modes.shape[0]):
        mx = float(np.max(np.abs(modes[m])))
        f = 32767.0 / mx
        v = (F32(f) * modes[m].astype(np.float32)).astype(np.float32)
        q[m] = np.clip(np.rint(v), -32768, 32767).astype(np.int64)
        sc[m] = F32(1.0 / f)
    return q, sc


def tilt_int16(E: np.ndarray, rx: np.ndarray, ry: np.ndarray, scale) -> np.ndarray:
    """``applyTilt16`` (digHolo.cpp:2701-2837) for one (frame, pol): the reference wave from
    its int16 factors (integer products, times float32 1/16383^2), V = -conj(E R) in
    float32, then V times the frame's scale rounded to int16.  Returns the int16 field
    as a complex array of integers."""
    E = np.asarray(E).astype(np.complex64)
    rr = (ry.real[:, None] * rx.real[None, :] - ry.imag[:, None] * rx.imag[None, :]).astype(np.float32)
    ri = (ry.real[:, None] * rx.imag[None, :] + ry.imag[:, None] * rx.real[None, :]).astype(np.float32)
    s = F32(1.0 / (16383.0 * 16383.0))
    ar = (s * rr).astype(np.float32)
    ai = (s * ri).astype(np.float32)
    Er, Ei = E.real.astype(np.float32), E.imag.astype(np.float32)
    vr = ((Ei * ai).astype(np.float32) - (Er * ar).astype(np.float32)).astype(np.float32)
    vi = ((Er * ai).astype(np.float32) + (Ei * ar).astype(np.float32)).astype(np.float32)
    scale = F32(scale)
    qr = np.clip(np.rint((scale * vr).astype(np.float32)), -32768, 32767)
    qi = np.clip(np.rint((scale * vi).astype(np.float32)), -32768, 32767)
    return qr.astype(np.int64) + 1j * qi.astype(np.int64)


def overlap16(f16: np.ndarray, fscale, qx: np.ndarray, sx: np.ndarray, qy: np.ndarray, sy: np.ndarray,
              G: int) -> np.ndarray:
    """``overlapFieldSeparable16`` (digHolo.cpp:2149-2341) for one (frame, pol): exact
    integer products of the int16 field and x-modes summed over x (T[y][m]), then per
    block of 16 rows the block's T times the int16 y-mode summed in float64, rounded to
    float32 and accumulated in float32 over the blocks; at the last block the float32
    product of the three scales multiplies the sum.  When the field has exactly one
    block (16 rows) that last-block step never runs (:2313-2329) and the coefficients
    stay unscaled, as there.  Returns [M] complex64 in mode order."""
    fh = f16.shape[0]
    Mi, Ni = hg_indices(G)
    T = np.einsum("mx,yx->my", qx.astype(np.float64), f16.real.astype(np.float64)), \
        np.einsum("mx,yx->my", qx.astype(np.float64), f16.imag.astype(np.float64))
    nb = fh // 16
    out = np.zeros(len(Mi), dtype=np.complex64)
    for j, (m, n) in enumerate(zip(Mi, Ni)):
        acc = [F32(0.0), F32(0.0)]
        for part in range(2):
            for blk in range(nb):
                ys = slice(16 * blk, 16 * blk + 16)
                ov = F32(float(np.dot(T[part][m, ys], qy[n, ys].astype(np.float64))))
                acc[part] = ov if blk == 0 else F32(acc[part] + ov)
            if nb > 1:
                sc = F32(float(sx[m]) * float(sy[n]) * float(F32(fscale)))
                acc[part] = F32(acc[part] * sc)
        out[j] = np.complex64(complex(float(acc[0]), float(acc[1])))
    return out


def quantise_int16(V: np.ndarray) -> np.ndarray:
    """Per-frame int16 storage of ``applyTilt16`` (digHolo.cpp:2820-2832, :9174, :4954-4976)."""
    out = np.empty_like(V)
    for b in range(V.shape[0]):
        mx = np.max(np.abs(V[b]))
        scale = np.float32(32767.0 / mx) if mx > 0 else np.float32(0)
        qr = np.clip(np.rint(scale * V[b].real), -32768, 32767)
        qi = np.clip(np.rint(scale * V[b].imag), -32768, 32767)
        inv = (1.0 / scale) if scale != 0 else 0.0
        out[b] = (qr + 1j * qi) * inv
    return out


# ---------------------------------------------------------------------------
# Plane analysis (digHolo.cpp:2354-2654, :8990-9188, :6931-6984, :8840-8870):
# power, centre of mass, peak and effective area of every (frame, pol) plane and
# of the batch-summed intensity
# ---------------------------------------------------------------------------

ANALYSIS_TOTALPOWER, ANALYSIS_COMX, ANALYSIS_COMY, ANALYSIS_MAXABS, ANALYSIS_MAXABSIDX, ANALYSIS_AEFF, \
    ANALYSIS_COMYWRAP = range(7)


def plane_analysis(planes: np.ndarray, x: np.ndarray, y: np.ndarray, include=None, skip_rows=(),
                   col_range=None):
    """Statistics of ``digHoloFieldAnalysis`` (digHolo.cpp:2354-2654) finished as in
    ``digHoloFieldAnalysisRoutine`` (:9112-9140).

    planes [B, P, Ny, Nx] complex; x [Nx], y [Ny] float32 axes; include [Ny, Nx]
    bool window test (None: every pixel); rows in ``skip_rows`` and columns
    outside ``col_range`` are not visited at all.  Per visited pixel the power is
    I = |V|^2 / (Nx Ny) (zero outside ``include``).  Returns (params, summed):
    params [7, P, B+1] float64 (index B is the batch-summed intensity map;
    MAXABSIDX holds the integer index 2*j*Nx + i of the peak) and summed
    [P, Ny, Nx], the per-pixel sum of I over the batch.
    """
    planes = np.asarray(planes)
    B, P, Ny, Nx = planes.shape
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    x0, x1 = col_range if col_range is not None else (0, Nx)
    visited = np.zeros((Ny, Nx), dtype=bool)
    visited[:, x0:x1] = True
    for j in skip_rows:
        visited[j, :] = False
    inc = visited if include is None else (visited & np.asarray(include, dtype=bool))
    v2 = np.where(inc, np.abs(planes) ** 2, 0.0)                  # |V|^2, masked
    pw = v2 / (Nx * Ny)
    y_shift = abs(float((F32(y[1]) - F32(y[0])) * F32(Ny)))
    ywrap = y + y_shift * (y < 0)
    dA = float(F32(abs(F32(x[2]) - F32(x[1]))) * F32(abs(F32(y[2]) - F32(y[1]))))
    flat_idx = (2 * np.arange(Ny)[:, None] * Nx + np.arange(Nx)[None, :])
    out = np.zeros((7, P, B + 1))
    summed = pw.sum(axis=0)

    def stats(I, peak_vals):
        tot = I.sum()
        e4 = (I * I).sum()
        pv = np.where(visited, peak_vals, -1.0)
        k = np.flatnonzero(pv.ravel() == pv.max())[-1]        # ties: the later pixel
        return (tot, (I * x[None, :]).sum() / tot, (I * y[:, None]).sum() / tot,
                32767.0 / math.sqrt(pv.ravel()[k]), flat_idx.ravel()[k], dA * tot * tot / e4,
                (I * ywrap[:, None]).sum() / tot)

    for p in range(P):
        for b in range(B):
            out[:, p, b] = stats(pw[b, p], v2[b, p])
        out[:, p, B] = stats(summed[p], summed[p])
    return out, summed


def fourier_plane_analysis(frames: np.ndarray, cfg: Config):
    """Analysis of the R2C Fourier plane of every window (digHolo.cpp:6931-6984):
    the zero-order region k < 2 k0 sin(r) is excluded, as are the kx = 0 column,
    the ky = 0 and Nyquist rows, and the columns past the last multiple of 8."""
    cfg = apply_pol_locks(cfg)
    plan = make_plan(cfg)
    W, H = plan.W, plan.H
    frames = np.asarray(frames)
    planes = np.stack([np.fft.rfft2(frames[:, plan.cy[p]:plan.cy[p] + H, plan.cx[p]:plan.cx[p] + W]
                                    .astype(np.float64), axes=(-2, -1)) for p in range(cfg.pol_count)], axis=1)
    Nx = W // 2 + 1
    kx = k_axis_f32(W, cfg.pixel_size)[:Nx]
    ky = k_axis_f32(H, cfg.pixel_size)
    k0 = F32(F32(2.0) * PI_F32 / F32(cfg.wavelength))
    r0 = F32(F32(cfg.aperture_deg) * UNIT_ANGLE_F32)
    R = F32(F32(F32(2.0) * k0) * F32(math.sin(float(r0))))
    inc = (kx[None, :] * kx[None, :] + ky[:, None] * ky[:, None]).astype(np.float32) >= F32(R * R)
    return plane_analysis(planes, kx, ky, include=inc, skip_rows=(0, H // 2), col_range=(1, 8 * (Nx // 8)))


def field_axes(cfg: Config, plan: Plan):
    """Reconstructed-plane axes (float32): x_i = (i - fw/2) * (px W / fw), digHolo.cpp:8184-8189,
    with fw = w (low-res IFFT) or W (full resolution)."""
    fw, fh = (plan.w, plan.h) if cfg.resolution_mode == 1 else (plan.W, plan.H)
    px = F32(cfg.pixel_size)
    pxx = F32(F32(px * F32(plan.W)) / F32(fw))
    pxy = F32(F32(px * F32(plan.H)) / F32(fh))
    x = (np.arange(fw) - fw // 2).astype(np.float32) * pxx
    y = (np.arange(fh) - fh // 2).astype(np.float32) * pxy
    return x.astype(np.float32), y.astype(np.float32)


def field_plane_analysis(fields: np.ndarray, cfg: Config, plan: Plan):
    """Analysis of the reconstructed fields (digHolo.cpp:8840-8870): every pixel."""
    x, y = field_axes(cfg, plan)
    return plane_analysis(fields, x, y)


def fields_int16(fields: np.ndarray, maxabs: np.ndarray):
    """int16 export of ``digHoloGetFields16`` (digHolo.cpp:4880-4906, :2820-2832):
    round(V * MAXABS) per (frame, pol), MAXABS = 32767 / max|V|."""
    s = np.asarray(maxabs, dtype=np.float32)[..., None, None]
    q = lambda v: np.clip(np.rint((s * v.astype(np.float32)).astype(np.float32)), -32768, 32767).astype(np.int16)
    return q(fields.real), q(fields.imag)


# ---------------------------------------------------------------------------
# Power-meter metrics (digHolo.cpp:10714-11129, :12183-12191)
# ---------------------------------------------------------------------------

METRIC_IL, METRIC_MDL, METRIC_DIAG, METRIC_SNRAVG, METRIC_DIAGBEST, METRIC_DIAGWORST, \
    METRIC_SNRBEST, METRIC_SNRWORST, METRIC_SNRMG = range(9)


def _group_of(j: np.ndarray) -> np.ndarray:
    g = np.floor((np.sqrt(8.0 * j + 1.0) - 1.0) / 2.0).astype(np.int64)
    return g


def power_meter(coefs: np.ndarray, independent: bool = False, mul_conj: bool = False,
                do_il: bool = True, do_mdl: bool = True, do_xtalk: bool = True) -> np.ndarray:
    """Linear metrics of a [B, pol, M] coefficient batch (digHolo.cpp:10714-11129, polIDX = -1).

    A_s is the (modes x batch) matrix the reference hands to cgesvd: column b is
    frame b's coefficient vector (both pols stacked unless the pols are
    independent, in which case there is one matrix per pol).
    """
    c = np.asarray(coefs, dtype=np.complex128)
    B, P, Mp = c.shape
    out = np.zeros(9)
    fmax = float(np.finfo(np.float32).max)
    if B == 1:
        pw = np.abs(c[0]) ** 2
        out[METRIC_IL] = pw.sum()
        out[METRIC_DIAG] = out[METRIC_DIAGBEST] = out[METRIC_DIAGWORST] = pw[:, 0].sum()
        out[METRIC_SNRAVG] = out[METRIC_SNRBEST] = out[METRIC_SNRWORST] = out[METRIC_SNRMG] = fmax
        return out
    indep = independent or P == 1
    if P == 2 and indep:
        mats = [c[:, 0, :], c[:, 1, :]]                      # rows = frames
    else:
        mats = [c.reshape(B, P * Mp)]
    S = len(mats)
    tot_sig = tot_noi = sig_mg = noi_mg = 0.0
    min_sig, max_sig, min_snr, max_snr = fmax, -fmax, fmax, -fmax
    count = 1.0
    if do_xtalk:
        for A in mats:
            X = A
            if mul_conj:
                X = A.conj() @ A.T                          # (A^H A) in the modes x batch sense, N x N
            N, M = X.shape
            count = float(M * N * S)
            for i in range(min(N, Mp)):
                pw = np.abs(X[i]) ** 2
                j = np.arange(M)
                diag = (j == i) | (j - Mp == i)
                sig, noi = pw[diag].sum(), pw[~diag].sum()
                gin = _group_of(np.array([i % Mp]))[0] + 1
                gout = _group_of(j % Mp) + 1
                sig_mg += (pw[gout == gin] / gin).sum()
                noi_mg += (pw[gout != gin] / gin).sum()
                snr = sig / noi
                tot_sig += sig
                tot_noi += noi
                max_sig, min_sig = max(max_sig, sig), min(min_sig, sig)
                max_snr, min_snr = max(max_snr, snr), min(min_snr, snr)
    il = mdl = 0.0
    if do_mdl:
        svals = np.concatenate([np.linalg.svd(A.T, compute_uv=False) for A in mats])
        il = float(np.mean(svals ** 2))
        mdl = float(svals.min() ** 2 / svals.max() ** 2)
    elif do_il:
        il = sum(float((np.abs(A) ** 2).sum()) / B for A in mats)
    out[METRIC_IL] = il
    out[METRIC_MDL] = mdl
    out[METRIC_DIAG] = tot_sig / count
    out[METRIC_SNRAVG] = tot_sig / tot_noi if do_xtalk else 0.0
    out[METRIC_DIAGBEST] = max_sig
    out[METRIC_DIAGWORST] = min_sig
    out[METRIC_SNRBEST] = max_snr
    out[METRIC_SNRWORST] = min_snr
    out[METRIC_SNRMG] = sig_mg / noi_mg if do_xtalk else float("nan")
    return out


def power_meter_sweep(coefs: np.ndarray, L: int, ord_out: int, **kw) -> np.ndarray:
    """Linear metrics per wavelength, [L, 9] (digHolo.cpp:10770-10790): wavelength
    l owns B // L output positions, l, l + L, ... (fast) or l B/L, ... (slow)."""
    c = np.asarray(coefs)
    n = c.shape[0] // L
    rows = [np.arange(n) * L + l if not ord_out else l * n + np.arange(n) for l in range(L)]
    return np.stack([power_meter(c[r], **kw) for r in rows])


def sweep_metrics_db(lin_l: np.ndarray) -> np.ndarray:
    """[9, L + 1] in dB: every wavelength, then their linear average
    (digHolo.cpp:12172-12191)."""
    lin_l = np.asarray(lin_l, dtype=np.float64)
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.concatenate([lin_l.T, lin_l.mean(axis=0)[:, None]], axis=1))


def metrics_db(lin: np.ndarray) -> np.ndarray:
    """10 log10 of the linear metrics (digHolo.cpp:12183-12191), one wavelength."""
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.asarray(lin, dtype=np.float64))


def relative_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / den) if den > 0 else float(np.linalg.norm(a.ravel()))


# ---------------------------------------------------------------------------
# AutoAlign parameter estimate (digHolo.cpp:11517-11930), focus search
# (:11181-11515) and the fitted beam-size relations (:2887-2900, :12040-12075)
# ---------------------------------------------------------------------------

def r2r_matrix(n: int, kind: int) -> np.ndarray:
    """Matrix of FFTW's unnormalised real-to-real transform of length n
    (out = M @ x): kind 10 = REDFT10 (DCT-II), Y_k = 2 sum_j x_j cos(pi (2j+1) k / 2n);
    kind 1 = REDFT01 (DCT-III), Y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / 2n)."""
    k = np.arange(n)[:, None].astype(np.float64)
    j = np.arange(n)[None, :].astype(np.float64)
    if kind == 10:
        return 2.0 * np.cos(np.pi * (2 * j + 1) * k / (2 * n))
    M = 2.0 * np.cos(np.pi * j * (2 * k + 1) / (2 * n))
    M[:, 0] = 1.0
    return M


def r2r_2d(m: np.ndarray, kind: int) -> np.ndarray:
    """2-D transform of an [Ny, Nx] map along both axes (fftwf_plan_r2r_2d, :7393-7420)."""
    Ny, Nx = m.shape
    return r2r_matrix(Ny, kind) @ m @ r2r_matrix(Nx, kind).T


def snap_gauss(m: np.ndarray, px: float, wx2: np.float32, wy2: np.float32) -> np.ndarray:
    """The Gaussian weighting of the DCT-domain map as the reference applies it
    (:11636-11660): its loop counters run over an (8 floor(Nx/8)) x Ny array while
    the map's rows are Nx long, so flat element f < Ny * NxR gets the weight of
    counters j = f // NxR, c = f % NxR; the rest keep their value."""
    Ny, Nx = m.shape
    NxR = 8 * (Nx // 8)
    out = m.ravel().copy()
    f = np.arange(Ny * NxR)
    j = f // NxR
    c = f % NxR
    pxf = F32(px)
    y2 = (j * j).astype(np.float32) * F32(pxf * pxf)
    x = c.astype(np.float32) * pxf
    e = (F32(1) * (x * x) * F32(wx2)).astype(np.float32) + (y2 * F32(wy2)).astype(np.float32)
    out[:Ny * NxR] *= np.exp(e.astype(np.float64))
    return out.reshape(Ny, Nx)


def snap_tilt(summed: np.ndarray, W: int, H: int, pixel_size: float, wavelength: float, ap_rad: np.float32):
    """Tilt (degrees, float32) from one batch-summed R2C Fourier map [H, W/2+1]
    (:11612-11708): DCT-II, Gaussian of the window size, DCT-III, then the first
    maximum outside k^2 > (2 k0 r)^2."""
    k0 = F32(F32(2.0) * PI_F32 / F32(wavelength))
    aperturek = F32(k0 * F32(ap_rad))
    wx = F32((1.0 * float(PI_F32) / float(aperturek)) * math.sqrt(2.0))
    wy = F32((2.0 * float(PI_F32) / float(aperturek)) * math.sqrt(2.0))
    wx2 = F32(-1.0 / float(F32(wx * wx)))
    wy2 = F32(-1.0 / float(F32(wy * wy)))
    m = r2r_2d(np.asarray(summed, dtype=np.float64), 10)
    m = snap_gauss(m, pixel_size, wx2, wy2)
    m = r2r_2d(m, 1)
    Nx = W // 2 + 1
    kx = k_axis_f32(W, pixel_size)[:Nx]
    ky = k_axis_f32(H, pixel_size)
    ad = F32(aperturek * F32(2))
    ad = F32(ad * ad)
    kr = (kx[None, :] * kx[None, :]).astype(np.float32) + (ky[:, None] * ky[:, None]).astype(np.float32)
    cand = np.where(kr > ad, m, -np.inf).ravel()
    if not np.isfinite(cand).any():
        kxs, kys = F32(0), F32(0)
    else:
        f = int(np.argmax(cand))                       # first maximum
        kxs, kys = kx[f % Nx], ky[f // Nx]
    tx = F32(F32(np.arcsin(F32(kxs / k0))) / UNIT_ANGLE_F32)
    ty = F32(F32(np.arcsin(F32(kys / k0))) / UNIT_ANGLE_F32)
    return tx, ty


def aeff_to_mfd(aeff: float, groups: int) -> np.float32:
    """Mode-field diameter from the effective area, the reference's fitted
    relation for a beam filling `groups` HG groups (:2887-2900)."""
    if not groups:
        return F32(0)
    A = F32(float(aeff) / (15.1059697707643 * groups + 6.65360122093789))
    return F32(float(F32(2.0) * F32(np.sqrt(F32(A / PI_F32)))) / 0.223008930683135)


def focus_estimate(w_camera: float, w_fourier: float) -> np.float32:
    """Defocus magnitude from the beam radii in the two planes (:12040-12075)."""
    if not (w_camera > 0 and w_fourier > 0):
        return F32(0)
    ratio = F32(F32(w_fourier) / F32(w_camera))
    a = F32(0.25)
    b = F32(PI_F32 / F32(2))
    with np.errstate(invalid="ignore"):
        r = F32(F32(2) * F32(b * F32(np.sqrt(F32(F32(F32(1) - a) / F32(ratio - a)) - F32(1)))))
    if not (r > 0) or np.isnan(r):
        r = F32(0)
    return r


def quadratic_vertex(x, y) -> float:
    """-b / 2a of the least-squares quadratic a x^2 + b x + c through the points
    (the fixed point of fitToQuadratic's Gauss-Newton iteration, :3146-3257)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if not (np.all(np.isfinite(x)) and np.all(np.isfinite(y))):
        return float("nan")
    A = np.stack([x * x, x, np.ones_like(x)], axis=1)
    (a, b, _), *_ = np.linalg.lstsq(A, y, rcond=None)
    return float(-b / (2 * a)) if a != 0 else float("nan")


def farfield_aeff(fields: np.ndarray, W: int, H: int, pixel_size: float) -> np.ndarray:
    """Effective area of the batch-summed far-field intensity of each pol
    (:11280-11380): the unnormalised inverse 2-D DFT of every field, summed |F|^2
    over the batch, dA (sum I)^2 / sum I^2 with dA the Fourier-window k-step area."""
    F = np.fft.ifft2(fields, axes=(-2, -1)) * (fields.shape[-1] * fields.shape[-2])
    I = (np.abs(F) ** 2).sum(axis=0)                   # [P, fh, fw]
    kx = k_axis_f32(W, pixel_size)
    ky = k_axis_f32(H, pixel_size)
    dA = F32(F32(abs(F32(kx[2] - kx[1]))) * F32(abs(F32(ky[2] - ky[1]))))
    v2 = I.sum(axis=(1, 2))
    v4 = (I * I).sum(axis=(1, 2))
    return np.array([float(F32(dA * F32(F32(v2[p]) * F32(v2[p]) / F32(v4[p])))) for p in range(I.shape[0])])


def _fourier_window_map(spec_p: np.ndarray, W: int, H: int, pixel_size: float, wcx, wcy, R):
    """Batch-summed Fourier intensity of one pol ([B, H, W/2+1] R2C planes) and
    its effective area: |k - wc|^2 >= R^2 (R > 0), <= R^2 (R < 0), all (R = 0);
    kx = 0 column, ky = 0 / Nyquist rows and columns past 8 floor(Nx/8) left out."""
    Nx = W // 2 + 1
    kx = k_axis_f32(W, pixel_size)[:Nx]
    ky = k_axis_f32(H, pixel_size)
    R = F32(R)
    sign = F32(1) if R >= 0 else F32(-1)
    dx = (kx + F32(-F32(wcx))).astype(np.float32)
    dy = (ky + F32(-F32(wcy))).astype(np.float32)
    d2 = ((dx * dx)[None, :] + (dy * dy)[:, None]).astype(np.float32)
    inc = (sign * d2).astype(np.float32) >= F32(sign * F32(R * R))
    params, summed = plane_analysis(spec_p[:, None], kx, ky, include=inc, skip_rows=(0, H // 2),
                                    col_range=(1, 8 * (Nx // 8)))
    return summed[0], float(F32(params[ANALYSIS_AEFF, 0, -1]))


def autoalign_estimate(frames: np.ndarray, cfg: Config, align_centre=True, align_tilt=True,
                       align_defocus=True, align_waist=True, align_radius=True):
    """The parameter estimate of AutoAlign's FULL / ESTIMATE modes
    (digHolo.cpp:12274-12289, :11517-11930, :11181-11515) from the frames alone.

    Returns the config it ends on (tilt, centres, waist, defocus, window radius;
    the FFT window is left at the whole half frame, as during the estimate) and a
    dict of the intermediate estimates."""
    import copy
    c = copy.deepcopy(cfg)
    P = c.pol_count
    px = F32(c.pixel_size)
    lam = F32(c.wavelength)
    info = {}
    if align_radius:                                   # :12274-12289
        if align_tilt:
            wc = F32((-2.0 + math.sqrt(68.0)) / 16.0)
            wmax = F32(float(lam) / (2.0 * float(px)))
            c.aperture_deg = float(F32(F32(wc * wmax) / UNIT_ANGLE_F32))
        else:
            t = min(float(F32(np.sqrt(F32(F32(c.tilt_x[p]) ** 2 + F32(c.tilt_y[p]) ** 2)))) for p in range(P))
            c.aperture_deg = float(F32(F32(t) / F32(3)))
    for p in range(P):
        if align_centre:
            c.beam_centre_x[p] = float(F32((-(P - 1) * (c.frame_width / 4.0) + (p * c.frame_width / 2.0)) * float(px)))
            c.beam_centre_y[p] = 0.0
        if align_tilt:
            c.tilt_x[p] = c.tilt_y[p] = 0.0
        if align_defocus:
            c.defocus[p] = 0.0
    c.fft_window_x = c.frame_width // P
    c.fft_window_y = c.frame_height
    W, H, cx, cy = fft_window(c)
    frames = np.asarray(frames)
    spec = [np.fft.rfft2(frames[:, cy[p]:cy[p] + H, cx[p]:cx[p] + W].astype(np.float64), axes=(-2, -1))
            for p in range(P)]
    k0 = F32(F32(2.0) * PI_F32 / lam)
    ap = F32(F32(c.aperture_deg) * UNIT_ANGLE_F32)
    maps = [None] * P

    def zero_order_maps():
        R = F32(F32(F32(2) * k0) * F32(math.sin(float(ap))))
        for q in range(P):
            maps[q] = _fourier_window_map(spec[q], W, H, c.pixel_size, 0.0, 0.0, R)[0]

    zero_order_maps()
    waists = np.zeros((2, 2), dtype=np.float32)
    tilts = []
    for _ in range(1 + int(bool(align_radius))):
        for p in range(P):
            tx, ty = snap_tilt(maps[p], W, H, c.pixel_size, c.wavelength, ap)
            tilts.append((float(tx), float(ty)))
            if align_radius:
                apmax = F32(F32(np.sqrt(F32(tx * tx + ty * ty))) / F32(3))
                if apmax < F32(c.aperture_deg):
                    c.aperture_deg = float(apmax)
                    ap = F32(apmax * UNIT_ANGLE_F32)
                    zero_order_maps()
            if align_tilt:
                c.tilt_x[p], c.tilt_y[p] = float(tx), float(ty)
            wcx = F32(k0 * F32(F32(c.tilt_x[p]) * UNIT_ANGLE_F32))
            wcy = F32(k0 * F32(F32(c.tilt_y[p]) * UNIT_ANGLE_F32))
            wr = F32(-F32(k0 * F32(math.sin(float(ap)))))
            maps[p], aeff = _fourier_window_map(spec[p], W, H, c.pixel_size, wcx, wcy, wr)
            west = aeff_to_mfd(aeff, c.group_count)
            waists[0, p] = F32((1.0 / float(west)) * float(PI_F32)) if west != 0 else F32(np.inf)
    info["tilts"] = tilts
    # reconstructed plane
    max_focus = F32(0.25)
    def_est = [F32(0), F32(0)]
    used_cx, used_cy = list(c.beam_centre_x), list(c.beam_centre_y)
    valid, attempts = True, 0
    while (not valid or attempts == 0) and attempts < 10:
        valid = True
        used_cx, used_cy = list(c.beam_centre_x), list(c.beam_centre_y)
        fields, _, plan = process_batch(frames, c)
        params, _ = field_plane_analysis(fields, c, plan)
        B = fields.shape[0]
        for p in range(P):
            pol_off = F32(((2 * p - 1) * float(F32(F32(px) * F32(c.frame_width))) / 4.0) * (P - 1))
            comx = F32(F32(params[ANALYSIS_COMX, p, B]) + pol_off)
            comy = F32(params[ANALYSIS_COMY, p, B])
            aeff = F32(params[ANALYSIS_AEFF, p, B])
            west = F32(float(aeff_to_mfd(aeff, c.group_count)) * 0.5)
            waists[1, p] = west
            def_est[p] = focus_estimate(waists[1, p], waists[0, p])
            if np.isnan(comx) or np.isnan(comy) or np.isnan(aeff):
                valid = False
                c.aperture_deg = float(F32(F32(c.aperture_deg) * F32(0.9)))
            if valid:
                if align_centre:
                    c.beam_centre_x[p], c.beam_centre_y[p] = float(comx), float(comy)
                if align_waist:
                    c.waist[p] = float(west)
                if align_defocus:
                    c.defocus[p] = 0.0
                max_focus = max(max_focus, F32(abs(def_est[p])))
            attempts += 1
    info["defocus_estimate"] = [float(v) for v in def_est[:P]]
    info["waists"] = waists.copy()
    if align_defocus:
        best, trials = autofocus(frames, c, [float(v) for v in def_est], used_cx, used_cy)
        info["focus_trials"] = trials
        for p in range(P):
            c.defocus[p] = float(best[p])
            max_focus = max(max_focus, F32(abs(best[p])))
    info["max_focus"] = float(max_focus)
    # the processing that follows writes the locked values into pol 1
    # (reference wave :10296-10318, basis :10527-10531)
    if P > 1:
        if c.pol_lock_tilt:
            c.tilt_x[1], c.tilt_y[1] = c.tilt_x[0], c.tilt_y[0]
        if c.pol_lock_defocus:
            c.defocus[1] = c.defocus[0]
        if c.pol_lock_waist:
            c.waist[1] = c.waist[0]
    return c, info


def autofocus(frames, cfg: Config, scale, centre_x, centre_y):
    """Defocus of least far-field effective area (digHoloAutoFocus, :11181-11515)."""
    import copy
    c = copy.deepcopy(cfg)
    P = c.pol_count
    c.beam_centre_x, c.beam_centre_y = list(centre_x), list(centre_y)
    W, H, _, _ = fft_window(c)
    steps, fit = 25, 5
    scale = [max(F32(s), F32(1.0)) for s in scale[:P]]
    defs = np.zeros((2, steps + 3), dtype=np.float32)
    aeffs = np.zeros((2, steps + 3), dtype=np.float32)
    s0 = list(scale)
    for st in range(0, steps, 2):
        for p in range(P):
            defs[p, st] = s0[p]
            defs[p, st + 1] = -s0[p]
            s0[p] = F32(s0[p] + F32(scale[p] * F32(0.25)))
    best = [F32(0)] * P
    best_aeff = [np.inf] * P
    conv = [False] * P
    trials = []
    done = False
    st = 0
    while st < steps and not done:
        done = True
        for p in range(P):
            c.defocus[p] = float(defs[p, st])
        fields, _, _ = process_batch(frames, c)
        ae = farfield_aeff(fields, W, H, c.pixel_size)
        trials.append([float(defs[p, st]) for p in range(P)])
        for p in range(P):
            aeffs[p, st] = ae[p]
            if aeffs[p, st] < best_aeff[p]:
                best_aeff[p] = aeffs[p, st]
                best[p] = defs[p, st]
            if st == 1:
                if aeffs[p, 1] < aeffs[p, 0]:
                    step_scale, base = F32(defs[p, 1] - defs[p, 0]), defs[p, 1]
                else:
                    step_scale, base = F32(defs[p, 0] - defs[p, 1]), defs[p, 0]
                k = 1
                for i in range(st + 1, steps, 2):
                    defs[p, i] = F32(base + F32(F32(k * F32(0.25)) * step_scale))
                    defs[p, i + 1] = F32(base - F32(F32(k * F32(0.25)) * step_scale))
                    k += 1
            if st >= fit - 1:
                order = sorted(range(st + 1), key=lambda i: aeffs[p, i])     # stable
                sel = order[:fit - 1]
                vx = F32(quadratic_vertex(defs[p, sel], aeffs[p, sel]))
                if st + 1 < steps - 1 and not np.isnan(vx):
                    defs[p, st + 1] = vx
                if abs(F32(vx - best[p])) < F32(0.01):
                    conv[p] = True
            done = done and conv[p]
        st += 1
    return best, trials


# ---------------------------------------------------------------------------
# AutoAlign tweak search (digHolo.cpp:12315-12885): one parameter at a time,
# three samples about the current value, then samples at the vertex of the
# least-squares quadratic through every sample so far, the best kept
# ---------------------------------------------------------------------------

AA_WAIST, AA_DEFOCUS, AA_TILTX, AA_CX, AA_TILTY, AA_CY = range(6)      # digHolo.cpp:2938-2943
_AA_FIELD = {AA_WAIST: "waist", AA_DEFOCUS: "defocus", AA_TILTX: "tilt_x", AA_CX: "beam_centre_x",
             AA_TILTY: "tilt_y", AA_CY: "beam_centre_y"}


def quadratic_vertex(x, y) -> float:
    """``fitToQuadratic`` (digHolo.cpp:3146-3257): Gauss-Newton steps on
    y = a x^2 + b x + c, solved by least squares (sgels) each step.  The model is
    linear in (a, b, c), so the steps reach the least-squares coefficients and
    the returned maximum is -b / (2 a); NaN when a sample is not finite.  Solved
    here in float64 on centred, scaled abscissae (the vertex is invariant under
    that change of variable)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if not (np.all(np.isfinite(x)) and np.all(np.isfinite(y))):
        return float("nan")
    xm = x.mean()
    sc = np.max(np.abs(x - xm))
    if not sc > 0:
        return float("nan")
    u = (x - xm) / sc
    A = np.stack([u * u, u, np.ones_like(u)], axis=1)
    coef, _, rank, _ = np.linalg.lstsq(A, y, rcond=None)
    if rank < 3 or coef[0] == 0:
        return float("nan")
    return float(xm + sc * (-coef[1] / (2 * coef[0])))


def window_radius_rec(cfg: Config) -> np.float32:
    """A third of the smallest tilt magnitude over the pols (digHolo.cpp:11978-11993)."""
    m = min(float(np.sqrt(F32(F32(cfg.tilt_x[p]) * F32(cfg.tilt_x[p]) + F32(cfg.tilt_y[p]) * F32(cfg.tilt_y[p]))))
            for p in range(cfg.pol_count))
    return F32(F32(m) / F32(3.0))


def window_radius_max(pixel_size: float, wavelength: float) -> np.float32:
    """Largest Fourier-window radius any tilt can use (digHolo.cpp:11960-11971): a
    fraction (sqrt(68) - 2) / 16 of the largest resolvable angle lambda / (2 px)."""
    wc = F32((-2.0 + math.sqrt(68.0)) / 16.0)
    wmax = F32(float(F32(wavelength)) / (2.0 * float(F32(pixel_size))))
    return F32(F32(wc * wmax) / UNIT_ANGLE_F32)


def autoalign_tweak(frames: np.ndarray, cfg: Config, tol: float = 0.0, goal: int = METRIC_IL,
                    independent: bool = False, mul_conj: bool = False, align_tilt: bool = True,
                    align_defocus: bool = True, align_centre: bool = True, align_waist: bool = True,
                    fourier_radius: bool = True, max_attempts: int = 20, trace=None):
    """``digHoloAutoAlign`` in TWEAK mode (digHolo.cpp:12276-12290, then the tweak
    stage :12315-12885) from the parameters in cfg.

    Steps (TWEAK mode, :12303-12320): tilt 10% of the aperture, centre and waist
    10% of pol 0's waist, defocus max(0.4, 10% of |defocus|); tried values are
    clamped to the limits of :12336-12359; samples go to the vertex of the
    least-squares quadratic through every sample of the sweep, held within ten
    steps of the best value.  The first evaluation fixes the FFT crops; a tilt
    sweep also holds the Fourier window of the last full evaluation (only the
    reference wave follows the tilt, :12620-12645); after each parameter the
    best values are loaded and the batch reprocessed.  Unlocked parameters of
    two pols are judged with the other pol's best coefficients (:10852-10871).
    A goal other than IL first runs an IL pass, then the goal with steps a tenth
    as large (:12463-12497).  Attempts repeat until the best goal value moves by
    at most tol (0.1 + 10 tol on an IL pre-pass), or 20 attempts.

    Returns (cfg with the final parameters, final metrics in dB [9], attempts
    per pass)."""
    import copy
    c = copy.deepcopy(cfg)
    P = c.pol_count
    FMAX = float(np.finfo(np.float32).max)
    # on entry the window radius opens to the largest usable one when the tilt is
    # searched, else to a third of the tilt (:12276-12290)
    if fourier_radius:
        c.aperture_deg = float(window_radius_max(c.pixel_size, c.wavelength) if align_tilt else window_radius_rec(c))
    px = F32(c.pixel_size)
    lam = F32(c.wavelength)
    base = [0.0] * 6
    base[AA_TILTX] = base[AA_TILTY] = F32(F32(c.aperture_deg) * F32(0.1))
    base[AA_CX] = base[AA_CY] = F32(F32(c.waist[0]) * F32(0.1))
    base[AA_WAIST] = F32(F32(c.waist[0]) * F32(0.1))
    base[AA_DEFOCUS] = F32(0.4)
    for p in range(P):
        base[AA_DEFOCUS] = max(base[AA_DEFOCUS], F32(abs(F32(c.defocus[p])) * F32(0.1)))
    tilt_max = F32(F32(lam / F32(F32(2) * px)) / UNIT_ANGLE_F32)
    lim = {AA_TILTX: (F32(0), tilt_max),
           AA_TILTY: (-(F32(c.aperture_deg) + tilt_max), F32(c.aperture_deg) + tilt_max),
           AA_DEFOCUS: (F32(-100), F32(100)),
           AA_WAIST: (F32(px / F32(3)), F32(px * F32(max(c.frame_height, c.frame_width)))),
           AA_CX: (-F32(F32(c.frame_width / 2.0) * px), F32(F32(c.frame_width / 2.0) * px)),
           AA_CY: (-F32(F32(c.frame_height / 2.0) * px), F32(F32(c.frame_height / 2.0) * px))}
    bound = [F32(b * F32(10)) for b in base]
    get = lambda p, par: F32(getattr(c, _AA_FIELD[par])[p])

    def put(p, par, v):
        getattr(c, _AA_FIELD[par])[p] = float(F32(v))

    def locked(par):
        return ((par in (AA_TILTX, AA_TILTY)) and c.pol_lock_tilt) or (par == AA_DEFOCUS and c.pol_lock_defocus) or \
            (par == AA_WAIST and c.pol_lock_waist)

    best = [[get(p, i) for i in range(6)] for p in range(2 if P > 1 else 1)]
    if P == 1:
        best.append(list(best[0]))
    st = {"crop": None, "win": None, "last": None}

    def process():
        cfg_l = apply_pol_locks(c)
        plan = make_plan(cfg_l)
        if st["crop"] is not None:
            plan.cx, plan.cy = list(st["crop"].cx), list(st["crop"].cy)
        masks = [None] * P
        if st["win"] is not None:
            wp, wm = st["win"]
            plan.w, plan.h, plan.SX, plan.SY = wp.w, wp.h, list(wp.SX), list(wp.SY)
            masks = list(wm)
        _, coefs, plan = process_batch(frames, c, hold=Hold(plan, masks))
        if st["crop"] is None:
            st["crop"] = plan
        st["last"] = (plan, [pol_tables(cfg_l, plan, p, mask=masks[p]).mask for p in range(P)])
        return coefs

    def db(lin):
        return 10.0 * np.log10(np.asarray(lin, dtype=np.float32).astype(np.float64))

    goal_count = 2 if goal else 1
    last_tol = 0.1 + tol * 10.0 if goal else tol
    step = list(base)
    out_best = [[-FMAX] * 9 for _ in range(2)]
    backup = None
    coefs = None
    attempts_log = []
    for gp in range(goal_count):
        g = goal if gp else METRIC_IL
        converged, first = False, True
        best_pwr = [-FMAX, -FMAX]
        last_err = [FMAX, FMAX]
        out_best[0][g] = out_best[1][g] = -FMAX
        attempts = 0
        if gp:
            step = [F32(b / F32(10)) for b in base]
            last_tol = tol
        while not converged:
            enabled = {AA_TILTX: align_tilt, AA_TILTY: align_tilt, AA_CX: align_centre, AA_CY: align_centre,
                       AA_WAIST: align_waist, AA_DEFOCUS: align_defocus}
            for par in range(6):
                if not enabled[par]:
                    continue
                lock = locked(par)
                var0 = [get(0, par), get(0, par) if lock or P == 1 else get(1, par)]
                max_pwr = [-FMAX, -FMAX]
                tilt_par = par in (AA_TILTX, AA_TILTY)
                if tilt_par and st["last"] is not None:
                    st["win"] = st["last"]
                dtx = [[0.0, 0.0] for _ in range(7)]
                xs = [[], []]
                ys = [[], []]
                for k in range(6):
                    valid = [True, True]
                    for p in range(P):
                        if k < 3:
                            dtx[k][p] = F32(var0[p] + ((k - 1.0) * float(step[par]) / 2.0))
                        if lock:
                            dtx[k][1] = dtx[k][0]
                        dtx[k][p] = min(max(F32(dtx[k][p]), lim[par][0]), lim[par][1])
                        valid[p] = bool(np.isfinite(dtx[k][p]))
                        if valid[p]:
                            put(p, par, dtx[k][p])
                    if fourier_radius:
                        c.aperture_deg = float(min(F32(c.aperture_deg), window_radius_rec(c)))
                    coefs = process()
                    if tilt_par and st["win"] is None:
                        st["win"] = st["last"]
                    if first:
                        backup = coefs.copy()
                        first = False
                    if (not lock and P > 1) or independent:
                        outs = []
                        for p in range(P):
                            if independent:
                                lin = power_meter(coefs[:, p:p + 1], False, mul_conj)
                            else:
                                mix = np.stack([coefs[:, p], backup[:, 1 - p]], axis=1)
                                lin = power_meter(mix, False, mul_conj)
                            outs.append(db(lin))
                    else:
                        o = db(power_meter(coefs, independent, mul_conj))
                        outs = [o, o]
                    vx0 = None
                    for p in range(P):
                        pwr = float(F32(outs[p][g]))
                        xs[p].append(float(dtx[k][p]))
                        ys[p].append(pwr)
                        if trace is not None:
                            trace.append(("sample", par, k, p, float(dtx[k][p]), pwr))
                        if pwr > max_pwr[p] and valid[p]:
                            max_pwr[p] = pwr
                            if pwr > best_pwr[p]:
                                best[p] = [get(p, i) for i in range(6)]
                                out_best[p] = [float(F32(v)) for v in outs[p]]
                                backup[:, p] = coefs[:, p]
                                best_pwr[p] = pwr
                        if k >= 2:
                            vx = quadratic_vertex(xs[p], ys[p])
                            if p == 0:
                                vx0 = vx
                            # the reference tests pol 0's vertex for every pol (:12752)
                            if np.isfinite(vx0):
                                d = vx - float(best[p][par])
                                nxt = float(best[p][par]) + (float(bound[par]) if d > 0 else -float(bound[par])) \
                                    if abs(d) > float(bound[par]) else vx
                            else:
                                nxt = float(best[p][par]) + ((k - 2.0) * float(step[par]) / 4.0)
                            dtx[k + 1][p] = F32(nxt)
                st["win"] = None
                for p in range(P):
                    for i in range(6):
                        put(p, i, best[p][i])
                coefs = process()
            if first:
                coefs = process()
                backup = coefs.copy()
                first = False
            backup = coefs.copy()
            err = 0.0
            for p in range(P):
                err += abs(float(out_best[p][g]) - last_err[p])
                last_err[p] = float(out_best[p][g])
            err /= P
            attempts += 1
            if trace is not None:
                trace.append((gp, attempts, [list(map(float, b)) for b in best], [o[g] for o in out_best], err))
            if err <= last_tol or attempts == max_attempts:
                converged = True
        attempts_log.append(attempts)
    st["crop"] = st["win"] = None
    coefs = process()
    final = db(power_meter(coefs, independent, mul_conj))
    return c, final, attempts_log


# ---------------------------------------------------------------------------
# Gradient of the alignment loss (no counterpart in the reference; the FP64
# restatement of the library's adjoint, pinned by central finite differences
# with the window placement held)
# ---------------------------------------------------------------------------

GRAD_WAIST, GRAD_DEFOCUS, GRAD_TILTX, GRAD_CX, GRAD_TILTY, GRAD_CY = range(6)


def hg_modes_1d_dwaist(nmax: int, x: np.ndarray, waist: float, dx: float) -> np.ndarray:
    """d/dw of psi_n(sqrt(2) x / w) sqrt(dx sqrt(2) / w): with u = sqrt(2) x / w,
    psi_n'(u) = sqrt(n/2) psi_{n-1} - sqrt((n+1)/2) psi_{n+1}, du/dw = -u/w and the
    norm contributes -1/(2w)."""
    u = math.sqrt(2.0) * np.asarray(x, dtype=np.float64) / waist
    psi = hermite_functions(nmax + 1, u)
    nrm = math.sqrt(dx * math.sqrt(2.0) / waist)
    out = np.zeros((nmax,) + u.shape)
    for n in range(nmax):
        dpsi = -math.sqrt((n + 1) / 2.0) * psi[n + 1]
        if n > 0:
            dpsi = dpsi + math.sqrt(n / 2.0) * psi[n - 1]
        out[n] = nrm * (dpsi * (-u / waist) - psi[n] / (2.0 * waist))
    return out


def _mdl_matrices(coefs: np.ndarray, independent: bool):
    """The matrices the MDL metric takes singular values of (frames as rows, both
    pols side by side unless independent, digHolo.cpp:10796-11089), with the
    (pol offset) of their columns."""
    c = np.asarray(coefs, dtype=np.complex128)
    B, P, M = c.shape
    if P == 2 and not independent:
        return [(c.reshape(B, 2 * M), 0)]
    return [(c[:, p, :], p) for p in range(P)]


def alignment_loss(coefs: np.ndarray, mdl_weight: float = 0.0, independent: bool = False) -> float:
    """L = -(1/B) sum |c|^2 + mu (ln s_max^2 - ln s_min^2): minus the linear
    insertion loss of the batch without the SVD (digHolo.cpp:11046-11063,
    IL = sum |A|^2 / B), plus a mode-dependent-loss term over the singular values
    of the matrices the MDL metric uses."""
    c = np.asarray(coefs)
    L = -float((np.abs(c) ** 2).sum()) / c.shape[0]
    if mdl_weight:
        sv = np.concatenate([np.linalg.svd(A, compute_uv=False) for A, _ in _mdl_matrices(c, independent)])
        L += mdl_weight * (math.log(sv.max() ** 2) - math.log(sv.min() ** 2))
    return L


def _mdl_seed(coefs: np.ndarray, mdl_weight: float, independent: bool) -> np.ndarray:
    """dL_mdl/dc (Wirtinger): d ln s^2 = 2 Re(u^H dA v) / s -> conj(u_b) v_j / s."""
    c = np.asarray(coefs, dtype=np.complex128)
    B, P, M = c.shape
    out = np.zeros_like(c)
    trip = []
    for A, off in _mdl_matrices(c, independent):
        U, S, Vh = np.linalg.svd(A, full_matrices=False)
        trip.append((S[0], U[:, 0], np.conj(Vh[0]), off, A.shape[1]))
        trip.append((S[-1], U[:, -1], np.conj(Vh[-1]), off, A.shape[1]))
    hi = max(trip, key=lambda t: t[0])
    lo = min(trip, key=lambda t: t[0])
    for (s, u, v, off, n), sign in ((hi, 1.0), (lo, -1.0)):
        gm = sign * mdl_weight * np.outer(np.conj(u), v) / s          # [B, n]
        flat = out.reshape(B, P * M)
        flat[:, off * M:off * M + n] += gm
    return out


def alignment_grad(frames: np.ndarray, cfg: Config, hold: Hold, mdl_weight: float = 0.0,
                   independent: bool = False):
    """Loss and its gradient with respect to [pol][waist, defocus (dioptre),
    tilt x (degree), centre x (m), tilt y (degree), centre y (m)] at the
    configuration, with the placement `hold` fixed.

    With g = dL/dc (Wirtinger, -conj(c)/B) and the adjoint field
    A(x, y) = sum_mn g_mn HGx_m(x) HGy_n(y):
      tilt:    dV/dtheta_x = i k0 (x + bcX) V          -> -2 k0 sum (x + bcX) Im(A V)
      defocus: dV/dD = i (pi/lambda) (x^2 + y^2) V      -> -2 (pi/lambda) sum r^2 Im(A V)
      waist:   2 Re sum_mn g_mn sum (dHGx HGy + HGx dHGy) V
      centre:  the window ramp exp(i kx bcx) shifts the field, dE/dbcx = IFFT(i kx W);
               -2 sum_k kx Im(X1 conj X2) / N with X1 = IFFT(-V R), X2 = IFFT(conj(A) R),
               plus the reference-wave piston -2 k0 theta_x sum Im(A V).
    """
    cl = apply_pol_locks(cfg)
    fields, coefs, plan = process_batch(frames, cfg, hold=hold)
    B, P = fields.shape[:2]
    loss = alignment_loss(coefs, mdl_weight, independent)
    g = -np.conj(coefs) / B
    if mdl_weight:
        g = g + _mdl_seed(coefs, mdl_weight, independent)
    low = cl.resolution_mode == 1
    W, H = plan.W, plan.H
    fw, fh = (plan.w, plan.h) if low else (W, H)
    N = fw * fh
    px = float(F32(cl.pixel_size))
    lam = float(F32(cl.wavelength))
    k0 = 2.0 * math.pi / lam
    dkx, dky = 2.0 * math.pi / (px * W), 2.0 * math.pi / (px * H)
    pxx, pxy = px * W / fw, px * H / fh
    x = (np.arange(fw) - fw // 2) * pxx
    y = (np.arange(fh) - fh // 2) * pxy
    G = cl.group_count
    Mi, Ni = hg_indices(G)
    grad = np.zeros((P, 6))
    for p in range(P):
        t = pol_tables(cl, plan, p, mask=hold.masks[p])
        tp, ip = _pol_index(cl, p)
        wp = 0 if cl.pol_lock_waist else p
        waist = float(F32(cl.waist[wp]))
        dhx = hg_modes_1d_dwaist(G, x, waist, pxx)
        dhy = hg_modes_1d_dwaist(G, y, waist, pxy)
        gm = np.zeros((B, G, G), dtype=np.complex128)          # g as [b, m, n]
        gm[:, Mi, Ni] = g[:, p]
        A = np.einsum("bmn,mx,ny->byx", gm, t.hgx, t.hgy)
        Aw = np.einsum("bmn,mx,ny->byx", gm, dhx, t.hgy) + np.einsum("bmn,mx,ny->byx", gm, t.hgx, dhy)
        V = fields[:, p]
        im = np.imag(A * V)
        S0 = im.sum()
        Sx = (im * x[None, None, :]).sum()
        Sy = (im * y[None, :, None]).sum()
        Sr2 = (im * (x[None, None, :] ** 2 + y[None, :, None] ** 2)).sum()
        Sw = np.real(Aw * V).sum()
        R = t.ref[None]
        X1 = np.fft.ifft2(-V * R, axes=(-2, -1)) * N
        X2 = np.fft.ifft2(np.conj(A) * R, axes=(-2, -1)) * N
        if low:
            kxv = (plan.SX[p] - W // 2 + np.arange(fw)) * dkx
            rows = np.mod(np.arange(fh) + plan.SY[p] + H // 2, H)
            kyv = np.where(rows <= H // 2, rows, rows - H) * dky
        else:
            kxv = np.where(np.arange(W) <= W // 2, np.arange(W), np.arange(W) - W) * dkx
            kyv = np.where(np.arange(H) <= H // 2, np.arange(H), np.arange(H) - H) * dky
        imx = np.imag(X1 * np.conj(X2))
        Tbx = (imx * kxv[None, None, :]).sum() / N
        Tby = (imx * kyv[None, :, None]).sum() / N
        thx = float(F32(cl.tilt_x[tp])) * math.pi / 180.0
        thy = float(F32(cl.tilt_y[tp])) * math.pi / 180.0
        bcx = float(F32(cl.beam_centre_x[p]))
        bcy = float(F32(cl.beam_centre_y[p]))
        deg = math.pi / 180.0
        grad[tp, GRAD_TILTX] += -2.0 * k0 * (Sx + bcx * S0) * deg
        grad[tp, GRAD_TILTY] += -2.0 * k0 * (Sy + bcy * S0) * deg
        grad[ip, GRAD_DEFOCUS] += -2.0 * (math.pi / lam) * Sr2
        grad[p, GRAD_CX] += -2.0 * Tbx - 2.0 * k0 * thx * S0
        grad[p, GRAD_CY] += -2.0 * Tby - 2.0 * k0 * thy * S0
        grad[wp, GRAD_WAIST] += 2.0 * Sw
    return loss, grad


# ---------------------------------------------------------------------------
# Basis transforms (digHolo.cpp:9469-9561, :9819-10030, :4999-5040) and the
# basis fields (:5233-5296)
# ---------------------------------------------------------------------------

def lg_blocks(G: int) -> List[np.ndarray]:
    """Per HG group g (size g+1) the matrix U_g with LG[k] = sum_c U_g[k, c] HG[c]
    (digHoloLGInit :9909-10031 and digHoloLGtoHGcoef :9819-9906): LG mode k of
    group g has the index pair of HG mode k (m = g-k, n = k), l = m - n,
    p = (g - |l|)/2; with (a, b) = (p, p - l) for l < 0 (so a = p, b = -l + p)
    and (a, b) = (l + p, p) otherwise, the entry for HG (hm = g - c, hn = c) is
      i^(g + hm) (-1)^[g odd and l > 0] 2^(-g/2) sqrt(hm! hn! / (a! b!))
        sum_t (-1)^t C(b, t) C(a, hm - t)."""
    from math import comb, factorial
    blocks = []
    for g in range(G):
        U = np.zeros((g + 1, g + 1), dtype=np.complex128)
        for k in range(g + 1):
            m_, n_ = g - k, k
            l = m_ - n_
            p = (g - abs(l)) // 2
            if l < 0:
                a, b = p, -l + p
            else:
                b, a = p, l + p
            for c in range(g + 1):
                hm, hn = g - c, c
                s = sum((-1) ** t * comb(b, t) * comb(a, hm - t) for t in range(0, min(hm, b) + 1) if hm - t <= a)
                val = s * math.sqrt(factorial(hm) * factorial(hn) / (factorial(a) * factorial(b))) * 2.0 ** (-g / 2)
                if g % 2 and l > 0:
                    val = -val
                U[k, c] = val * (1j ** ((g + hm) % 4))
        blocks.append(U)
    return blocks


def lg_transform(coefs: np.ndarray, G: int) -> np.ndarray:
    """HG -> LG coefficients, group by group (HGtoLG :4999-5040)."""
    c = np.asarray(coefs, dtype=np.complex128)
    out = np.zeros_like(c)
    off = 0
    for g, U in enumerate(lg_blocks(G)):
        out[..., off:off + g + 1] = c[..., off:off + g + 1] @ U.T
        off += g + 1
    return out


def custom_transform(coefs: np.ndarray, T: np.ndarray, n_in: int, n_out: int) -> np.ndarray:
    """Custom basis (:9519-9561, :9566-9605): T is the caller's column-major
    n_out x n_in matrix (element (j, i) at T[i * n_out + j]); the part used is
    min(M, n_in) inputs by min(M, n_out) outputs, c_out[j] = sum_i T(j, i) c_HG[i]."""
    c = np.asarray(coefs, dtype=np.complex128)
    M = c.shape[-1]
    di, do = min(M, n_in), min(M, n_out)
    Tm = np.asarray(T, dtype=np.complex128).reshape(n_in, n_out)[:di, :do]    # [i, j]
    return c[..., :di] @ Tm


def basis_field_weights(G: int, basis: int, T=None, n_in: int = 0, n_out: int = 0) -> np.ndarray:
    """Rows of HG weights the basis fields are built from (GetFieldsBasis
    :5233-5296): identity (HG), conj(U_g) rows (LG, the inverse transform of the
    identity), T(j, :) unconjugated (custom)."""
    M = G * (G + 1) // 2
    if basis == 1:
        Wt = np.zeros((M, M), dtype=np.complex128)
        off = 0
        for g, U in enumerate(lg_blocks(G)):
            Wt[off:off + g + 1, off:off + g + 1] = np.conj(U)
            off += g + 1
        return Wt
    if basis == 2:
        di, do = min(M, n_in), min(M, n_out)
        Tm = np.asarray(T, dtype=np.complex128).reshape(n_in, n_out)[:di, :do]
        Wt = np.zeros((do, M), dtype=np.complex128)
        Wt[:, :di] = Tm.T
        return Wt
    return np.eye(M, dtype=np.complex128)


def basis_fields(cfg: Config, weights: np.ndarray) -> np.ndarray:
    """Basis fields [modes, pol, fh, fw]: sum_i w_i HGx_{m_i}(x) HGy_{n_i}(y)
    on each pol's field plane."""
    cfg = apply_pol_locks(cfg)
    plan = make_plan(cfg)
    Mi, Ni = hg_indices(cfg.group_count)
    out = []
    for p in range(cfg.pol_count):
        t = pol_tables(cfg, plan, p)
        out.append(np.einsum("ki,ix,iy->kyx", weights, t.hgx[Mi], t.hgy[Ni]))
    return np.stack(out, axis=1)
