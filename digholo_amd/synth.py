"""Seeded off-axis hologram generator for tests and benchmarks.

Each polarisation's frame region holds I = |E_sig + E_ref|^2 where

* E_sig = sum_{m+n<G} c_mn HG_m(x - x_p) HG_n(y - y_p), continuous-normalised
  Hermite-Gaussian modes of waist w (so that the integral of |E_sig|^2 equals
  sum |c_mn|^2), sampled at the camera pixel centres
  x_i = (i - frameWidth/2) * px, y_j = (j - frameHeight/2) * px
  (the camera coordinate convention of digHolo.cpp:7341-7344);
* E_ref = A exp(i k0 (theta_x x + theta_y y) + i (pi D / lambda) |r - r_p|^2),
  a plane wave whose carrier is k0*theta (theta in radians), i.e. exactly the
  carrier that the reference-wave model removes (digHolo.cpp:10420-10485),
  with optional curvature D (dioptre) centred on the beam;
* A = ref_ratio * RMS(E_sig), RMS taken over the signal's effective area
  (sum|E|^2)^2 / sum|E|^4.

Intensities are scaled so the batch maximum maps to ``peak_counts``, then
Gaussian read noise is added and the result rounded and clipped to 12 bits
(uint16 output), or returned unquantised as float64 when ``quantise=False``.

In dual polarisation, pol p is rendered in the x-half p of the sensor.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np


def hermite_functions(nmax: int, u: np.ndarray) -> np.ndarray:
    """psi_n(u), n < nmax, by the stable three-term recurrence (FP64)."""
    u = np.asarray(u, dtype=np.float64)
    out = np.zeros((max(nmax, 1),) + u.shape)
    out[0] = math.pi ** -0.25 * np.exp(-0.5 * u * u)
    if nmax > 1:
        out[1] = math.sqrt(2.0) * u * out[0]
    for n in range(2, nmax):
        out[n] = math.sqrt(2.0 / n) * u * out[n - 1] - math.sqrt((n - 1.0) / n) * out[n - 2]
    return out[:nmax]


def mode_indices(G: int):
    m, n = [], []
    for g in range(G):
        for k in range(g + 1):
            m.append(g - k)
            n.append(k)
    return np.array(m), np.array(n)


@dataclass
class HologramSpec:
    frame_width: int = 320
    frame_height: int = 256
    pol_count: int = 1
    pixel_size: float = 20e-6
    wavelength: float = 1565e-9
    waist: float = 200e-6
    group_count: int = 10
    beam_centre_x: List[float] = field(default_factory=lambda: [0.0, 0.0])
    beam_centre_y: List[float] = field(default_factory=lambda: [0.0, 0.0])
    tilt_x_deg: List[float] = field(default_factory=lambda: [1.5, 1.5])
    tilt_y_deg: List[float] = field(default_factory=lambda: [1.5, 1.5])
    defocus: List[float] = field(default_factory=lambda: [0.0, 0.0])
    ref_ratio: float = 3.0
    peak_counts: float = 4000.0
    noise_counts: float = 2.0
    # power of group g proportional to exp(-g / decay); None = flat
    group_decay: Optional[float] = None
    # Gauss-Legendre points per axis for integrating the intensity over each
    # pixel (100% fill factor); 0 samples the pixel centre
    pixel_quadrature: int = 0
    # reference-wave amplitude envelope exp(-r^2 / w_r^2) about the frame centre
    # (a non-uniform reference, the case a reference calibration corrects); None = flat
    ref_envelope_waist: Optional[float] = None


def random_coefs(spec: HologramSpec, batch: int, rng: np.random.Generator) -> np.ndarray:
    """i.i.d. CN(0,1) coefficients per (frame, pol), normalised to unit power."""
    M = spec.group_count * (spec.group_count + 1) // 2
    c = (rng.standard_normal((batch, spec.pol_count, M)) +
         1j * rng.standard_normal((batch, spec.pol_count, M))) / math.sqrt(2.0)
    if spec.group_decay is not None:
        mi, ni = mode_indices(spec.group_count)
        c = c * np.exp(-(mi + ni) / (2.0 * spec.group_decay))[None, None, :]
    c /= np.linalg.norm(c, axis=-1, keepdims=True)
    return c


def render(spec: HologramSpec, coefs: np.ndarray, rng: Optional[np.random.Generator] = None,
           quantise: bool = True, chunk: int = 64):
    """Render frames [B, frameHeight, frameWidth].

    Returns (frames, scale, ref_amplitude) where ``scale`` converts |E|^2 to
    counts and ``ref_amplitude[p]`` is A for pol p (field units).
    """
    B = coefs.shape[0]
    P = spec.pol_count
    G = spec.group_count
    Fw, Fh = spec.frame_width, spec.frame_height
    px = spec.pixel_size
    k0 = 2.0 * math.pi / spec.wavelength
    mi, ni = mode_indices(G)
    half = Fw // P
    intens = np.zeros((B, Fh, Fw))
    if spec.pixel_quadrature > 0:
        nodes, weights = np.polynomial.legendre.leggauss(spec.pixel_quadrature)
        nodes = nodes / 2.0
        weights = weights / 2.0
    else:
        nodes, weights = np.zeros(1), np.ones(1)
    amps = None
    for ax, wx in zip(nodes, weights):
        for ay, wy in zip(nodes, weights):
            x = (np.arange(Fw) - Fw // 2 + ax) * px
            y = (np.arange(Fh) - Fh // 2 + ay) * px
            sub, amps_sub = _render_points(spec, coefs, x, y, half, chunk, amps)
            intens += wx * wy * sub
            if amps is None:
                amps = amps_sub
    scale = spec.peak_counts / float(intens.max())
    frames = intens * scale
    if not quantise:
        return frames, scale, amps
    if rng is None:
        rng = np.random.default_rng(0)
    frames = frames + spec.noise_counts * rng.standard_normal(frames.shape)
    frames = np.clip(np.rint(frames), 0, 4095).astype(np.uint16)
    return frames, scale, amps


def _render_points(spec, coefs, x, y, half, chunk, amps_in):
    B = coefs.shape[0]
    P = spec.pol_count
    G = spec.group_count
    k0 = 2.0 * math.pi / spec.wavelength
    mi, ni = mode_indices(G)
    intens = np.zeros((B, len(y), len(x)))
    amps = []
    for p in range(P):
        xs = x[p * half:(p + 1) * half]
        ux = math.sqrt(2.0) * (xs - spec.beam_centre_x[p]) / spec.waist
        uy = math.sqrt(2.0) * (y - spec.beam_centre_y[p]) / spec.waist
        hx = hermite_functions(G, ux) * math.sqrt(math.sqrt(2.0) / spec.waist)   # [G, nx]
        hy = hermite_functions(G, uy) * math.sqrt(math.sqrt(2.0) / spec.waist)   # [G, ny]
        thx = spec.tilt_x_deg[p] * math.pi / 180.0
        thy = spec.tilt_y_deg[p] * math.pi / 180.0
        curv = math.pi * spec.defocus[p] / spec.wavelength
        rx = (xs - spec.beam_centre_x[p]) ** 2
        ry = (y - spec.beam_centre_y[p]) ** 2
        carrier = (np.exp(1j * (k0 * thy * y + curv * ry))[:, None] *
                   np.exp(1j * (k0 * thx * xs + curv * rx))[None, :])
        carrier = carrier * reference_envelope(spec, xs, y)
        amp_p = None if amps_in is None else amps_in[p]
        for b0 in range(0, B, chunk):
            b1 = min(B, b0 + chunk)
            C = np.zeros((b1 - b0, G, G), dtype=np.complex128)
            C[:, mi, ni] = coefs[b0:b1, p]
            # E[y, x] = sum_mn c_mn hy[n, y] hx[m, x]
            E = np.matmul(hy.T[None], np.matmul(np.transpose(C, (0, 2, 1)), hx[None]))
            if amp_p is None:
                # mean |E|^2 over the effective area (sum|E|^2)^2/sum|E|^4 is sum|E|^4/sum|E|^2
                pw = np.sum(np.abs(E) ** 2, axis=(1, 2))
                p4 = np.sum(np.abs(E) ** 4, axis=(1, 2))
                rms = np.sqrt(np.mean(p4 / np.maximum(pw, 1e-300)))
                amp_p = spec.ref_ratio * float(rms)
            tot = E + amp_p * carrier[None]
            intens[b0:b1, :, p * half:(p + 1) * half] = np.abs(tot) ** 2
        amps.append(amp_p)
    return intens, amps


def reference_envelope(spec: HologramSpec, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Reference-wave amplitude [len(y), len(x)] (1 without an envelope)."""
    if spec.ref_envelope_waist is None:
        return np.ones((len(y), len(x)))
    wr2 = spec.ref_envelope_waist ** 2
    return np.exp(-(y[:, None] ** 2 + x[None, :] ** 2) / wr2)


def reference_intensity_frame(spec: HologramSpec, peak: float = 4000.0) -> np.ndarray:
    """The reference wave alone on the camera as uint16 counts (peak counts at its
    maximum): the frame a reference-intensity calibration is taken from."""
    Fw, Fh, px = spec.frame_width, spec.frame_height, spec.pixel_size
    x = (np.arange(Fw) - Fw // 2) * px
    y = (np.arange(Fh) - Fh // 2) * px
    I = reference_envelope(spec, x, y) ** 2
    return np.clip(np.rint(I * (peak / I.max())), 0, 65535).astype(np.uint16)


def make_batch(spec: HologramSpec, batch: int, seed: int = 0, quantise: bool = True):
    """Coefficients and frames for one seeded batch: (frames, coefs, scale, amps)."""
    rng = np.random.default_rng(seed)
    coefs = random_coefs(spec, batch, rng)
    frames, scale, amps = render(spec, coefs, rng, quantise=quantise)
    return frames, coefs, scale, amps


def make_sweep_batch(spec: HologramSpec, wavelengths, batch: int, seed: int = 0, order_in: int = 0,
                     quantise: bool = True):
    """Frames of a wavelength sweep: frame b is rendered at wavelengths[l(b)],
    l(b) = b mod L (wavelength the fast axis) or b // (batch / L) (slow).
    Returns (frames, coefs, scales [B], amps [B][pol])."""
    rng = np.random.default_rng(seed)
    coefs = random_coefs(spec, batch, rng)
    L = len(wavelengths)
    frames, scales, amps = [], [], []
    for b in range(batch):
        lam = b % L if not order_in else b // max(1, batch // L)
        s = replace(spec, wavelength=float(wavelengths[lam]))
        f, sc, am = render(s, coefs[b:b + 1], rng, quantise=quantise)
        frames.append(f[0])
        scales.append(sc)
        amps.append(am)
    return np.stack(frames), coefs, np.array(scales), amps


# The configurations the benchmark and parity tests are quoted on.
def spec_small() -> HologramSpec:
    """320x256 single-pol, G=10 (55 modes), waist 200 um, FFT window 128."""
    return HologramSpec(frame_width=320, frame_height=256, pol_count=1, waist=200e-6, group_count=10)


def spec_dual_640() -> HologramSpec:
    """640x512 dual-pol, G=45 (1035 modes), waist 300 um, beams at -/+1.6 mm."""
    return HologramSpec(frame_width=640, frame_height=512, pol_count=2, waist=300e-6, group_count=45,
                        beam_centre_x=[-1.6e-3, 1.6e-3], beam_centre_y=[0.0, 0.0])
