"""Pins the FP64 oracle against closed forms, brute force and known answers.

The reference ships no tests or golden vectors and cannot be compiled here, so
each oracle function is pinned by mathematics chosen so that a plausible
mistake (dropped term, wrong sign, swapped index, wrong normalisation) fails.
The IFFT-window sizes are the ones the reference itself produces for these
windows (320x256/128 -> 48, 640x512/192 -> 64, /256 -> 80, 1024x1024/512 -> 160).
"""
import math

import numpy as np
import pytest

from oracle import digholo_oracle as O
from digholo_amd import synth


def small_cfg(**kw):
    c = O.Config(frame_width=320, frame_height=256, pol_count=1, fft_window_x=128, fft_window_y=128,
                 aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=10,
                 waist=[200e-6, 200e-6], fill_factor=0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def recover(spec, cfg, B=2, seed=1):
    """Fit coefs = alpha * c_true per (pol, frame); return worst residual and alpha/prediction."""
    frames, c, s, A = synth.make_batch(spec, B, seed=seed, quantise=False)
    fields, coefs, plan = O.process_batch(frames, cfg)
    errs, ratios = [], []
    for p in range(cfg.pol_count):
        c0x = plan.SX[p] - plan.W // 2
        c0y = plan.SY[p] - plan.H // 2
        # closed form: V = -(-1)^(c0x+c0y) s A E_sig / px  in the HG-normalised overlap
        pred = -((-1) ** (c0x + c0y)) * s * A[p] / spec.pixel_size
        for b in range(B):
            ct, cr = c[b, p], coefs[b, p]
            alpha = np.vdot(ct, cr) / np.vdot(ct, ct)
            errs.append(np.linalg.norm(cr - alpha * ct) / np.linalg.norm(cr))
            ratios.append(alpha / pred)
    return max(errs), np.array(ratios), plan


def test_rfft2_matches_direct_dft():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 8))
    H, W = x.shape
    ky = np.arange(H)[:, None, None, None]
    kx = np.arange(W // 2 + 1)[None, :, None, None]
    yy = np.arange(H)[None, None, :, None]
    xx = np.arange(W)[None, None, None, :]
    direct = np.sum(x[None, None] * np.exp(-2j * np.pi * (ky * yy / H + kx * xx / W)), axis=(2, 3))
    assert np.allclose(O.fft_plane(x), direct, atol=1e-12)


def test_hermite_functions_closed_form():
    u = np.linspace(-6, 6, 241)
    psi = O.hermite_functions(21, u)
    for n in range(21):
        c = np.zeros(n + 1)
        c[n] = 1.0
        Hn = np.polynomial.hermite.hermval(u, c)
        ref = Hn * np.exp(-u * u / 2) / math.sqrt(2.0 ** n * math.factorial(n) * math.sqrt(math.pi))
        assert np.allclose(psi[n], ref, rtol=1e-10, atol=1e-13), n


def test_hermite_functions_orthonormal_high_order():
    # 80 orders on a fine grid: sum psi_m psi_n du = delta
    u = np.linspace(-20, 20, 8001)
    du = u[1] - u[0]
    psi = O.hermite_functions(80, u)
    gram = psi @ psi.T * du
    assert np.max(np.abs(gram - np.eye(80))) < 1e-9
    # and agrees with the plain recurrence used by the generator
    assert np.allclose(psi, synth.hermite_functions(80, u), rtol=1e-9, atol=1e-300)


def test_hg_indices_order():
    M, N = O.hg_indices(4)
    assert list(zip(M, N)) == [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2),
                               (3, 0), (2, 1), (1, 2), (0, 3)]


@pytest.mark.parametrize("frame,pols,win,G,w_ifft", [
    ((320, 256), 1, 128, 10, 48),
    ((640, 512), 2, 192, 30, 64),
    ((640, 512), 2, 256, 45, 80),
    ((1024, 1024), 2, 512, 80, 160),
])
def test_ifft_window_sizes_match_reference_runs(frame, pols, win, G, w_ifft):
    cfg = O.Config(frame_width=frame[0], frame_height=frame[1], pol_count=pols, fft_window_x=win,
                   fft_window_y=win, aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5],
                   group_count=G, beam_centre_x=[-1.6e-3, 1.6e-3] if pols == 2 else [0, 0])
    plan = O.make_plan(cfg)
    assert (plan.w, plan.h) == (w_ifft, w_ifft)
    assert plan.W == win and plan.H == win
    for p in range(pols):
        # window column range stays inside the r2c half plane (no skipped columns)
        assert plan.SX[p] >= plan.W // 2 and plan.SX[p] + plan.w <= plan.W


def test_fft_window_corner_rounding():
    # corner x snaps to 16 px, y rounds to the nearest pixel, both clamp to the frame
    cfg = small_cfg(beam_centre_x=[37e-6, 0], beam_centre_y=[-55e-6, 0])
    W, H, cx, cy = O.fft_window(cfg)
    assert (W, H) == (128, 128)
    assert cx == [96]          # round((160 + 1.85 - 64)/16)*16
    assert cy == [61]          # round(128 - 2.75 - 64)
    cfg = small_cfg(beam_centre_x=[-5e-3, 0], beam_centre_y=[9e-3, 0])
    W, H, cx, cy = O.fft_window(cfg)
    assert cx == [0] and cy == [256 - 128]


def test_known_answer_single_pol():
    spec = synth.spec_small()
    spec.beam_centre_x = [37e-6, 0]
    spec.beam_centre_y = [-55e-6, 0]
    cfg = small_cfg(beam_centre_x=[37e-6, 0], beam_centre_y=[-55e-6, 0])
    err, ratio, _ = recover(spec, cfg, B=3)
    assert err < 1e-5
    assert np.allclose(ratio, 1.0, atol=1e-6)


def test_known_answer_fill_factor_correction():
    # pixel-integrated (100% fill factor) frames are recovered exactly only with
    # the inverse-sinc envelope on both axes
    spec = synth.spec_small()
    spec.pixel_quadrature = 4
    spec.beam_centre_x = [37e-6, 0]
    cfg = small_cfg(beam_centre_x=[37e-6, 0], fill_factor=1)
    err, ratio, _ = recover(spec, cfg, B=2)
    assert err < 1e-5
    assert np.allclose(ratio, 1.0, atol=1e-5)
    cfg.fill_factor = 0
    err0, ratio0, _ = recover(spec, cfg, B=1)
    assert abs(ratio0[0]) < 0.9                    # uncorrected amplitude droops


def test_known_answer_dual_pol_defocus_wrapped_window():
    spec = synth.spec_dual_640()
    spec.group_count = 12
    spec.defocus = [0.3, -0.2]
    spec.beam_centre_y = [40e-6, -70e-6]
    spec.tilt_x_deg = [1.5, 1.4]
    spec.tilt_y_deg = [1.5, -1.6]
    cfg = O.Config(frame_width=640, frame_height=512, pol_count=2, fft_window_x=256, fft_window_y=256,
                   aperture_deg=0.7, tilt_x=[1.5, 1.4], tilt_y=[1.5, -1.6], defocus=[0.3, -0.2],
                   group_count=12, waist=[300e-6, 300e-6], beam_centre_x=[-1.6e-3, 1.6e-3],
                   beam_centre_y=[40e-6, -70e-6], fill_factor=0)
    err, ratio, plan = recover(spec, cfg, B=2)
    assert plan.SY[1] < 0                          # pol 1 window wraps past ky = -pi/px
    assert err < 1e-5
    assert np.allclose(ratio, 1.0, atol=1e-4)


def test_planted_mistakes_fail_known_answer(monkeypatch):
    spec = synth.spec_small()
    cfg = small_cfg()
    # swapping the x/y mode orders breaks recovery
    real = O.hg_indices
    monkeypatch.setattr(O, "hg_indices", lambda G: real(G)[::-1])
    err, _, _ = recover(spec, cfg, B=1)
    assert err > 0.1
    monkeypatch.setattr(O, "hg_indices", real)
    # using the wrong sign on the window's beam-centre ramp breaks recovery
    spec.beam_centre_x = [150e-6, 0]
    cfg = small_cfg(beam_centre_x=[150e-6, 0])
    base, _, _ = recover(spec, cfg, B=1)
    assert base < 1e-5
    cfg_bad = small_cfg(beam_centre_x=[-150e-6, 0])
    frames, c, s, A = synth.make_batch(spec, 1, seed=1, quantise=False)
    _, coefs, _ = O.process_batch(frames, cfg_bad)
    ct, cr = c[0, 0], coefs[0, 0]
    alpha = np.vdot(ct, cr) / np.vdot(ct, ct)
    assert np.linalg.norm(cr - alpha * ct) / np.linalg.norm(cr) > 0.1


def test_int16_emulation_error_budget():
    # the reference stores fields as int16 with a per-frame scale; that storage
    # alone moves fields by ~1e-5..1e-4 relative L2 on realistic frames
    spec = synth.spec_small()
    frames, c, s, A = synth.make_batch(spec, 4, seed=2, quantise=True)
    cfg = small_cfg(fill_factor=1)
    f0, c0, _ = O.process_batch(frames, cfg)
    f1, c1, _ = O.process_batch(frames, cfg, emulate_int16=True)
    e_f = O.relative_l2(f1, f0)
    e_c = O.relative_l2(c1, c0)
    assert 1e-6 < e_f < 2e-4
    assert e_c < 2e-4


def test_sincos_f32_reference_form():
    # the reference's float32 sincos (digHolo.cpp:716-812): its cos polynomial
    # multiplies by x where x^2 is meant (:786), an error of 1/8! x^8 (1 - x) that
    # reaches ~6.4e-6 at a reduced argument of -pi/4 (odd quadrants put it into sin)
    x = np.linspace(-60.0, 60.0, 240001).astype(np.float32)
    s, c = O.sincos_f32(x)
    es = np.max(np.abs(s - np.sin(x.astype(np.float64))))
    ec = np.max(np.abs(c - np.cos(x.astype(np.float64))))
    assert 5e-6 < es < 7e-6 and 5e-6 < ec < 7e-6, (es, ec)
    # first quadrant: the reduced argument is |x| >= 0, where the slip stays below 8e-7
    xr = np.linspace(0.0, np.pi / 4 * 0.999, 10001).astype(np.float32)
    s, c = O.sincos_f32(xr)
    assert np.max(np.abs(s - np.sin(xr.astype(np.float64)))) < 3e-7
    e1 = np.max(np.abs(c - np.cos(xr.astype(np.float64))))
    assert 3e-7 < e1 < 8e-7, e1
    # odd symmetry of sin, even of cos
    s1, c1 = O.sincos_f32(-x)
    s0, c0 = O.sincos_f32(x)
    assert np.array_equal(s1, -s0) and np.array_equal(c1, c0)


def test_reference_wave_int16():
    # int16 factors of the separable reference wave (digHolo.cpp:10386-10497) against
    # the FP64 wave pinned by the known-answer tests: quantisation (1/(2*16383) per
    # component) and float32 phases, nothing more
    cfg = O.apply_pol_locks(small_cfg(fill_factor=1, defocus=[0.3, 0.3], beam_centre_x=[1e-4, 1e-4]))
    plan = O.make_plan(cfg)
    t = O.pol_tables(cfg, plan, 0)
    rx, ry = O.ref_wave_int16(cfg, plan, 0)
    for f in (rx, ry):
        assert np.all(np.abs(f.real) <= 16383) and np.all(np.abs(f.imag) <= 16383)
        assert np.all(f.real == np.rint(f.real)) and np.all(f.imag == np.rint(f.imag))
    R = ry[:, None] * rx[None, :] / 16383.0 ** 2
    assert np.max(np.abs(R - t.ref)) < 1.5e-4
    assert 5e-6 < O.relative_l2(R, t.ref) < 1e-4


def test_hg_int16_per_mode_scale():
    # digHolo.cpp:1444-1471: each mode spans the int16 range, dequantised within half a step
    cfg = small_cfg()
    plan = O.make_plan(cfg)
    t = O.pol_tables(O.apply_pol_locks(cfg), plan, 0)
    q, sc = O.hg_int16(t.hgx)
    assert np.all(np.max(np.abs(q), axis=1) == 32767)
    err = np.abs(q * sc[:, None].astype(np.float64) - t.hgx) / sc[:, None]
    assert 0.4 < np.max(err) <= 0.51         # rounded to the nearest step


def test_overlap16_block_accumulation():
    # digHolo.cpp:2149-2341: integer products, float64 sums within a 16-row block,
    # float32 across blocks, the scales applied at the last block -- and with a
    # single 16-row block the scales are never applied (:2313-2329)
    rng = np.random.default_rng(7)
    G, w = 6, 32
    M = G * (G + 1) // 2
    Mi, Ni = O.hg_indices(G)
    qx = rng.integers(-32767, 32768, (G, w))
    sx = rng.uniform(1e-6, 1e-5, G).astype(np.float32)
    for h in (48, 16):
        f = rng.integers(-32767, 32768, (h, w)) + 1j * rng.integers(-32767, 32768, (h, w))
        qy = rng.integers(-32767, 32768, (G, h))
        sy = rng.uniform(1e-6, 1e-5, G).astype(np.float32)
        fs = np.float32(3e-5)
        c = O.overlap16(f, fs, qx, sx, qy, sy, G)
        exact_int = np.array([(qy[n] * (f @ qx[m])).sum() for m, n in zip(Mi, Ni)])
        if h == 16:
            assert np.array_equal(c, exact_int.astype(np.complex64))
            continue
        scale = np.array([float(np.float32(float(sx[m]) * float(sy[n]) * float(fs))) for m, n in zip(Mi, Ni)])
        exact = exact_int * scale
        blocks = np.array([[np.abs((qy[n, 16 * b:16 * b + 16] * (f[16 * b:16 * b + 16] @ qx[m])).sum())
                            for b in range(h // 16)] for m, n in zip(Mi, Ni)]).sum(axis=1)
        bound = 4 * 2.0 ** -24 * blocks * scale
        assert np.all(np.abs(c - exact) <= bound + 1e-30)
        assert len(c) == M


def test_reference_emulation_error_budget():
    # the reference's int16 tail against the exact path on 12-bit holograms: what a
    # comparison with the reference's own output can be held to
    spec = synth.spec_small()
    frames, c, s, A = synth.make_batch(spec, 3, seed=2, quantise=True)
    cfg = small_cfg(fill_factor=1)
    f0, c0, _ = O.process_batch(frames, cfg)
    f1, c1, _ = O.process_batch(frames, cfg, emulate_reference=True)
    ef, ec = O.relative_l2(f1, f0), O.relative_l2(c1, c0)
    assert 2e-5 < ef < 2e-4, ef
    assert 3e-6 < ec < 1e-4, ec


def test_linearity():
    spec = synth.spec_small()
    f1, _, _, _ = synth.make_batch(spec, 2, seed=4, quantise=False)
    f2, _, _, _ = synth.make_batch(spec, 2, seed=5, quantise=False)
    cfg = small_cfg(fill_factor=1)
    a, b = 0.7, -1.3
    _, cA, _ = O.process_batch(a * f1 + b * f2, cfg)
    _, c1, _ = O.process_batch(f1, cfg)
    _, c2, _ = O.process_batch(f2, cfg)
    assert O.relative_l2(cA, a * c1 + b * c2) < 1e-12


def test_power_meter_known_answers():
    # identity transfer matrix: every frame is exactly its own mode
    M = 10
    G = 4
    assert G * (G + 1) // 2 == M
    c = np.zeros((M, 1, M), dtype=complex)
    c[np.arange(M), 0, np.arange(M)] = 1.0
    lin = O.power_meter(c)
    assert np.isclose(lin[O.METRIC_IL], 1.0) and np.isclose(lin[O.METRIC_MDL], 1.0)
    assert np.isclose(lin[O.METRIC_DIAG], M / (M * M))        # total diagonal / (M N)
    assert np.isinf(lin[O.METRIC_SNRAVG])
    # diagonal gains: singular values are the gains
    gains = np.linspace(0.5, 2.0, M)
    c2 = c * gains[:, None, None]
    lin2 = O.power_meter(c2)
    assert np.isclose(lin2[O.METRIC_MDL], (0.5 / 2.0) ** 2)
    assert np.isclose(lin2[O.METRIC_IL], np.mean(gains ** 2))
    assert np.isclose(lin2[O.METRIC_DIAGBEST], 4.0) and np.isclose(lin2[O.METRIC_DIAGWORST], 0.25)
    # hand-computed crosstalk: 2 frames x 3 modes (G = 2)
    c3 = np.array([[[1.0, 0.5, 0.0]], [[0.25, 2.0, 1.0]]], dtype=complex)
    lin3 = O.power_meter(c3, do_mdl=False)
    sig = [1.0, 4.0]
    noi = [0.25, 0.0625 + 1.0]
    assert np.isclose(lin3[O.METRIC_SNRAVG], sum(sig) / sum(noi))
    assert np.isclose(lin3[O.METRIC_SNRBEST], max(s / n for s, n in zip(sig, noi)))
    assert np.isclose(lin3[O.METRIC_IL], (1 + 0.25 + 0.0625 + 4 + 1) / 2)
    # mode-group crosstalk: frame 0 is mode (0,0), group 1; frame 1 is mode (1,0), group 2
    mg_sig = 1.0 / 1 + (4.0 + 1.0) / 2
    mg_noi = 0.25 / 1 + 0.0625 / 2
    assert np.isclose(lin3[O.METRIC_SNRMG], mg_sig / mg_noi)


def test_power_meter_gram_eigenvalues_equal_svd():
    # the GPU path takes the singular values from the Gram matrix; both orientations agree
    rng = np.random.default_rng(9)
    for B, M in ((7, 15), (15, 7)):
        A = rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))
        s2 = np.linalg.svd(A.T, compute_uv=False) ** 2
        if B <= M:
            G = A @ A.conj().T
        else:
            G = A.T.conj() @ A
        ev = np.linalg.eigvalsh(G)
        assert np.allclose(np.sort(ev)[::-1][:len(s2)], s2)


def test_power_meter_independent_pols_and_single_frame():
    rng = np.random.default_rng(1)
    c = rng.standard_normal((6, 2, 10)) + 1j * rng.standard_normal((6, 2, 10))
    joint = O.power_meter(c)
    indep = O.power_meter(c, independent=True)
    s0 = np.linalg.svd(c[:, 0, :].T, compute_uv=False)
    s1 = np.linalg.svd(c[:, 1, :].T, compute_uv=False)
    s = np.concatenate([s0, s1])
    assert np.isclose(indep[O.METRIC_MDL], s.min() ** 2 / s.max() ** 2)
    assert not np.isclose(joint[O.METRIC_MDL], indep[O.METRIC_MDL])
    one = O.power_meter(c[:1])
    assert np.isclose(one[O.METRIC_IL], (np.abs(c[0]) ** 2).sum())
    assert np.isclose(one[O.METRIC_DIAG], (np.abs(c[0, :, 0]) ** 2).sum())


# ---------------------------------------------------------------------------
# plane analysis (power, centre of mass, peak, effective area)
# ---------------------------------------------------------------------------
def _axes(Nx, Ny, d=1.5):
    return ((np.arange(Nx) - Nx // 2) * d).astype(np.float32), ((np.arange(Ny) - Ny // 2) * d).astype(np.float32)


def test_analysis_single_pixel_known_answers():
    Nx, Ny = 16, 12
    x, y = _axes(Nx, Ny)
    planes = np.zeros((1, 1, Ny, Nx), complex)
    a = 3.0 - 4.0j                                        # |a| = 5
    planes[0, 0, 2, 11] = a
    out, summed = O.plane_analysis(planes, x, y)
    for b in (0, 1):                                      # the frame and the batch sum agree
        assert out[O.ANALYSIS_TOTALPOWER, 0, b] == pytest.approx(25.0 / (Nx * Ny))
        assert out[O.ANALYSIS_COMX, 0, b] == pytest.approx(x[11])
        assert out[O.ANALYSIS_COMY, 0, b] == pytest.approx(y[2])
        assert out[O.ANALYSIS_COMYWRAP, 0, b] == pytest.approx(y[2] + 1.5 * Ny)   # y < 0 wraps by Ny*dy
        assert out[O.ANALYSIS_AEFF, 0, b] == pytest.approx(1.5 * 1.5)            # one pixel of area dA
        assert out[O.ANALYSIS_MAXABSIDX, 0, b] == 2 * 2 * Nx + 11
    assert out[O.ANALYSIS_MAXABS, 0, 0] == pytest.approx(32767.0 / 5.0)          # peak of |V|
    assert out[O.ANALYSIS_MAXABS, 0, 1] == pytest.approx(32767.0 / math.sqrt(25.0 / (Nx * Ny)))
    assert summed[0, 2, 11] == pytest.approx(25.0 / (Nx * Ny))


def test_analysis_gaussian_effective_area_and_centre():
    # |E|^2 = exp(-2 r^2 / w^2): Petermann II area (int I)^2 / int I^2 = pi w^2
    Nx = Ny = 128
    d = 0.25
    x, y = _axes(Nx, Ny, d)
    w, x0, y0 = 4.0, 1.75, -2.5
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    E = np.exp(-((xd[None, :] - x0) ** 2 + (yd[:, None] - y0) ** 2) / w ** 2)
    out, _ = O.plane_analysis(E[None, None], x, y)
    assert out[O.ANALYSIS_AEFF, 0, 0] == pytest.approx(math.pi * w * w, rel=1e-6)
    assert out[O.ANALYSIS_COMX, 0, 0] == pytest.approx(x0, abs=1e-9)
    assert out[O.ANALYSIS_COMY, 0, 0] == pytest.approx(y0, abs=1e-9)
    assert out[O.ANALYSIS_TOTALPOWER, 0, 0] == pytest.approx(math.pi * w * w / 2 / d / d / (Nx * Ny), rel=1e-6)


def test_analysis_batch_sum_is_power_weighted():
    Nx, Ny = 8, 8
    x, y = _axes(Nx, Ny)
    planes = np.zeros((2, 1, Ny, Nx), complex)
    planes[0, 0, 1, 1] = 1.0
    planes[1, 0, 6, 5] = 2.0                              # 4x the power
    out, summed = O.plane_analysis(planes, x, y)
    assert out[O.ANALYSIS_COMX, 0, 2] == pytest.approx((x[1] + 4 * x[5]) / 5)
    assert out[O.ANALYSIS_COMY, 0, 2] == pytest.approx((y[1] + 4 * y[6]) / 5)
    assert out[O.ANALYSIS_MAXABSIDX, 0, 2] == 2 * 6 * Nx + 5
    assert out[O.ANALYSIS_TOTALPOWER, 0, 2] == pytest.approx(out[O.ANALYSIS_TOTALPOWER, 0, :2].sum())
    assert out[O.ANALYSIS_AEFF, 0, 2] == pytest.approx(1.5 * 1.5 * 25 / 17)      # (1+4)^2 / (1+16)
    assert np.allclose(summed[0], (np.abs(planes[:, 0]) ** 2).sum(0) / (Nx * Ny))


def test_fourier_plane_analysis_exclusions():
    # a DC offset, a pure-x and pure-y fringe and a Nyquist-row fringe all land on
    # excluded lines; an off-axis fringe outside 2 k0 sin(r) is the only power seen
    cfg = O.Config(frame_width=64, frame_height=64, pol_count=1, aperture_deg=0.3)
    W = H = 64
    yy, xx = np.mgrid[0:H, 0:W]
    base = 100.0 + 7 * np.cos(2 * np.pi * 3 * xx / W) + 5 * np.cos(2 * np.pi * 4 * yy / H) + 2 * (-1.0) ** yy
    out, summed = O.fourier_plane_analysis(base[None], cfg)
    assert out[O.ANALYSIS_TOTALPOWER, 0, 0] == pytest.approx(0.0, abs=1e-9)
    fr = base + 3 * np.cos(2 * np.pi * (12 * xx / W + 10 * yy / H))
    out, summed = O.fourier_plane_analysis(fr[None], cfg)
    Nx = W // 2 + 1
    # the off-axis term lands at (kx, ky) = (12, 10): power (3/2 * W * H)^2 / (Nx * H)
    assert out[O.ANALYSIS_TOTALPOWER, 0, 0] == pytest.approx((1.5 * W * H) ** 2 / (Nx * H), rel=1e-9)
    assert out[O.ANALYSIS_MAXABSIDX, 0, 0] == 2 * 10 * Nx + 12
    kx = O.k_axis_f32(W, cfg.pixel_size)
    assert out[O.ANALYSIS_COMX, 0, 0] == pytest.approx(kx[12], rel=1e-6)
    assert out[O.ANALYSIS_COMY, 0, 0] == pytest.approx(kx[10], rel=1e-6)


def test_fields_int16_scale_uses_full_range():
    rng = np.random.default_rng(5)
    V = rng.standard_normal((3, 2, 8, 8)) + 1j * rng.standard_normal((3, 2, 8, 8))
    cfg = O.Config()
    plan = O.Plan(W=64, H=64, cx=[0, 0], cy=[0, 0], w=8, h=8, SX=[0, 0], SY=[0, 0])
    out, _ = O.field_plane_analysis(V, cfg, plan)
    s = out[O.ANALYSIS_MAXABS, :, :3].T                   # [B, P]
    R, I = O.fields_int16(V, s)
    assert np.max(np.maximum(np.abs(R), np.abs(I))) <= 32767
    back = (R + 1j * I) / s[..., None, None]
    assert np.max(np.abs(back - V)) <= 0.51 / s.min() * math.sqrt(2)


# ---- input variants: transposed uint16, in-phase averaging, batch calibration ----

def test_transpose_u16_matches_index_definition():
    """convert_int16tofloat32 with transpose (digHolo.cpp:1206-1229) writes the
    input pixel at flat j*W + i to flat i*H + j; checked by an explicit loop on a
    non-square frame, and for square frames it is the matrix transpose."""
    rng = np.random.default_rng(0)
    H, W = 6, 10
    fr = rng.integers(0, 65535, (2, H, W)).astype(np.uint16)
    out = O.transpose_u16(fr)
    for f in range(2):
        raw, got = fr[f].reshape(-1), out[f].reshape(-1)
        for i in range(W):
            for j in range(H):
                assert got[i * H + j] == raw[j * W + i]
    sq = rng.integers(0, 65535, (1, 8, 8)).astype(np.uint16)
    assert np.array_equal(O.transpose_u16(sq)[0], sq[0].T)


def test_transposed_square_frames_give_the_untransposed_result():
    spec = synth.HologramSpec(frame_width=256, frame_height=256, pol_count=1, waist=200e-6, group_count=8)
    frames, _, _, _ = synth.make_batch(spec, 2, seed=3)
    cfg = small_cfg(frame_width=256, frame_height=256, group_count=8)
    f0, c0, _ = O.process_batch(frames, cfg)
    f1, c1, _ = O.process_batch(np.ascontiguousarray(frames.transpose(0, 2, 1)), cfg, transpose=True)
    assert O.relative_l2(f1, f0) < 1e-12 and O.relative_l2(c1, c0) < 1e-12


def test_average_windows_aligns_phases():
    """Copies of one window under arbitrary global phases average back to the first
    copy (the phase of each is aligned to the running sum); with the conjugate on the
    wrong operand the phases would add instead and this fails."""
    rng = np.random.default_rng(1)
    o = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    th = rng.uniform(0, 2 * np.pi, 5)
    stack = np.stack([o * np.exp(1j * t) for t in th])
    assert np.allclose(O.average_windows(stack), stack[0], rtol=1e-12, atol=1e-12)
    # a sign flip is a phase too: o and -o average to o, not to zero
    assert np.allclose(O.average_windows(np.stack([o, -o])), o)
    # independent noise averages down: the mean of A aligned noisy copies
    noisy = np.stack([o + 0.3 * (rng.normal(size=o.shape) + 1j * rng.normal(size=o.shape)) for _ in range(16)])
    assert O.relative_l2(O.average_windows(noisy), o) < 0.5 * O.relative_l2(noisy[0], o)


@pytest.mark.parametrize("mode", [0, 1])
def test_average_frame_layouts(mode):
    """AVGMODE_SEQUENTIAL reads [A A A B B B], INTERLACED [A B A B] (digHolo.h:655-657):
    phase-rotated copies of two holograms laid out per mode average back to each
    hologram's own single-frame result."""
    spec = synth.spec_small()
    frames, _, _, _ = synth.make_batch(spec, 2, seed=4, quantise=False)
    cfg = small_cfg()
    f_ref, c_ref, _ = O.process_batch(frames, cfg)
    A = 3
    # frames are real intensities: repeat them (the Fourier windows then differ by nothing)
    order = [b * A + a for b in range(2) for a in range(A)] if mode == 0 else [a * 2 + b for a in range(A) for b in range(2)]
    stacked = np.empty((2 * A,) + frames.shape[1:], frames.dtype)
    for b in range(2):
        for a in range(A):
            stacked[O.average_frame_index(b, a, 2, A, mode)] = frames[b]
    assert sorted(O.average_frame_index(b, a, 2, A, mode) for b in range(2) for a in range(A)) == sorted(order)
    f_avg, c_avg, _ = O.process_batch(stacked, cfg, avg_count=A, avg_mode=mode)
    assert O.relative_l2(f_avg, f_ref) < 1e-12 and O.relative_l2(c_avg, c_ref) < 1e-12


def test_batch_calibration_multiplies_fields_by_its_conjugate():
    """The calibration multiplies the Fourier window; the tilt step V = -conj(E ref)
    turns that into conj(cal) on the field and coefficients, cal[p % P][b % B]."""
    spec = synth.spec_dual_640()
    frames, _, _, _ = synth.make_batch(spec, 3, seed=6, quantise=False)
    cfg = O.Config(frame_width=640, frame_height=512, pol_count=2, fft_window_x=256, fft_window_y=256,
                   aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=10,
                   waist=[300e-6, 300e-6], beam_centre_x=[-1.6e-3, 1.6e-3])
    f0, c0, _ = O.process_batch(frames, cfg)
    cal = np.array([[0.5 + 0.2j, -1.1j], [2.0, 0.3 - 0.4j]])      # [pol][batch], batch wraps at 2
    f1, c1, _ = O.process_batch(frames, cfg, batch_cal=cal)
    for b in range(3):
        for p in range(2):
            k = np.conj(cal[p, b % 2])
            assert np.allclose(f1[b, p], k * f0[b, p]) and np.allclose(c1[b, p], k * c0[b, p])
    # a pol-independent calibration (one row) applies to both pols
    f2, _, _ = O.process_batch(frames, cfg, batch_cal=cal[:1])
    assert np.allclose(f2[1, 1], np.conj(cal[0, 1]) * f0[1, 1])


# ---- full-resolution IFFT (IFFTResolutionMode 0, the reference default) ----

def test_known_answer_full_resolution():
    """Full resolution keeps the window at its own frequencies in a W x H plane,
    drops the window-offset tilt from the reference wave and normalises by
    1/(W H): the rendered coefficients come back with the same closed-form
    scale as the low-res mode (a wrong norm or a leftover offset tilt fails)."""
    spec = synth.spec_small()
    spec.beam_centre_x = [37e-6, 0]
    spec.beam_centre_y = [-55e-6, 0]
    cfg = small_cfg(beam_centre_x=[37e-6, 0], beam_centre_y=[-55e-6, 0], resolution_mode=0)
    err, ratio, plan = recover(spec, cfg, B=3)
    assert err < 1e-5
    assert np.allclose(ratio, 1.0, atol=1e-6)
    fields, _, _ = O.process_batch(synth.make_batch(spec, 1, seed=2, quantise=False)[0], cfg)
    assert fields.shape[-2:] == (plan.H, plan.W)


def test_full_resolution_field_interpolates_the_low_res_field():
    """Both modes reconstruct the same band-limited field with the same total
    power (the norm 1/sqrt(W H fw fh) makes window size not matter): with an
    integer ratio W/w = H/h = 3 (640x512, FFT window 192 -> IFFT window 64) the
    full-resolution field at the low-res pixel positions is the low-res field
    times sqrt(w h / (W H)), after each mode's own tilt removal."""
    spec = synth.spec_dual_640()
    spec.group_count = 20
    frames, _, _, _ = synth.make_batch(spec, 1, seed=8, quantise=False)
    kw = dict(frame_width=640, frame_height=512, pol_count=2, fft_window_x=192, fft_window_y=192,
              aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=20,
              waist=[300e-6, 300e-6], beam_centre_x=[-1.6e-3, 1.6e-3])
    lo, c_lo, plan = O.process_batch(frames, O.Config(resolution_mode=1, **kw))
    hi, c_hi, _ = O.process_batch(frames, O.Config(resolution_mode=0, **kw))
    assert (plan.W, plan.w) == (192, 64)
    assert abs(np.sum(np.abs(hi) ** 2) / np.sum(np.abs(lo) ** 2) - 1) < 1e-9
    r = plan.W // plan.w
    ks = (np.arange(plan.w) - plan.w // 2) * r + plan.W // 2
    js = (np.arange(plan.h) - plan.h // 2) * r + plan.H // 2
    for p in range(2):
        sub = hi[0, p][np.ix_(js, ks)] / math.sqrt(plan.w * plan.h / (plan.W * plan.H))
        assert O.relative_l2(sub, lo[0, p]) < 1e-9
    # the HG coefficients are a property of the field, not of its sampling
    assert O.relative_l2(c_hi, c_lo) < 1e-4


# ---------------------------------------------------------------------------
# AutoAlign parameter estimate (digHolo.cpp:11517-11930, :11181-11515)
# ---------------------------------------------------------------------------

def test_r2r_matrices_are_fftw_redft10_and_redft01():
    """The unnormalised DCT-II / DCT-III FFTW defines equal scipy's type-2 /
    type-3 DCT without normalisation, and REDFT01 inverts REDFT10 up to 2n."""
    import scipy.fft
    rng = np.random.default_rng(0)
    for n in (5, 16, 129):
        x = rng.standard_normal(n)
        assert np.allclose(O.r2r_matrix(n, 10) @ x, scipy.fft.dct(x, type=2), rtol=1e-12, atol=1e-10)
        assert np.allclose(O.r2r_matrix(n, 1) @ x, scipy.fft.dct(x, type=3), rtol=1e-12, atol=1e-10)
        assert np.allclose(O.r2r_matrix(n, 1) @ O.r2r_matrix(n, 10), 2 * n * np.eye(n), atol=1e-9)
    m = rng.standard_normal((12, 9))
    assert np.allclose(O.r2r_2d(m, 10), scipy.fft.dctn(m, type=2), atol=1e-10)
    assert np.allclose(O.r2r_2d(O.r2r_2d(m, 10), 1), 4 * 12 * 9 * m, atol=1e-9)


def test_snap_gauss_layout():
    """Rows of 8 floor(Nx/8) counters laid over rows of Nx = 8k + 1: the weight
    sequence along the flat map equals the one of an 8k-wide map, the tail keeps
    its values, and for an 8k-wide map the weight is the separable Gaussian."""
    px = 20e-6
    wx2, wy2 = np.float32(-1.0 / (3.1e-4) ** 2), np.float32(-1.0 / (6.3e-4) ** 2)
    Ny = 10
    aligned = O.snap_gauss(np.ones((Ny, 16)), px, wx2, wy2)
    skewed = O.snap_gauss(np.ones((Ny, 17)), px, wx2, wy2)
    assert np.array_equal(skewed.ravel()[:Ny * 16], aligned.ravel())
    assert np.all(skewed.ravel()[Ny * 16:] == 1.0)
    gx = np.exp(float(wx2) * (np.arange(16) * px) ** 2)
    gy = np.exp(float(wy2) * (np.arange(Ny) * px) ** 2)
    assert np.allclose(aligned, np.outer(gy, gx), rtol=1e-5)


def _estimate_spec():
    spec = synth.spec_small()
    spec.tilt_x_deg = [1.52, 1.52]
    spec.tilt_y_deg = [1.47, 1.47]
    spec.beam_centre_x = [30e-6, 0.0]
    spec.beam_centre_y = [-20e-6, 0.0]
    spec.defocus = [0.5, 0.0]
    return spec


def test_snap_tilt_finds_the_rendered_carrier():
    """The peak of the filtered Fourier power sits on the rendered carrier: the
    tilt comes back within one Fourier pixel (k step 2 pi / (px W))."""
    spec = _estimate_spec()
    frames, _, _, _ = synth.make_batch(spec, 20, seed=3)
    cfg = small_cfg(fft_window_x=320, fft_window_y=256, aperture_deg=0.8)
    W, H, cx, cy = O.fft_window(cfg)
    spec_p = np.fft.rfft2(frames[:, cy[0]:cy[0] + H, cx[0]:cx[0] + W].astype(np.float64), axes=(-2, -1))
    k0 = 2 * math.pi / 1565e-9
    R = 2 * k0 * math.sin(0.8 * math.pi / 180)
    summed, _ = O._fourier_window_map(spec_p, W, H, 20e-6, 0.0, 0.0, R)
    tx, ty = O.snap_tilt(summed, W, H, 20e-6, 1565e-9, O.F32(0.8 * math.pi / 180))
    dkx, dky = 2 * math.pi / (20e-6 * W), 2 * math.pi / (20e-6 * H)
    assert abs(k0 * math.sin(math.radians(float(tx))) - k0 * math.radians(1.52)) <= dkx
    assert abs(k0 * math.sin(math.radians(float(ty))) - k0 * math.radians(1.47)) <= dky


def test_beam_size_relations():
    """AeffToMFD reads the rendered waist off the effective area of a beam that
    fills its groups (within 5 % for several waists); FocusEstimate is zero when
    both planes agree and grows as the Fourier-plane size shrinks."""
    for waist in (150e-6, 200e-6, 260e-6):
        spec = synth.spec_small()
        spec.waist = waist
        frames, _, _, _ = synth.make_batch(spec, 40, seed=7)
        cfg = small_cfg(waist=[waist, waist])
        fields, _, plan = O.process_batch(frames, cfg)
        params, _ = O.field_plane_analysis(fields, cfg, plan)
        w = float(O.aeff_to_mfd(params[O.ANALYSIS_AEFF, 0, -1], cfg.group_count)) * 0.5
        assert abs(w / waist - 1) < 0.05, (waist, w)
    assert O.focus_estimate(1e-4, 1e-4) == 0
    est = [float(O.focus_estimate(1e-4, 1e-4 * r)) for r in (0.9, 0.7, 0.5, 0.3)]
    assert all(b > a > 0 for a, b in zip(est, est[1:]))
    assert O.focus_estimate(0.0, 1e-4) == 0 and O.aeff_to_mfd(1.0, 0) == 0


def test_farfield_aeff_shift_invariant_and_least_without_curvature():
    """A circular shift of the fields leaves the far-field area unchanged; a
    quadratic phase of either sign enlarges it."""
    n = 64
    x = (np.arange(n) - n // 2) * 50e-6
    X, Y = np.meshgrid(x, x)
    E = np.exp(-(X ** 2 + Y ** 2) / (300e-6) ** 2)[None, None]
    a0 = O.farfield_aeff(E, 128, 128, 20e-6)[0]
    assert O.farfield_aeff(np.roll(E, (5, -3), axis=(-2, -1)), 128, 128, 20e-6)[0] == pytest.approx(a0, rel=1e-5)
    for alpha in (-2e6, 2e6, 6e6):
        q = np.exp(1j * alpha * (X ** 2 + Y ** 2))[None, None]
        assert O.farfield_aeff(E * q, 128, 128, 20e-6)[0] > a0 * 1.01


def test_quadratic_vertex_exact_for_parabola():
    x = np.array([-1.0, 0.3, 2.0, 2.5])
    assert O.quadratic_vertex(x, 3 * (x - 0.7) ** 2 + 1) == pytest.approx(0.7, abs=1e-12)
    assert math.isnan(O.quadratic_vertex([0.0, 1.0, 2.0], [1.0, 2.0, float("nan")]))


def test_autoalign_estimate_lands_near_the_rendered_parameters():
    """The estimate stage alone (reset, tilt search, window analysis, focus
    search) lands near the rendered beam: tilt within 0.02 deg, centre within a
    pixel, waist within 5 %, defocus on the right side."""
    spec = _estimate_spec()
    frames, _, _, _ = synth.make_batch(spec, 30, seed=5)
    cfg = small_cfg(tilt_x=[0.9, 0.9], tilt_y=[-0.4, -0.4], waist=[250e-6, 250e-6], fill_factor=1)
    c, info = O.autoalign_estimate(frames, cfg)
    assert abs(c.tilt_x[0] - 1.52) < 0.02 and abs(c.tilt_y[0] - 1.47) < 0.02
    assert abs(c.beam_centre_x[0] - 30e-6) < 20e-6 and abs(c.beam_centre_y[0] + 20e-6) < 20e-6
    assert abs(c.waist[0] / 200e-6 - 1) < 0.05
    assert 0.0 < c.defocus[0] < 1.0
    assert c.aperture_deg == pytest.approx(math.hypot(c.tilt_x[0], c.tilt_y[0]) / 3, rel=1e-5)


# ---------------------------------------------------------------------------
# Alignment-loss gradient (the adjoint the library computes on the GPU)
# ---------------------------------------------------------------------------

def _f32(v):
    return float(np.float32(v))


# central-difference steps: powers of two, so base +- step is exact in float32
# (the configuration is float32, as in the reference)
_GRAD_STEPS = (2.0 ** -28, 2.0 ** -14, 2.0 ** -16, 2.0 ** -30, 2.0 ** -16, 2.0 ** -30)
_GRAD_FIELDS = ("waist", "defocus", "tilt_x", "beam_centre_x", "tilt_y", "beam_centre_y")


def _fd_grad(frames, cfg, hold, p):
    import copy
    out = np.zeros(6)
    for k, name in enumerate(_GRAD_FIELDS):
        vals = []
        for sgn in (1, -1):
            c = copy.deepcopy(cfg)
            arr = getattr(c, name)
            arr[p] = arr[p] + sgn * _GRAD_STEPS[k]
            assert np.float32(arr[p]) == arr[p]
            _, coefs, _ = O.process_batch(frames, c, hold=hold)
            vals.append(O.alignment_loss(coefs))
        out[k] = (vals[0] - vals[1]) / (2 * _GRAD_STEPS[k])
    return out


@pytest.mark.parametrize("resolution_mode", [1, 0])
def test_alignment_gradient_matches_central_differences(resolution_mode):
    """The FP64 adjoint (tilt/defocus phase, window-ramp shift, waist basis
    derivative) equals central differences of the loss with the placement held."""
    spec = synth.spec_small()
    spec.tilt_x_deg, spec.tilt_y_deg = [1.52, 1.52], [1.47, 1.47]
    spec.beam_centre_x, spec.beam_centre_y, spec.defocus = [30e-6, 0], [-20e-6, 0], [0.5, 0]
    frames, _, _, _ = synth.make_batch(spec, 6, seed=11)
    cfg = small_cfg(tilt_x=[_f32(1.51)] * 2, tilt_y=[_f32(1.48)] * 2, beam_centre_x=[_f32(25e-6), 0.0],
                    beam_centre_y=[_f32(-15e-6), 0.0], defocus=[_f32(0.3), 0.0], waist=[_f32(210e-6)] * 2,
                    fill_factor=1, resolution_mode=resolution_mode)
    hold = O.hold_of(cfg)
    _, g = O.alignment_grad(frames, cfg, hold)
    fd = _fd_grad(frames, cfg, hold, 0)
    assert np.allclose(g[0], fd, rtol=2e-6, atol=0), (g[0], fd)


def test_alignment_gradient_dual_pol_locks():
    """Dual polarisation with tilt and waist locked: pol 0 collects both pols'
    terms of a locked parameter; the unlocked ones are per pol."""
    spec = synth.HologramSpec(frame_width=320, frame_height=256, pol_count=2, waist=150e-6, group_count=6)
    spec.beam_centre_x, spec.beam_centre_y = [-1.6e-3, 1.6e-3], [10e-6, -10e-6]
    spec.tilt_x_deg, spec.tilt_y_deg, spec.defocus = [1.5, 1.5], [1.5, 1.5], [0.2, -0.3]
    frames, _, _, _ = synth.make_batch(spec, 4, seed=2)
    cfg = O.Config(frame_width=320, frame_height=256, pol_count=2, fft_window_x=128, fft_window_y=128,
                   aperture_deg=0.7, tilt_x=[_f32(1.49), _f32(1.49)], tilt_y=[_f32(1.51), _f32(1.51)],
                   beam_centre_x=[_f32(-1.59e-3), _f32(1.605e-3)], beam_centre_y=[_f32(15e-6), _f32(-5e-6)],
                   defocus=[_f32(0.1), _f32(-0.2)], waist=[_f32(160e-6)] * 2, group_count=6, pol_lock_tilt=1,
                   pol_lock_waist=1)
    hold = O.hold_of(cfg)
    _, g = O.alignment_grad(frames, cfg, hold)
    for p in range(2):
        fd = _fd_grad(frames, cfg, hold, p)
        if p == 1:   # locked: pol 1's own values are not read
            fd[[0, 2, 4]] = 0.0
        assert np.allclose(g[p], fd, rtol=2e-6, atol=1e-9 * np.abs(g).max()), (p, g[p], fd)


@pytest.mark.parametrize("independent", [False, True])
def test_alignment_gradient_with_mdl_term(independent):
    """The mode-dependent-loss term's seed (extreme singular triplets) carried
    through the same adjoint equals central differences of the combined loss."""
    import copy
    spec = synth.HologramSpec(frame_width=320, frame_height=256, pol_count=2, waist=150e-6, group_count=5)
    spec.beam_centre_x, spec.beam_centre_y = [-1.6e-3, 1.6e-3], [10e-6, -10e-6]
    spec.defocus = [0.2, -0.3]
    frames, _, _, _ = synth.make_batch(spec, 5, seed=4)
    cfg = O.Config(frame_width=320, frame_height=256, pol_count=2, fft_window_x=128, fft_window_y=128,
                   aperture_deg=0.7, tilt_x=[_f32(1.49), _f32(1.52)], tilt_y=[_f32(1.51), _f32(1.48)],
                   beam_centre_x=[_f32(-1.59e-3), _f32(1.605e-3)], beam_centre_y=[_f32(15e-6), _f32(-5e-6)],
                   defocus=[_f32(0.1), _f32(-0.2)], waist=[_f32(160e-6)] * 2, group_count=5, pol_lock_waist=0)
    hold = O.hold_of(cfg)
    mu = 3e7
    _, g = O.alignment_grad(frames, cfg, hold, mdl_weight=mu, independent=independent)
    for p in range(2):
        fd = np.zeros(6)
        for k, name in enumerate(_GRAD_FIELDS):
            vals = []
            for sgn in (1, -1):
                c = copy.deepcopy(cfg)
                getattr(c, name)[p] += sgn * _GRAD_STEPS[k]
                _, coefs, _ = O.process_batch(frames, c, hold=hold)
                vals.append(O.alignment_loss(coefs, mu, independent))
            fd[k] = (vals[0] - vals[1]) / (2 * _GRAD_STEPS[k])
        assert np.allclose(g[p], fd, rtol=1e-5, atol=1e-8 * np.abs(g).max()), (p, g[p], fd)
    # the MDL term moves the gradient: it is not the power-only one
    _, g0 = O.alignment_grad(frames, cfg, hold)
    assert np.abs(g - g0).max() > 1e-3 * np.abs(g0).max()


# ---------------------------------------------------------------------------
# Basis transforms
# ---------------------------------------------------------------------------

def test_lg_blocks_are_unitary():
    for g, U in enumerate(O.lg_blocks(12)):
        assert np.allclose(U @ U.conj().T, np.eye(g + 1), atol=1e-12), g


def test_lg_basis_fields_are_laguerre_gauss_modes():
    """conj(U_g) rows applied to continuous HG modes give the textbook LG_{p,l}
    (up to one phase per mode), with l = m - n of the mode's index pair and one
    handedness convention for every mode."""
    from scipy.special import eval_genlaguerre
    G, w = 7, 1.0
    x = np.linspace(-6, 6, 241)
    X, Y = np.meshgrid(x, x)
    dx = x[1] - x[0]
    hx = O.hg_modes_1d(G, x, w, dx)
    Mi, Ni = O.hg_indices(G)
    Wt = O.basis_field_weights(G, 1)
    R2, PHI = X ** 2 + Y ** 2, np.arctan2(Y, X)
    signs = set()
    for k in range(len(Mi)):
        field = np.einsum("i,ix,iy->yx", Wt[k], hx[Mi], hx[Ni])
        l = Mi[k] - Ni[k]
        g = Mi[k] + Ni[k]
        p = (g - abs(l)) // 2
        best = 0.0
        for s in (1, -1):
            lg = ((np.sqrt(2 * R2) / w) ** abs(l) * eval_genlaguerre(p, abs(l), 2 * R2 / w ** 2) *
                  np.exp(-R2 / w ** 2) * np.exp(1j * s * l * PHI))
            lg = lg / np.sqrt((np.abs(lg) ** 2).sum())
            ov = abs(np.vdot(lg, field)) / np.sqrt((np.abs(field) ** 2).sum())
            if ov > best:
                best, bs = ov, s
        assert best > 1 - 1e-9, (k, l, p, best)
        if l != 0:
            signs.add(bs)
    assert len(signs) == 1


def test_transformed_coefficients_are_the_basis_field_overlaps():
    """c_LG[k] = sum conj(LG field_k) E and c_custom[j] = sum field_j E with the
    same fields the library exports (HG overlap without conjugation)."""
    spec = synth.spec_small()
    spec.group_count = 6
    frames, _, _, _ = synth.make_batch(spec, 2, seed=1)
    cfg = small_cfg(group_count=6)
    fields, coefs, _ = O.process_batch(frames, cfg)
    G, M = 6, 21
    lg = O.lg_transform(coefs, G)
    F = O.basis_fields(cfg, O.basis_field_weights(G, 1))
    proj = np.einsum("kpyx,bpyx->bpk", np.conj(F), fields)
    assert np.allclose(proj, lg, rtol=1e-9, atol=1e-9 * np.abs(lg).max())
    rng = np.random.default_rng(0)
    n_in, n_out = 30, 12                   # more inputs than modes: the extra rows are not used
    T = rng.standard_normal(n_in * n_out) + 1j * rng.standard_normal(n_in * n_out)
    cu = O.custom_transform(coefs, T, n_in, n_out)
    assert cu.shape[-1] == n_out
    F2 = O.basis_fields(cfg, O.basis_field_weights(G, 2, T, n_in, n_out))
    proj2 = np.einsum("kpyx,bpyx->bpk", F2, fields)
    assert np.allclose(proj2, cu, rtol=1e-9, atol=1e-9 * np.abs(cu).max())
    # identity transform leaves the HG coefficients
    eye = np.eye(M).reshape(-1)
    assert np.allclose(O.custom_transform(coefs, eye, M, M), coefs)


# ---------------------------------------------------------------------------
# wavelength sweeps (digHolo.cpp:2656-2700, :6577-6680, :7823-7865, :8644-8670)
# ---------------------------------------------------------------------------

def test_wavelength_ordering_is_the_sweep_transpose():
    L, n = 3, 4
    B = L * n
    # fast -> slow: element (sub, lam) of a [n, L] array moves to (lam, sub)
    grid = np.arange(B).reshape(n, L)
    out = np.array([O.wavelength_ordering_calc(b, L, B, 0, 1)[0] for b in range(B)])
    assert np.array_equal(grid.T.reshape(-1)[out], np.arange(B))   # slot out[b] holds input b
    lam = [O.wavelength_ordering_calc(b, L, B, 0, 0)[1] for b in range(B)]
    assert lam == [b % L for b in range(B)]
    lam_s = [O.wavelength_ordering_calc(b, L, B, 1, 1)[1] for b in range(B)]
    assert lam_s == [b // n for b in range(B)]
    for oi in (0, 1):   # equal orderings keep every element in place
        assert [O.wavelength_ordering_calc(b, L, B, oi, oi)[0] for b in range(B)] == list(range(B))
    # slow -> fast is the inverse of fast -> slow
    back = [O.wavelength_ordering_calc(o, L, B, 1, 0)[0] for o in out]
    assert back == list(range(B))


def test_frequency_sweep_linear_is_linear_in_frequency():
    lam = O.frequency_sweep_linear(1530e-9, 1600e-9, 8, 1565e-9)
    f = 299792458.0 / lam.astype(np.float64)
    df = np.diff(f)
    assert np.allclose(df, df[0], rtol=0, atol=3e-7 * f[0])   # float32 wavelengths: ~1e-7 of f each
    assert abs(lam[0] - 1530e-9) < 2e-15 and abs(lam[-1] - 1600e-9) < 2e-15
    assert np.all(np.diff(lam) > 0)
    # one step: the mean of the two ends, or the centre when neither is valid
    assert np.isclose(O.frequency_sweep_linear(1500e-9, 1600e-9, 1, 1565e-9)[0], 1550e-9, rtol=1e-6)
    assert O.frequency_sweep_linear(0, 0, 1, 1565e-9)[0] == np.float32(1565e-9)


def _sweep_cfg(lams, order=(0, 0), **kw):
    return small_cfg(wavelengths=list(lams), wavelength_ordering=order, defocus=[0.2, 0.2], **kw)


def _sweep_residuals(frames, c_true, cfg):
    _, coefs, _ = O.process_batch(frames, cfg)
    L = len(O.wavelength_axis(cfg))
    B = frames.shape[0]
    errs = []
    for b in range(B):
        o = O.wavelength_ordering_calc(b, L, B, *cfg.wavelength_ordering)[0]
        ct, cr = c_true[b, 0], coefs[o, 0]
        alpha = np.vdot(ct, cr) / np.vdot(ct, ct)
        errs.append(np.linalg.norm(cr - alpha * ct) / np.linalg.norm(cr))
    return np.array(errs)


@pytest.mark.parametrize("order", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sweep_recovers_frames_rendered_at_each_wavelength(order):
    lams = [1530e-9, 1565e-9, 1600e-9]
    spec = synth.spec_small()
    spec.defocus = [0.2, 0.2]
    frames, c, _, _ = synth.make_sweep_batch(spec, lams, 6, seed=5, order_in=order[0], quantise=False)
    errs = _sweep_residuals(frames, c, _sweep_cfg(lams, order))
    assert errs.max() < 1e-5
    # the centre wavelength alone leaves a tilt and a curvature on the off-centre frames
    errs_c = _sweep_residuals(frames, c, small_cfg(defocus=[0.2, 0.2]))
    lam_b = [b % 3 if not order[0] else b // 2 for b in range(6)]
    assert all(errs_c[b] > 1e-2 for b in range(6) if lam_b[b] != 1)


def test_sweep_aperture_follows_k0_and_the_window_spans_the_axis():
    # the circular aperture of each wavelength is centred at k0 sin(tilt) with
    # radius k0 sin(r): centroid and pixel count against the circle
    def stats(lams):
        cfg = _sweep_cfg(lams)
        plan = O.make_plan(cfg)
        kx = O.k_axis_f32(plan.W, cfg.pixel_size)
        dk = float(kx[1] - kx[0])
        out = []
        for lam in lams:
            t = O.pol_tables(O.apply_pol_locks(cfg), plan, 0, wavelength=lam)
            k0 = 2 * math.pi / lam
            cols = kx[t.col].astype(np.float64)
            centre = (t.mask * cols[None, :]).sum() / t.mask.sum() / (k0 * math.sin(math.radians(1.5)))
            area = t.mask.sum() / (math.pi * (k0 * math.sin(math.radians(0.7))) ** 2 / dk ** 2)
            out.append((centre, area))
        return plan, out
    plan, st = stats([1540e-9, 1565e-9, 1590e-9])
    for centre, area in st:
        assert abs(centre - 1) < 2e-3 and abs(area - 1) < 1e-2
    # the window is wider than either end wavelength's own window
    for lam in (1540e-9, 1590e-9):
        p1 = O.make_plan(small_cfg(wavelength=lam))
        assert plan.w >= p1.w and plan.h >= p1.h
    # the reference sizes the outer edge with the longest wavelength's k0 and the
    # inner edge with the shortest's (digHolo.cpp:7913-7955), so a wide sweep
    # clips the shortest wavelength's aperture on the outer side
    _, st = stats([1450e-9, 1565e-9, 1700e-9])
    assert st[0][1] < 0.95 and st[0][0] < 0.99
    assert abs(st[2][1] - 1) < 1e-2


def test_power_meter_sweep_known_answer():
    # wavelength l: an identity transfer scaled by a_l -> IL_l = a_l^2, MDL_l = 1
    L, M = 2, 6
    a = np.array([0.5, 0.9])
    c = np.zeros((L * M, 1, M), dtype=np.complex128)
    for l in range(L):
        for k in range(M):
            c[k * L + l, 0, k] = a[l]          # wavelength the fast axis
    lin = O.power_meter_sweep(c, L, 0)
    assert np.allclose(lin[:, O.METRIC_IL], a ** 2)
    assert np.allclose(lin[:, O.METRIC_MDL], 1.0)
    db = O.sweep_metrics_db(lin)
    assert np.isclose(db[O.METRIC_IL, L], 10 * np.log10(np.mean(a ** 2)))
    c_slow = c.reshape(M, L, 1, M).transpose(1, 0, 2, 3).reshape(L * M, 1, M)
    assert np.allclose(O.power_meter_sweep(c_slow, L, 1), lin)


# ---------------------------------------------------------------------------
# reference-wave calibration (digHolo.cpp:5309-5555, :6729-6850, :8253-8292)
# ---------------------------------------------------------------------------

def test_ref_cal_stored_forms():
    # intensity: sqrt(max / I), zero where I = 0
    st = O.ref_cal_from_intensity(np.array([[[0, 100, 400]]], dtype=np.uint16))
    assert np.allclose(st, [[[0, 2, 1]]])
    # field: peak magnitude 1, then (1 / re, -1 / im); a zero part zeroes the pixel
    z = np.array([[[3 + 4j, 1 + 0j, 0.5 - 0.5j]]], dtype=np.complex64)
    st = O.ref_cal_from_field(z)
    assert np.allclose(st[0, 0, 0], 5 / 3 - 1j * 5 / 4)
    assert st[0, 0, 1] == 0
    assert np.allclose(st[0, 0, 2], 10 - 1j * (-10))


@pytest.mark.parametrize("resolution_mode", [1, 0])
def test_uniform_ref_calibration_is_the_identity(resolution_mode):
    cfg = small_cfg(resolution_mode=resolution_mode)
    plan = O.make_plan(cfg)
    flat = O.ref_cal_from_intensity(np.full((1, 256, 320), 1234, np.uint16))
    planes = O.ref_cal_planes(flat, cfg, plan)
    assert np.allclose(planes, 1.0, atol=1e-12)          # only the zero frequency survives, with gain 1
    spec = synth.spec_small()
    frames, _, _, _ = synth.make_batch(spec, 2, seed=3, quantise=False)
    f0, c0, _ = O.process_batch(frames, cfg)
    f1, c1, _ = O.process_batch(frames, cfg, ref_cal=flat)
    assert O.relative_l2(c1, c0) < 1e-12 and O.relative_l2(f1, f0) < 1e-12


@pytest.mark.parametrize("resolution_mode", [1, 0])
def test_intensity_calibration_undoes_a_nonuniform_reference(resolution_mode):
    spec = synth.spec_small()
    spec.ref_envelope_waist = 4e-3
    cfg = small_cfg(resolution_mode=resolution_mode)
    frames, c_true, _, _ = synth.make_batch(spec, 3, seed=8, quantise=False)
    cal = O.ref_cal_from_intensity(synth.reference_intensity_frame(spec)[None])

    def worst(coefs):
        errs = []
        for b in range(3):
            ct, cr = c_true[b, 0], coefs[b, 0]
            alpha = np.vdot(ct, cr) / np.vdot(ct, ct)
            errs.append(np.linalg.norm(cr - alpha * ct) / np.linalg.norm(cr))
        return max(errs)
    e_raw = worst(O.process_batch(frames, cfg)[1])
    e_cal = worst(O.process_batch(frames, cfg, ref_cal=cal)[1])
    print(e_raw, e_cal)
    assert e_raw > 1e-3 and e_cal < e_raw / 20
    # multiplying by the envelope instead of dividing by it makes things worse
    wrong = (1.0 / np.where(np.abs(cal) > 0, cal, 1)).astype(np.complex64)
    assert worst(O.process_batch(frames, cfg, ref_cal=wrong)[1]) > e_raw


def test_constant_field_calibration_scales_by_the_conjugate():
    cfg = small_cfg()
    z = np.full((1, 256, 320), 0.6 + 0.8j, np.complex64)
    cal = O.ref_cal_from_field(z)
    C = complex(cal[0, 0, 0])
    assert np.isclose(C, 1 / 0.6 - 1j / 0.8, rtol=1e-6)
    spec = synth.spec_small()
    frames, _, _, _ = synth.make_batch(spec, 2, seed=4, quantise=False)
    _, c0, _ = O.process_batch(frames, cfg)
    _, c1, _ = O.process_batch(frames, cfg, ref_cal=cal)
    assert np.allclose(c1, np.conj(C) * c0, rtol=1e-9, atol=1e-12 * np.abs(c0).max())


def test_quadratic_vertex_and_window_radii():
    # fitToQuadratic (digHolo.cpp:3146-3257) returns the vertex of the
    # least-squares parabola; any three samples of a parabola give its vertex
    for x0, scale in ((2.0, 1.0), (1.2e-4, 1e-5), (-0.4, 0.2)):
        xs = x0 + scale * np.array([-1.3, 0.2, 0.9, 2.1])
        ys = -3.0 * ((xs - x0) / scale) ** 2 + 5.0
        assert O.quadratic_vertex(xs[:3], ys[:3]) == pytest.approx(x0, rel=1e-9, abs=1e-12 * scale)
        assert O.quadratic_vertex(xs, ys) == pytest.approx(x0, rel=1e-9, abs=1e-12 * scale)
    # least squares through noisy samples: numpy's polyfit on the same points
    rng = np.random.default_rng(3)
    xs = np.linspace(-1, 1, 6)
    ys = -(xs - 0.3) ** 2 + 0.05 * rng.standard_normal(6)
    a, b, _ = np.polyfit(xs, ys, 2)
    assert O.quadratic_vertex(xs, ys) == pytest.approx(-b / (2 * a), rel=1e-9)
    assert np.isnan(O.quadratic_vertex([1.0, 2.0, np.nan], [0.0, 1.0, 0.0]))
    assert np.isnan(O.quadratic_vertex([1.0, 1.0, 1.0], [0.0, 1.0, 0.0]))
    # window radii (digHolo.cpp:11960-11993): (sqrt(68) - 2)/16 of lambda/(2 px) in
    # degrees, and a third of the smallest tilt magnitude
    r = O.window_radius_max(20e-6, 1565e-9)
    assert float(r) == pytest.approx((np.sqrt(68) - 2) / 16 * 1565e-9 / 40e-6 * 180 / np.pi, rel=1e-6)
    cfg = small_cfg(pol_count=2, tilt_x=[3.0, 1.2], tilt_y=[4.0, 0.5])
    assert float(O.window_radius_rec(cfg)) == pytest.approx(1.3 / 3, rel=1e-6)
