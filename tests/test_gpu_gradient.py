"""Alignment-loss gradient on the GPU path (digHoloAutoAlignGradient; no
counterpart in the reference) and the gradient-based alignment beside the
reference's search.

* the adjoint kernels' gradient equals the FP64 oracle's at the same window
  placement (each component within 2e-3 of the largest of its kind, in units of
  the parameter's search step and the loss), for low-res and full-resolution
  fields and dual polarisation with locks;
* central differences of the library's own FP32 loss with the placement held
  (holdPlacement) agree with it (5e-3);
* the check-only torch.autograd.Function returns the same gradient through
  backward();
* from a perturbed start, the gradient search ends at an insertion loss within
  0.01 dB of the reference's tweak search from the same start.
"""
import time

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O

pytestmark = pytest.mark.gpu

# search steps per parameter (AutoAlign order): the units the comparison uses
SCALE = np.array([10e-6, 0.1, 0.01, 10e-6, 0.01, 10e-6])


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def f32(v):
    return float(np.float32(v))


def small_case(resolution_mode=1):
    spec = synth.spec_small()
    spec.tilt_x_deg, spec.tilt_y_deg = [1.52, 1.52], [1.47, 1.47]
    spec.beam_centre_x, spec.beam_centre_y, spec.defocus = [30e-6, 0], [-20e-6, 0], [0.5, 0]
    cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), group_count=10,
               resolution_mode=resolution_mode)
    start = dict(tilt_x=(f32(1.51), f32(1.51)), tilt_y=(f32(1.48), f32(1.48)), beam_centre_x=(f32(25e-6), 0.0),
                 beam_centre_y=(f32(-15e-6), 0.0), defocus=(f32(0.3), 0.0), waist=f32(210e-6))
    return spec, cfg, start


def dual_case():
    spec = synth.spec_dual_640()
    spec.group_count = 15
    spec.tilt_x_deg, spec.tilt_y_deg = [1.50, 1.53], [1.50, 1.46]
    spec.beam_centre_y, spec.defocus = [10e-6, -15e-6], [0.3, -0.4]
    cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), group_count=15,
               pol_lock_tilt=0, pol_lock_waist=1)
    start = dict(tilt_x=(f32(1.51), f32(1.52)), tilt_y=(f32(1.49), f32(1.47)),
                 beam_centre_x=(f32(-1.605e-3), f32(1.59e-3)), beam_centre_y=(f32(0.0), f32(-5e-6)),
                 defocus=(f32(0.2), f32(-0.3)), waist=f32(290e-6))
    return spec, cfg, start


def oracle_cfg(cfg, start):
    P = cfg["pol_count"]
    pad = lambda v: list(v) + [v[0]] * (2 - len(v))
    return O.Config(frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                    fft_window_x=cfg["fft_window"][0], fft_window_y=cfg["fft_window"][1], aperture_deg=0.7,
                    resolution_mode=cfg.get("resolution_mode", 1), beam_centre_x=pad(start["beam_centre_x"]),
                    beam_centre_y=pad(start["beam_centre_y"]), tilt_x=pad(start["tilt_x"]),
                    tilt_y=pad(start["tilt_y"]), defocus=pad(start["defocus"]), waist=[start["waist"]] * 2,
                    group_count=cfg["group_count"], pol_lock_tilt=cfg.get("pol_lock_tilt", 0),
                    pol_lock_waist=cfg.get("pol_lock_waist", 1))


@pytest.mark.parametrize("case,mdl,B", [("small", 0, 6), ("small_fullres", 0, 6), ("dual", 0, 6),
                                         ("dual", 1, 6), ("small", 1, 60)])
def test_gradient_matches_oracle(lib, case, mdl, B):
    """mdl = 1 adds the mode-dependent-loss term with a weight of 5 % of the
    power term (B = 6 < modes: Gram of the frames; B = 60 > 55 modes: Gram of
    the modes)."""
    spec, cfg, start = dual_case() if case == "dual" else small_case(0 if case == "small_fullres" else 1)
    frames = np.ascontiguousarray(synth.make_batch(spec, B, seed=11)[0])
    P = cfg["pol_count"]
    print(f"case {case}: {frames.shape} mdl {mdl}", flush=True)
    oc = oracle_cfg(cfg, start)
    hold = O.hold_of(oc)
    mu = 0.05 * abs(O.alignment_grad(frames, oc, hold)[0]) if mdl else 0.0
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **start})
        assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
        loss, grad = api.alignment_gradient(lib, h, mdl_weight=mu)
    finally:
        lib.digHoloDestroy(h)
    want_loss, want = O.alignment_grad(frames, oc, hold, mdl_weight=mu)
    print("library", loss, grad[:P], "\noracle ", want_loss, want, flush=True)
    assert loss == pytest.approx(want_loss, rel=2e-6)
    a = grad[:P] * SCALE / abs(want_loss)
    b = want * SCALE / abs(want_loss)
    tol = 2e-3 * np.maximum(np.abs(b).max(axis=0), 1e-6)
    assert np.all(np.abs(a - b) <= tol), (a, b)


def test_gradient_central_differences_with_held_placement(lib):
    spec, cfg, start = small_case()
    B = 8
    frames = np.ascontiguousarray(synth.make_batch(spec, B, seed=3)[0])
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **start})
        assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
        base = api.get_align_params(lib, h, 1)
        loss, grad = api.alignment_gradient(lib, h)
        steps = np.array([2e-6, 0.02, 2e-3, 1e-6, 2e-3, 1e-6])
        fd = np.zeros(6)
        for k in range(6):
            vals = []
            for sgn in (1, -1):
                prm = base.copy()
                prm[0, k] += sgn * steps[k]
                api.set_align_params(lib, h, prm)
                vals.append(api.alignment_gradient(lib, h, hold=True)[0])
            fd[k] = (vals[0] - vals[1]) / (2 * steps[k])
            api.set_align_params(lib, h, base)
    finally:
        lib.digHoloDestroy(h)
    print("adjoint", grad[0], "\ncentral", fd, flush=True)
    a, b = grad[0] * SCALE / abs(loss), fd * SCALE / abs(loss)
    assert np.all(np.abs(a - b) <= 5e-3 * np.abs(a).max()), (a, b)


def test_torch_autograd_function(lib):
    import torch
    from digholo_amd.autograd import alignment_loss
    spec, cfg, start = small_case()
    B = 4
    frames = np.ascontiguousarray(synth.make_batch(spec, B, seed=5)[0])
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **start})
        assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
        params = torch.tensor(api.get_align_params(lib, h, 1), dtype=torch.float64, requires_grad=True)
        loss = alignment_loss(params, lib, h)
        loss.backward()
        _, grad = api.alignment_gradient(lib, h, hold=True)
        assert torch.allclose(params.grad, torch.as_tensor(grad[:1]), rtol=1e-6, atol=0)
        # scaled central difference along the gradient through the wrapper
        d = params.grad / params.grad.norm()
        d = d * torch.as_tensor(SCALE) * 0.05
        lp = alignment_loss(params.detach() + d, lib, h, True).item()
        lm = alignment_loss(params.detach() - d, lib, h, True).item()
        assert (lp - lm) / 2 == pytest.approx(float((params.grad * d).sum()), rel=5e-3)
    finally:
        lib.digHoloDestroy(h)


def test_gradient_search_matches_tweak(lib):
    spec, cfg, start = dual_case()
    B = 60
    frames = np.ascontiguousarray(synth.make_batch(spec, B, seed=7)[0])
    out = {}
    for method in ("tweak", "gradient"):
        h = lib.digHoloCreate()
        try:
            api.configure(lib, h, **{**cfg, **start})
            lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_TWEAK)
            assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
            api.process_batch(lib, h, frames)
            t0 = time.perf_counter()
            ret = lib.digHoloAutoAlign(h) if method == "tweak" else lib.digHoloAutoAlignGradientSearch(h, 40, 0.0)
            dt = time.perf_counter() - t0
            il = lib.digHoloAutoAlignGetMetric(h, O.METRIC_IL)
            out[method] = (ret, il, dt, api.get_align_params(lib, h, 2))
        finally:
            lib.digHoloDestroy(h)
        print(method, out[method][:3], "\n", out[method][3], flush=True)
    assert np.isfinite(out["gradient"][0])
    assert out["gradient"][1] >= out["tweak"][1] - 0.01


def c3_case(B=6, seed=3):
    """The differentiable-alignment workload: 640x512 dual-pol frames, 45 groups,
    true tilt (+-0.2 deg), defocus (+-0.5 dioptre), waist (+-10 %) and beam
    centre (+-50 um) drawn from the seed; the gradient is taken at the nominal
    start (tilt 1.5 deg, centres -/+1.6 mm, defocus 0, waist 300 um) with every
    pol lock off, so each pol has its own 6 parameters."""
    rng = np.random.default_rng(seed)
    spec = synth.spec_dual_640()
    spec.tilt_x_deg = list(1.5 + rng.uniform(-0.2, 0.2, 2))
    spec.tilt_y_deg = list(1.5 + rng.uniform(-0.2, 0.2, 2))
    spec.defocus = list(rng.uniform(-0.5, 0.5, 2))
    spec.waist = 300e-6 * (1 + rng.uniform(-0.1, 0.1))
    spec.beam_centre_x = [(-1.6e-3, 1.6e-3)[p] + rng.uniform(-50e-6, 50e-6) for p in range(2)]
    spec.beam_centre_y = list(rng.uniform(-50e-6, 50e-6, 2))
    frames = np.ascontiguousarray(synth.make_batch(spec, B, seed=seed)[0])
    cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), group_count=45,
               pol_lock_tilt=0, pol_lock_defocus=0, pol_lock_waist=0)
    start = dict(tilt_x=(1.5, 1.5), tilt_y=(1.5, 1.5), beam_centre_x=(-1.6e-3, 1.6e-3),
                 beam_centre_y=(0.0, 0.0), defocus=(0.0, 0.0), waist=300e-6)
    return frames, cfg, start


def oracle_params(oc):
    """[pol][waist, defocus, tilt x, centre x, tilt y, centre y] of an oracle Config."""
    return np.array([[oc.waist[p], oc.defocus[p], oc.tilt_x[p], oc.beam_centre_x[p], oc.tilt_y[p],
                      oc.beam_centre_y[p]] for p in range(2)], dtype=np.float64)


def oracle_with(oc, prm):
    import dataclasses
    return dataclasses.replace(oc, waist=[float(prm[0, 0]), float(prm[1, 0])],
                               defocus=[float(prm[0, 1]), float(prm[1, 1])],
                               tilt_x=[float(prm[0, 2]), float(prm[1, 2])],
                               beam_centre_x=[float(prm[0, 3]), float(prm[1, 3])],
                               tilt_y=[float(prm[0, 4]), float(prm[1, 4])],
                               beam_centre_y=[float(prm[0, 5]), float(prm[1, 5])])


def oracle_central_differences(frames, oc, hold, mu, steps):
    """FP64 central differences of the oracle loss with the placement held,
    Richardson-extrapolated (h and h/2, error O(h^4)).  Each perturbed value is
    rounded to float32 first (the oracle reads its parameters as float32) and
    the difference quotient uses the rounded values."""
    base = oracle_params(oc)
    out = np.zeros((2, 6))

    def loss_at(prm):
        _, coefs, _ = O.process_batch(frames, oracle_with(oc, prm), hold=hold)
        return O.alignment_loss(coefs, mu)

    for p in range(2):
        for k in range(6):
            d = []
            for hstep in (steps[k], steps[k] / 2):
                vals, xs = [], []
                for sgn in (1, -1):
                    prm = base.copy()
                    prm[p, k] = float(np.float32(base[p, k] + sgn * hstep))
                    xs.append(prm[p, k])
                    vals.append(loss_at(prm))
                d.append((vals[0] - vals[1]) / (xs[0] - xs[1]))
            out[p, k] = (4 * d[1] - d[0]) / 3
    return out


def test_c3_gradient_g45_mdl_vs_fp64_central_differences(lib):
    """Differentiable-alignment bar: at 640x512 dual pol, 45 groups, with the
    mode-dependent-loss term on (weight 5 % of the power term), every
    component of the GPU adjoint gradient (12 = 2 pols x 6 parameters) is within
    1e-3 relative of FP64 central differences of the oracle's loss with the
    placement held (the FP64 oracle's adjoint meets the same bar).  The same
    gradient against central differences of the library's own loss is printed
    beside it: that loss comes back as a float32 (about 4.8e9 here, one unit in
    the last place = 512), so differencing noise and step truncation set that
    figure, and its bar is 3e-2."""
    frames, cfg, start = c3_case()
    oc = O.Config(frame_width=640, frame_height=512, pol_count=2, fft_window_x=256, fft_window_y=256,
                  aperture_deg=0.7, beam_centre_x=list(start["beam_centre_x"]),
                  beam_centre_y=list(start["beam_centre_y"]), tilt_x=list(start["tilt_x"]),
                  tilt_y=list(start["tilt_y"]), defocus=list(start["defocus"]), waist=[start["waist"]] * 2,
                  group_count=45, pol_lock_tilt=0, pol_lock_defocus=0, pol_lock_waist=0)
    hold = O.hold_of(oc)
    mu = 0.05 * abs(O.alignment_grad(frames, oc, hold)[0])
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **start})
        assert lib.digHoloSetBatchUint16(h, frames.shape[0], frames.ctypes.data, 0) == 0
        loss, grad = api.alignment_gradient(lib, h, mdl_weight=mu)
        # the library's own FP32 loss differenced with the placement held; the
        # C API's waist setter writes every pol (digHolo.cpp:14058-14063), so
        # the waist is differenced for both pols at once (d/dw0 + d/dw1)
        base = api.get_align_params(lib, h, 2)
        fsteps = np.array([2e-6, 0.02, 2e-3, 1e-6, 2e-3, 1e-6])
        fd32 = np.zeros((2, 6))
        for p in range(2):
            for k in range(6):
                if k == 0 and p == 1:
                    continue
                vals, xs = [], []
                for sgn in (1, -1):
                    prm = base.copy()
                    cols = [0, 1] if k == 0 else [p]
                    for q in cols:
                        prm[q, k] = float(np.float32(base[q, k] + sgn * fsteps[k]))
                    xs.append(prm[p, k])
                    api.set_align_params(lib, h, prm)
                    vals.append(api.alignment_gradient(lib, h, hold=True, mdl_weight=mu)[0])
                fd32[p, k] = (vals[0] - vals[1]) / (xs[0] - xs[1])
                api.set_align_params(lib, h, base)
    finally:
        lib.digHoloDestroy(h)
    want_loss, want_adj = O.alignment_grad(frames, oc, hold, mdl_weight=mu)
    fd64 = oracle_central_differences(frames, oc, hold, mu, np.array([1e-6, 0.01, 1e-3, 0.5e-6, 1e-3, 0.5e-6]))
    print("loss", loss, want_loss, flush=True)
    print("GPU adjoint\n", grad, "\nFP64 central differences\n", fd64, "\nFP64 adjoint\n", want_adj,
          "\nFP32 central differences (library)\n", fd32, flush=True)
    rel64 = np.abs(grad - fd64) / np.abs(fd64)
    g32, f32w = grad.copy(), fd32.copy()
    g32[0, 0] = grad[0, 0] + grad[1, 0]        # both waists moved together
    g32[1, 0] = f32w[1, 0] = 0.0
    ref32 = np.abs(fd64).copy()
    ref32[0, 0] = abs(fd64[0, 0] + fd64[1, 0])
    ref32[1, 0] = 1.0
    rel32 = np.abs(g32 - f32w) / ref32
    print(f"max relative error per component: vs FP64 central differences {rel64.max():.2e}, "
          f"vs FP32 central differences {rel32.max():.2e}, oracle adjoint vs FP64 differences "
          f"{(np.abs(want_adj - fd64) / np.abs(fd64)).max():.2e}", flush=True)
    assert loss == pytest.approx(want_loss, rel=2e-6)
    assert rel64.max() <= 1e-3, rel64
    assert (np.abs(want_adj - fd64) / np.abs(fd64)).max() <= 1e-3
    assert rel32.max() <= 3e-2, rel32
