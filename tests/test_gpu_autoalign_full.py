"""AutoAlign FULL and ESTIMATE modes (digHolo.cpp:12237-12970) on the GPU path:
the parameter estimate (:11517-11930, the DCT-domain tilt search, window
analysis and fitted beam-size relations) and the far-field focus search
(:11181-11515), then in FULL mode the tweak search from the estimate.

* ESTIMATE: what the library estimates from the frames alone equals the
  oracle's restatement of the same estimate on the same frames -- tilt to 1e-4
  degree (the same Fourier pixel), window radius to 1e-5 degree, beam centre to
  0.1 um, waist to 1e-4 relative, defocus to 2e-3 dioptre -- and the goal metric
  it returns equals the oracle's power meter at those parameters (1e-3 dB), with
  the whole-half-frame FFT window the estimate runs under;
* FULL: from tilt and centres far off (the estimate resets them anyway) the
  parameters come back to the rendered ones (tilt 0.01 degree, centre 5 um,
  waist 3 %, defocus 0.1 dioptre) and the goal is within 0.02 dB of its value at
  the rendered parameters.
"""
import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O
from test_gpu_autoalign import lib_metrics, oracle_metrics, read_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def small_spec():
    spec = synth.spec_small()
    spec.tilt_x_deg = [1.52, 1.52]
    spec.tilt_y_deg = [1.47, 1.47]
    spec.beam_centre_x = [30e-6, 0.0]
    spec.beam_centre_y = [-20e-6, 0.0]
    spec.defocus = [0.5, 0.0]
    cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), group_count=10)
    return spec, cfg


def dual_spec():
    spec = synth.spec_dual_640()
    spec.group_count = 20
    spec.tilt_x_deg = [1.50, 1.53]
    spec.tilt_y_deg = [1.50, 1.46]
    spec.beam_centre_y = [10e-6, -15e-6]
    spec.defocus = [0.3, -0.4]
    cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), group_count=20)
    return spec, cfg


START = dict(tilt_x=(0.9, 0.9), tilt_y=(-0.4, -0.4), beam_centre_x=(0.0, 0.0), beam_centre_y=(0.0, 0.0),
             defocus=(1.0, 1.0), waist=250e-6)


@pytest.mark.parametrize("case", ["small", "dual"])
def test_estimate_matches_oracle(lib, case):
    spec, cfg = small_spec() if case == "small" else dual_spec()
    B = 30 if case == "small" else 24
    frames, _, _, _ = synth.make_batch(spec, B, seed=5)
    frames = np.ascontiguousarray(frames)
    P = cfg["pol_count"]
    print(f"case {case}: {frames.shape}", flush=True)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **START})
        lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_ESTIMATE)
        assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
        ret = lib.digHoloAutoAlign(h)
        prm = read_params(lib, h, P)
        got = lib_metrics(lib, h)
        wx, wy = api.c_int(), api.c_int()
        lib.digHoloConfigGetfftWindowSize(h, api.C.byref(wx), api.C.byref(wy))
        assert (wx.value, wy.value) == tuple(cfg["fft_window"])     # restored afterwards
    finally:
        lib.digHoloDestroy(h)
    pad = lambda v: list(v) + [v[0]] * (2 - len(v))
    oc = O.Config(frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                  fft_window_x=cfg["fft_window"][0], fft_window_y=cfg["fft_window"][1], aperture_deg=0.7,
                  beam_centre_x=pad(START["beam_centre_x"]), beam_centre_y=pad(START["beam_centre_y"]),
                  tilt_x=pad(START["tilt_x"]), tilt_y=pad(START["tilt_y"]), defocus=pad(START["defocus"]),
                  waist=[START["waist"]] * 2, group_count=cfg["group_count"])
    want, info = O.autoalign_estimate(frames, oc)
    print("library", prm, flush=True)
    print("oracle ", dict(tilt_x=want.tilt_x[:P], tilt_y=want.tilt_y[:P], centre_x=want.beam_centre_x[:P],
                          centre_y=want.beam_centre_y[:P], waist=want.waist[:P], defocus=want.defocus[:P],
                          aperture=want.aperture_deg), info, flush=True)
    assert prm["aperture_deg"] == pytest.approx(want.aperture_deg, abs=1e-5)
    for p in range(P):
        assert prm["tilt_x"][p] == pytest.approx(want.tilt_x[p], abs=1e-4)
        assert prm["tilt_y"][p] == pytest.approx(want.tilt_y[p], abs=1e-4)
        assert prm["beam_centre_x"][p] == pytest.approx(want.beam_centre_x[p], abs=1e-7)
        assert prm["beam_centre_y"][p] == pytest.approx(want.beam_centre_y[p], abs=1e-7)
        assert prm["waist"][p] == pytest.approx(want.waist[p], rel=1e-4)
        assert prm["defocus"][p] == pytest.approx(want.defocus[p], abs=2e-3)
    # the metrics of the final evaluation, under the whole-half-frame window
    _, coefs, _ = O.process_batch(frames, want)
    ref = O.metrics_db(O.power_meter(coefs))
    finite = np.isfinite(ref)
    assert np.allclose(got[finite], ref[finite], atol=1e-3, rtol=0)
    assert ret == pytest.approx(got[0], abs=1e-6)


@pytest.mark.parametrize("case", ["small", "dual"])
def test_full_mode_recovers_rendered_parameters(lib, case):
    spec, cfg = small_spec() if case == "small" else dual_spec()
    B = 40 if case == "small" else 40
    frames, _, _, _ = synth.make_batch(spec, B, seed=11)
    frames = np.ascontiguousarray(frames)
    P = cfg["pol_count"]
    print(f"case {case}: {frames.shape}", flush=True)
    truth = dict(tilt_x=spec.tilt_x_deg[:P], tilt_y=spec.tilt_y_deg[:P], beam_centre_x=spec.beam_centre_x[:P],
                 beam_centre_y=spec.beam_centre_y[:P], defocus=spec.defocus[:P])
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **{**cfg, **START})
        lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_FULL)
        assert lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0) == 0
        ret = lib.digHoloAutoAlign(h)
        after = lib_metrics(lib, h)
        prm = read_params(lib, h, P)
        api.configure(lib, h, **{**cfg, **truth, "waist": spec.waist, "aperture_deg": prm["aperture_deg"]})
        api.process_batch(lib, h, frames)
        assert lib.digHoloAutoAlignCalcMetrics(h) == 0
        at_truth = lib_metrics(lib, h)
    finally:
        lib.digHoloDestroy(h)
    print(f"goal: truth {at_truth[0]:.4f} end {after[0]:.4f} (returned {ret:.4f})", flush=True)
    print("params", prm, flush=True)
    assert np.isfinite(ret) and ret == pytest.approx(after[0], abs=1e-6)
    assert after[0] >= at_truth[0] - 0.02
    for p in range(P):
        assert abs(prm["tilt_x"][p] - spec.tilt_x_deg[p]) < 0.01
        assert abs(prm["tilt_y"][p] - spec.tilt_y_deg[p]) < 0.01
        assert abs(prm["beam_centre_x"][p] - spec.beam_centre_x[p]) < 5e-6
        assert abs(prm["beam_centre_y"][p] - spec.beam_centre_y[p]) < 5e-6
        assert abs(prm["waist"][p] / spec.waist - 1) < 0.03
        assert abs(prm["defocus"][p] - spec.defocus[p]) < 0.1
