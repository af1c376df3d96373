"""Wavelength sweeps on the GPU path against the FP64 oracle
(digHolo.cpp:2656-2700 ordering, :6577-6680 axis, :7868-8048 window over the
axis, :8644-8670 per-wavelength aperture, :10386-10420 per-wavelength reference
wave, :10770-10790 and :12172-12191 per-wavelength metrics).  Frames are
rendered at their own wavelength; every window path, both orderings on either
side, float frames, full resolution and the pipeline steps are covered.  Bar:
1e-5 relative L2 for fields and coefficients (as the other parity tests).
"""
import ctypes as C

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O
from test_gpu_parity import TOL, oracle_for

pytestmark = pytest.mark.gpu

LAMS = [1530e-9, 1565e-9, 1600e-9]


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def sweep_case(name, B, lams, order_in, seed, quantise=True):
    if name == "small":
        spec = synth.spec_small()
        cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), waist=200e-6,
                   group_count=10)
    else:
        spec = synth.spec_dual_640()
        spec.group_count = 30
        cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), waist=300e-6,
                   group_count=30, beam_centre_x=(-1.6e-3, 1.6e-3))
    spec.defocus = [0.2, -0.1]
    cfg["defocus"] = (0.2, -0.1)
    frames, _, _, _ = synth.make_sweep_batch(spec, lams, B, seed=seed, order_in=order_in, quantise=quantise)
    if not quantise:
        frames = frames.astype(np.float32)
    return frames, cfg


def configure(lib, cfg, lams, order, window_mode=3, precision=0):
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    api.set_wavelengths(lib, h, lams)
    assert lib.digHoloConfigSetWavelengthOrdering(h, 0, order[0]) == 0
    assert lib.digHoloConfigSetWavelengthOrdering(h, 1, order[1]) == 0
    assert lib.digHoloConfigSetFusedWindow(h, window_mode) == 0
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    return h


def oracle(frames, cfg, lams, order):
    c = oracle_for(cfg)
    c.wavelengths = list(lams)
    c.wavelength_ordering = tuple(order)
    return O.process_batch(frames, c)


def compare(frames, cfg, lams, order, coefs, fields):
    f_ref, c_ref, _ = oracle(frames, cfg, lams, order)
    ef, ec = O.relative_l2(fields, f_ref), O.relative_l2(coefs, c_ref)
    print(f"fields {ef:.2e} coefs {ec:.2e}", flush=True)
    assert ef <= TOL and ec <= TOL
    return c_ref


@pytest.mark.parametrize("window_mode,order", [(3, (0, 0)), (3, (0, 1)), (3, (1, 0)), (3, (1, 1)),
                                               (2, (0, 1)), (1, (1, 0)), (0, (0, 1))])
def test_sweep_every_window_path(lib, window_mode, order):
    B = 6
    frames, cfg = sweep_case("small", B, LAMS, order[0], seed=60 + window_mode)
    h = configure(lib, cfg, LAMS, order, window_mode)
    coefs, fields, _, _ = api.process_batch(lib, h, frames)
    print(f"path {lib.digHoloGetWindowPath(h)}", flush=True)
    assert lib.digHoloGetWindowPath(h) == window_mode
    compare(frames, cfg, LAMS, order, coefs, fields)
    lib.digHoloDestroy(h)


def test_sweep_dual_pol_fused_split_overlap(lib):
    lams = [1550e-9, 1580e-9]
    frames, cfg = sweep_case("dual", 4, lams, 0, seed=70)
    h = configure(lib, cfg, lams, (0, 1), 3, precision=1)
    coefs, fields, _, _ = api.process_batch(lib, h, frames)
    assert lib.digHoloGetWindowPath(h) == 3
    compare(frames, cfg, lams, (0, 1), coefs, fields)
    lib.digHoloDestroy(h)


def test_sweep_float_frames_and_full_resolution(lib):
    frames, cfg = sweep_case("small", 6, LAMS, 1, seed=71, quantise=False)
    cfg["resolution_mode"] = 0
    h = configure(lib, cfg, LAMS, (1, 0), 3)
    coefs, fields, _, _ = api.process_batch(lib, h, frames)
    compare(frames, cfg, LAMS, (1, 0), coefs, fields)
    lib.digHoloDestroy(h)


def test_sweep_pipeline_steps(lib):
    frames, cfg = sweep_case("small", 6, LAMS, 0, seed=72)
    h = configure(lib, cfg, LAMS, (0, 1), 3)
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchUint16(h, 6, buf.ctypes.data_as(C.c_void_p), 0) == 0
    for step in (lib.digHoloProcessFFT, lib.digHoloProcessIFFT, lib.digHoloProcessRemoveTilt,
                 lib.digHoloProcessBasisExtractCoefs):
        assert step(h) == 0
    compare(buf, cfg, LAMS, (0, 1), api.get_coefs(lib, h), api.get_fields(lib, h)[0])
    lib.digHoloDestroy(h)


def test_frequency_sweep_entry_point_and_metrics(lib):
    B, L = 8, 4
    lams = O.frequency_sweep_linear(1520e-9, 1610e-9, L, 1565e-9)
    frames, cfg = sweep_case("small", B, lams, 0, seed=73)
    h = configure(lib, cfg, [1565e-9], (0, 1), 3)
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchUint16(h, B, buf.ctypes.data_as(C.c_void_p), 0) == 0
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    ptr = lib.digHoloProcessBatchFrequencySweepLinear(h, C.byref(b), C.byref(m), C.byref(p),
                                                      C.c_float(1520e-9), C.c_float(1610e-9), L)
    assert ptr and b.value == B
    assert np.array_equal(api.get_wavelengths(lib, h), lams)
    coefs = api.get_coefs(lib, h)
    c_ref = compare(buf, cfg, lams, (0, 1), coefs, api.get_fields(lib, h)[0])
    # per-wavelength metrics and their average, [metric][L + 1] in dB
    assert lib.digHoloAutoAlignCalcMetrics(h) == 0
    ref_db = O.sweep_metrics_db(O.power_meter_sweep(c_ref, L, 1))
    for metric in (O.METRIC_IL, O.METRIC_MDL, O.METRIC_DIAG):
        got = np.ctypeslib.as_array(lib.digHoloAutoAlignGetMetrics(h, metric), shape=(L + 1,))
        print(metric, got, ref_db[metric], flush=True)
        assert np.allclose(got, ref_db[metric], atol=2e-3)
    # the arbitrary-axis entry point with the same axis gives the same coefficients
    lam32 = np.ascontiguousarray(lams, dtype=np.float32)
    ptr2 = lib.digHoloProcessBatchWavelengthSweepArbitrary(h, C.byref(b), C.byref(m), C.byref(p),
                                                           lam32.ctypes.data_as(api.P_float), L)
    assert ptr2 and np.array_equal(api.get_coefs(lib, h), coefs)
    lib.digHoloDestroy(h)


def test_sweep_rejections_and_single_entry(lib):
    frames, cfg = sweep_case("small", 6, LAMS, 0, seed=74)
    buf = np.ascontiguousarray(frames)
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    h = configure(lib, cfg, LAMS, (0, 0), 3)
    # averaging with a sweep is refused: the batch fails, nothing is published
    assert lib.digHoloSetBatchAvgUint16(h, 3, buf.ctypes.data_as(C.c_void_p), 2, 0, 0) == 0
    assert not lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p)) and b.value == 0
    # a slow ordering with a batch that is not whole sweeps (5 frames, 3 wavelengths)
    assert lib.digHoloConfigSetWavelengthOrdering(h, 1, 1) == 0
    assert lib.digHoloSetBatchUint16(h, 5, buf.ctypes.data_as(C.c_void_p), 0) == 0
    assert not lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p)) and b.value == 0
    lib.digHoloDestroy(h)
    # a one-entry axis is the centre wavelength
    h = configure(lib, cfg, [1.6e-6], (0, 0), 3)
    coefs, fields, _, _ = api.process_batch(lib, h, frames)
    f_ref, c_ref, _ = O.process_batch(frames, oracle_for(cfg))
    assert O.relative_l2(coefs, c_ref) <= TOL and O.relative_l2(fields, f_ref) <= TOL
    lib.digHoloDestroy(h)
