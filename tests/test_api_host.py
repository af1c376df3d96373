"""Host-side tests of libdigholo.so that need no GPU.

Covers: every symbol of include/digHolo.h is exported with C linkage; setter
validation and return conventions match the reference wrappers
(digHolo.cpp:13413-15415); window geometry computed by the C++ host code equals
the oracle's float32 restatement over a randomised sweep modelled on the
reference's stressTest (digHolo.cpp:4593-4633).
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

from digholo_amd import api
from oracle import digholo_oracle as O


@pytest.fixture(scope="module")
def lib():
    return api.load()


def test_exports_every_declared_symbol(lib):
    syms = api.declared_symbols()
    assert len(syms) >= 90
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        assert hasattr(lib, s)
    assert set(syms) == set(api.SIGNATURES)


def test_invalid_handle_conventions(lib):
    assert lib.digHoloConfigSetPolCount(12345, 1) == api.DIGHOLO_ERROR_INVALIDHANDLE
    assert lib.digHoloConfigGetPolCount(12345) == 0
    b, m, p = C.c_int(7), C.c_int(7), C.c_int(7)
    assert not lib.digHoloProcessBatch(-1, C.byref(b), C.byref(m), C.byref(p))
    assert (b.value, m.value, p.value) == (0, 0, 0)      # out-params zeroed (digHolo.cpp:15125-15131)
    h = lib.digHoloCreate()
    h2 = lib.digHoloCreate()
    assert h2 == h + 1
    assert lib.digHoloDestroy(h) == 0
    assert lib.digHoloConfigSetPolCount(h, 1) == api.DIGHOLO_ERROR_INVALIDHANDLE
    assert lib.digHoloCreate() == h2 + 1                  # handles are never reused
    # digHoloBenchmark: 0 for an invalid handle or a non-positive duration
    # (digHolo.cpp:13248, :13404), before touching any device
    info = (C.c_float * 7)(*([9.0] * 7))
    assert lib.digHoloBenchmark(-1, C.c_float(1.0), info) == 0.0
    assert lib.digHoloBenchmark(h2, C.c_float(0.0), info) == 0.0
    assert list(info) == [9.0] * 7


def test_setter_validation(lib):
    h = lib.digHoloCreate()
    E = api
    assert lib.digHoloConfigSetFrameDimensions(h, 330, 256) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetFrameDimensions(h, 640, 512) == 0
    w, hh = C.c_int(), C.c_int()
    lib.digHoloConfigGetFrameDimensions(h, C.byref(w), C.byref(hh))
    assert (w.value, hh.value) == (640, 512)
    assert lib.digHoloConfigSetFramePixelSize(h, -1.0) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetFramePixelSize(h, float("nan")) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetPolCount(h, 3) == E.DIGHOLO_ERROR_INVALIDPOLARISATION
    assert lib.digHoloConfigSetPolCount(h, 2) == 0
    assert lib.digHoloConfigSetfftWindowSizeX(h, 200) == 192      # returns the floored value
    assert lib.digHoloConfigGetfftWindowSizeX(h) == 192
    assert lib.digHoloConfigSetfftWindowSize(h, -1, 16) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetFourierWindowRadius(h, 0.0) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetIFFTResolutionMode(h, 2) == E.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetWavelengthCentre(h, 0.0) == E.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetBeamCentre(h, 2, 0, 1e-3) == E.DIGHOLO_ERROR_INVALIDAXIS
    assert lib.digHoloConfigSetBeamCentre(h, 0, 2, 1e-3) == E.DIGHOLO_ERROR_INVALIDPOLARISATION
    assert lib.digHoloConfigSetBeamCentre(h, 0, 1, float("inf")) == E.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetTilt(h, 1, 1, 1.25) == 0
    assert abs(lib.digHoloConfigGetTilt(h, 1, 1) - 1.25) < 1e-7
    assert lib.digHoloConfigSetBasisGroupCount(h, -7) == 0
    assert lib.digHoloConfigGetBasisGroupCount(h) == 7            # abs() like the reference
    # SetBasisWaist writes every pol regardless of polIdx (digHolo.cpp:14058-14063)
    assert lib.digHoloConfigSetBasisWaist(h, 1, 250e-6) == 0
    assert abs(lib.digHoloConfigGetBasisWaist(h, 0) - 250e-6) < 1e-10
    assert lib.digHoloConfigSetBasisWaist(h, 0, -1.0) == E.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetAutoAlignBeamCentre(h, 5) == 0
    assert lib.digHoloConfigGetAutoAlignBeamCentre(h) == 1
    assert lib.digHoloConfigSetAutoAlignMode(h, 3) == E.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetOverlapPrecision(h, 7) == E.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloSetBatch(h, 10, None) == E.DIGHOLO_ERROR_NULLPOINTER
    buf = np.zeros(640 * 512, dtype=np.float32)
    assert lib.digHoloSetBatch(h, 0, buf.ctypes.data) == 0
    assert lib.digHoloConfigGetBatchCount(h) == 1                  # batchCount < 1 -> 1
    # defaults of the reference config (digHolo.cpp:3327-3425)
    h3 = lib.digHoloCreate()
    assert lib.digHoloConfigGetIFFTResolutionMode(h3) == 0
    assert lib.digHoloConfigGetPolLockBasisWaist(h3) == 1
    assert abs(lib.digHoloConfigGetFourierWindowRadius(h3) - 0.4) < 1e-7
    assert abs(lib.digHoloConfigGetTilt(h3, 0, 0) + 0.75) < 1e-7


def oracle_cfg(d):
    # the setter floors the window to a multiple of 16 (digHolo.cpp:14280-14300)
    return O.Config(frame_width=d["fw"], frame_height=d["fh"], pol_count=d["P"], pixel_size=d["px"],
                    fft_window_x=16 * (d["win"][0] // 16), fft_window_y=16 * (d["win"][1] // 16),
                    aperture_deg=d["r"],
                    wavelength=d["lam"], beam_centre_x=list(d["bcx"]), beam_centre_y=list(d["bcy"]),
                    tilt_x=list(d["tx"]), tilt_y=list(d["ty"]), group_count=d["G"],
                    pol_lock_tilt=d["lock"])


def random_case(rng):
    P = int(rng.integers(1, 3))
    fw = int(rng.choice([128, 256, 320, 640, 1024]))
    fh = int(rng.choice([128, 256, 512, 1024]))
    win = (int(rng.integers(0, fw // P + 1)), int(rng.integers(0, fh + 1)))
    return dict(P=P, fw=fw, fh=fh, px=float(rng.uniform(3e-6, 30e-6)), win=win,
                r=float(rng.uniform(0.1, 1.5)), lam=float(rng.uniform(0.5e-6, 2e-6)),
                bcx=tuple(rng.uniform(-3e-3, 3e-3, 2)), bcy=tuple(rng.uniform(-3e-3, 3e-3, 2)),
                tx=tuple(rng.uniform(-4, 4, 2)), ty=tuple(rng.uniform(-4, 4, 2)),
                G=int(rng.integers(0, 20)), lock=int(rng.integers(0, 2)))


def test_window_geometry_matches_oracle_fuzz(lib):
    rng = np.random.default_rng(2024)
    checked = 0
    for _ in range(400):
        d = random_case(rng)
        h = lib.digHoloCreate()
        try:
            api.configure(lib, h, frame_width=d["fw"], frame_height=d["fh"], pol_count=d["P"], pixel_size=d["px"],
                          fft_window=d["win"], aperture_deg=d["r"], wavelength=d["lam"],
                          beam_centre_x=d["bcx"], beam_centre_y=d["bcy"], tilt_x=d["tx"], tilt_y=d["ty"],
                          group_count=d["G"], pol_lock_tilt=d["lock"])
        except ValueError:
            continue
        cfg = oracle_cfg(d)
        if d["lock"]:
            cfg = O.apply_pol_locks(cfg)
        plan = O.make_plan(cfg)
        out = (C.c_int * 12)()
        rc = lib.digHoloGetWindowGeometry(h, out)
        if plan.w > plan.W // 2:
            assert rc == api.DIGHOLO_ERROR_INVALIDDIMENSION      # window wider than the r2c half plane
            continue
        assert rc == 0
        geo = api.window_geometry(lib, h)
        assert (geo["W"], geo["H"], geo["w"], geo["h"]) == (plan.W, plan.H, plan.w, plan.h), d
        for p in range(d["P"]):
            assert geo["cx"][p] == plan.cx[p] and geo["cy"][p] == plan.cy[p], d
            assert geo["SX"][p] == plan.SX[p] and geo["SY"][p] == plan.SY[p], d
        checked += 1
        lib.digHoloDestroy(h)
    assert checked > 150


def test_process_without_gpu_fails_loudly(lib, gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=320, frame_height=256, fft_window=(128, 128), group_count=4)
    frames = np.zeros((2, 256, 320), dtype=np.uint16)
    with pytest.raises(RuntimeError):
        api.process_batch(lib, h, frames)


def test_autoalign_modes_without_estimate_stage(lib):
    fmax = np.finfo(np.float32).max
    assert lib.digHoloAutoAlign(12345) == 0.0                       # invalid handle (digHolo.cpp:14989-14999)
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=320, frame_height=256, fft_window=(128, 128), group_count=4)
    frames = np.zeros((2, 256, 320), dtype=np.uint16)
    assert lib.digHoloSetBatchUint16(h, 2, frames.ctypes.data_as(C.POINTER(C.c_ushort)), 0) == 0
    # FULL (the default) and ESTIMATE need the parameter-estimate stage: refused before any GPU work
    assert lib.digHoloConfigGetAutoAlignMode(h) == api.DIGHOLO_AUTOALIGNMODE_FULL
    assert lib.digHoloAutoAlign(h) == -fmax
    assert lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_ESTIMATE) == 0
    assert lib.digHoloAutoAlign(h) == -fmax
    # TWEAK with no modes (group count 0) needs it too
    assert lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_TWEAK) == 0
    assert lib.digHoloConfigSetBasisGroupCount(h, 0) == 0
    assert lib.digHoloAutoAlign(h) == -fmax
    lib.digHoloDestroy(h)


# ---- wavelength sweeps (digHolo.cpp:6577-6680, :14611-14700) ----

def test_wavelength_axis_setters_match_oracle(lib):
    h = lib.digHoloCreate()
    rng = np.random.default_rng(7)
    for _ in range(40):
        a, b = (float(v) for v in rng.uniform(0.6e-6, 1.9e-6, 2))
        n = int(rng.integers(2, 200))
        assert lib.digHoloConfigSetWavelengthsLinearFrequency(h, a, b, n) == 0
        got = api.get_wavelengths(lib, h)
        assert np.array_equal(got, O.frequency_sweep_linear(a, b, n, 1565e-9))   # bit-exact float32
    # fewer than two entries: the centre wavelength (digHolo.cpp:7855-7863)
    assert lib.digHoloConfigSetWavelengthCentre(h, C.c_float(1.31e-6)) == 0
    assert lib.digHoloConfigSetWavelengthsLinearFrequency(h, 1.5e-6, 1.6e-6, 1) == 0
    assert np.array_equal(api.get_wavelengths(lib, h), np.array([1.31e-6], np.float32))
    lam = np.array([1.5e-6, 1.52e-6, 1.58e-6], np.float32)
    api.set_wavelengths(lib, h, lam)
    assert np.array_equal(api.get_wavelengths(lib, h), lam)
    # error conventions
    assert lib.digHoloConfigSetWavelengths(h, lam.ctypes.data_as(api.P_float), 0) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetWavelengths(h, None, 3) == api.DIGHOLO_ERROR_NULLPOINTER
    assert lib.digHoloConfigSetWavelengthsLinearFrequency(h, 1.5e-6, 1.6e-6, 0) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetWavelengthsLinearFrequency(h, 1.5e-6, 1.6e-6, -2) == api.DIGHOLO_ERROR_INVALIDDIMENSION
    assert np.array_equal(api.get_wavelengths(lib, h), lam)      # failed calls keep the axis
    for io in (0, 1):
        assert lib.digHoloConfigGetWavelengthOrdering(h, io) == 0
        assert lib.digHoloConfigSetWavelengthOrdering(h, io, 1) == 0
        assert lib.digHoloConfigGetWavelengthOrdering(h, io) == 1
    assert lib.digHoloConfigSetWavelengthOrdering(h, 2, 0) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigSetWavelengthOrdering(h, 0, 2) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    assert lib.digHoloConfigGetWavelengthOrdering(h, 5) == 0
    assert not lib.digHoloConfigGetWavelengths(10 ** 6, None)       # NULL for an invalid handle
    lib.digHoloDestroy(h)


def test_sweep_window_geometry_matches_oracle(lib):
    rng = np.random.default_rng(99)
    checked = 0
    for _ in range(150):
        d = random_case(rng)
        L = int(rng.integers(2, 9))
        lam = np.sort(rng.uniform(0.5e-6, 2e-6, L)).astype(np.float32)
        if rng.integers(0, 2):
            lam = lam[::-1].copy()
        h = lib.digHoloCreate()
        try:
            api.configure(lib, h, frame_width=d["fw"], frame_height=d["fh"], pol_count=d["P"], pixel_size=d["px"],
                          fft_window=d["win"], aperture_deg=d["r"], wavelength=d["lam"],
                          beam_centre_x=d["bcx"], beam_centre_y=d["bcy"], tilt_x=d["tx"], tilt_y=d["ty"],
                          group_count=d["G"], pol_lock_tilt=d["lock"])
        except ValueError:
            continue
        api.set_wavelengths(lib, h, lam)
        cfg = oracle_cfg(d)
        cfg.wavelengths = list(lam)
        if d["lock"]:
            cfg = O.apply_pol_locks(cfg)
        plan = O.make_plan(cfg)
        out = (C.c_int * 12)()
        rc = lib.digHoloGetWindowGeometry(h, out)
        lib.digHoloDestroy(h)
        if plan.w > plan.W // 2 or plan.h > plan.H:     # window beyond the r2c half plane
            assert rc == api.DIGHOLO_ERROR_INVALIDDIMENSION
            continue
        assert rc == 0
        assert (out[0], out[1], out[2], out[3]) == (plan.W, plan.H, plan.w, plan.h), d
        for p in range(d["P"]):
            assert (out[8 + 2 * p], out[9 + 2 * p]) == (plan.SX[p], plan.SY[p]), d
        checked += 1
    assert checked > 50


def test_slow_wavelength_ordering_needs_whole_sweeps(lib):
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=320, frame_height=256, fft_window=(128, 128))
    api.set_wavelengths(lib, h, [1.55e-6, 1.56e-6, 1.57e-6])
    out = (C.c_int * 12)()
    assert lib.digHoloConfigSetBatchCount(h, 7) == 0
    assert lib.digHoloGetWindowGeometry(h, out) == 0                # fast/fast: any batch
    assert lib.digHoloConfigSetWavelengthOrdering(h, 1, 1) == 0
    assert lib.digHoloGetWindowGeometry(h, out) == api.DIGHOLO_ERROR_INVALIDDIMENSION
    assert lib.digHoloConfigSetBatchCount(h, 9) == 0
    assert lib.digHoloGetWindowGeometry(h, out) == 0
    lib.digHoloDestroy(h)
