"""ctypes binding of libdigholo.so with the digHolo C API names.

Every function declared in include/digHolo.h is available as an attribute of
the object returned by :func:`load` (e.g. ``lib.digHoloCreate()``), with the
argument and return types of the header.  Helpers at the end wrap the pointer
returns into numpy arrays for tests and benchmarks.

There is no fallback: if the library has not been built this raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdigholo.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "digHolo.h")

c_int, c_float, c_void_p = C.c_int, C.c_float, C.c_void_p
P_int = C.POINTER(C.c_int)
P_float = C.POINTER(C.c_float)
PP_float = C.POINTER(P_float)
P_u16 = C.POINTER(C.c_ushort)

# name -> (restype, argtypes)
SIGNATURES: Dict[str, Tuple[object, List[object]]] = {
    "digHoloCreate": (c_int, []),
    "digHoloDestroy": (c_int, [c_int]),
    "digHoloSetFrameBuffer": (c_int, [c_int, c_void_p]),
    "digHoloGetFrameBuffer": (c_void_p, [c_int]),
    "digHoloSetFrameBufferUint16": (c_int, [c_int, c_void_p, c_int]),
    "digHoloGetFrameBufferUint16": (c_void_p, [c_int, P_int]),
    "digHoloConfigSetFrameDimensions": (c_int, [c_int, c_int, c_int]),
    "digHoloConfigGetFrameDimensions": (c_int, [c_int, P_int, P_int]),
    "digHoloConfigSetFramePixelSize": (c_int, [c_int, c_float]),
    "digHoloConfigGetFramePixelSize": (c_float, [c_int]),
    "digHoloConfigSetPolCount": (c_int, [c_int, c_int]),
    "digHoloConfigGetPolCount": (c_int, [c_int]),
    "digHoloConfigSetFillFactorCorrectionEnabled": (c_int, [c_int, c_int]),
    "digHoloConfigGetFillFactorCorrectionEnabled": (c_int, [c_int]),
    "digHoloSetBatch": (c_int, [c_int, c_int, c_void_p]),
    "digHoloSetBatchAvg": (c_int, [c_int, c_int, c_void_p, c_int, c_int]),
    "digHoloSetBatchUint16": (c_int, [c_int, c_int, c_void_p, c_int]),
    "digHoloSetBatchAvgUint16": (c_int, [c_int, c_int, c_void_p, c_int, c_int, c_int]),
    "digHoloConfigSetBatchCount": (c_int, [c_int, c_int]),
    "digHoloConfigSetBatchCalibration": (c_int, [c_int, c_void_p, c_int, c_int]),
    "digHoloConfigGetBatchCalibration": (c_void_p, [c_int, P_int, P_int]),
    "digHoloConfigSetBatchCalibrationFromFile": (c_int, [c_int, C.c_char_p, c_int, c_int]),
    "digHoloConfigSetRefCalibrationIntensity": (c_int, [c_int, c_void_p, c_int, c_int, c_int]),
    "digHoloConfigSetRefCalibrationField": (c_int, [c_int, c_void_p, c_int, c_int, c_int]),
    "digHoloConfigSetRefCalibrationFromFile": (c_int, [c_int, C.c_char_p, c_int, c_int, c_int]),
    "digHoloConfigSetRefCalibrationEnabled": (c_int, [c_int, c_int]),
    "digHoloConfigGetRefCalibrationEnabled": (c_int, [c_int]),
    "digHoloConfigGetRefCalibrationFields": (c_void_p, [c_int, P_int, P_int, PP_float, PP_float, P_int, P_int]),
    "digHoloConfigSetBatchCalibrationEnabled": (c_int, [c_int, c_int]),
    "digHoloConfigGetBatchCalibrationEnabled": (c_int, [c_int]),
    "digHoloConfigGetBatchCount": (c_int, [c_int]),
    "digHoloConfigSetBatchAvgCount": (c_int, [c_int, c_int]),
    "digHoloConfigGetBatchAvgCount": (c_int, [c_int]),
    "digHoloConfigSetBatchAvgMode": (c_int, [c_int, c_int]),
    "digHoloConfigGetBatchAvgMode": (c_int, [c_int]),
    "digHoloConfigSetfftWindowSize": (c_int, [c_int, c_int, c_int]),
    "digHoloConfigGetfftWindowSize": (c_int, [c_int, P_int, P_int]),
    "digHoloConfigSetfftWindowSizeX": (c_int, [c_int, c_int]),
    "digHoloConfigGetfftWindowSizeX": (c_int, [c_int]),
    "digHoloConfigGetfftWindowSizeY": (c_int, [c_int]),
    "digHoloConfigSetfftWindowSizeY": (c_int, [c_int, c_int]),
    "digHoloConfigSetFourierWindowRadius": (c_int, [c_int, c_float]),
    "digHoloConfigGetFourierWindowRadius": (c_float, [c_int]),
    "digHoloConfigSetIFFTResolutionMode": (c_int, [c_int, c_int]),
    "digHoloConfigGetIFFTResolutionMode": (c_int, [c_int]),
    "digHoloConfigSetWavelengthCentre": (c_int, [c_int, c_float]),
    "digHoloConfigGetWavelengthCentre": (c_float, [c_int]),
    "digHoloConfigSetWavelengths": (c_int, [c_int, P_float, c_int]),
    "digHoloConfigSetWavelengthsLinearFrequency": (c_int, [c_int, c_float, c_float, c_int]),
    "digHoloConfigGetWavelengths": (P_float, [c_int, P_int]),
    "digHoloConfigSetWavelengthOrdering": (c_int, [c_int, c_int, c_int]),
    "digHoloConfigGetWavelengthOrdering": (c_int, [c_int, c_int]),
    "digHoloProcessBatchFrequencySweepLinear": (c_void_p, [c_int, P_int, P_int, P_int, c_float, c_float, c_int]),
    "digHoloProcessBatchWavelengthSweepArbitrary": (c_void_p, [c_int, P_int, P_int, P_int, P_float, c_int]),
    "digHoloConfigSetAutoAlignBeamCentre": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignBeamCentre": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignDefocus": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignDefocus": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignTilt": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignTilt": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignBasisWaist": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignBasisWaist": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignFourierWindowRadius": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignFourierWindowRadius": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignTol": (c_int, [c_int, c_float]),
    "digHoloConfigGetAutoAlignTol": (c_float, [c_int]),
    "digHoloConfigSetAutoAlignPolIndependence": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignPolIndependence": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignBasisMulConjTrans": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignBasisMulConjTrans": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignMode": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignMode": (c_int, [c_int]),
    "digHoloConfigSetAutoAlignGoalIdx": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoAlignGoalIdx": (c_int, [c_int]),
    "digHoloAutoAlign": (c_float, [c_int]),
    "digHoloGetFields16": (c_int, [c_int, P_int, P_int, C.POINTER(C.POINTER(C.c_short)),
                                   C.POINTER(C.POINTER(C.c_short)), C.POINTER(P_float), C.POINTER(P_float),
                                   C.POINTER(P_float), P_int, P_int]),
    "digHoloBatchGetSummary": (c_int, [c_int, c_int, P_int, P_int, P_int, C.POINTER(P_float), P_int, P_int,
                                       C.POINTER(P_float), C.POINTER(P_float), C.POINTER(P_float)]),
    "digHoloAutoAlignGetMetric": (c_float, [c_int, c_int]),
    "digHoloAutoAlignGetMetrics": (P_float, [c_int, c_int]),
    "digHoloAutoAlignCalcMetrics": (c_int, [c_int]),
    "digHoloConfigSetBeamCentre": (c_int, [c_int, c_int, c_int, c_float]),
    "digHoloConfigGetBeamCentre": (c_float, [c_int, c_int, c_int]),
    "digHoloConfigSetTilt": (c_int, [c_int, c_int, c_int, c_float]),
    "digHoloConfigGetTilt": (c_float, [c_int, c_int, c_int]),
    "digHoloConfigSetDefocus": (c_int, [c_int, c_int, c_float]),
    "digHoloConfigGetDefocus": (c_float, [c_int, c_int]),
    "digHoloConfigSetPolLockTilt": (c_int, [c_int, c_int]),
    "digHoloConfigGetPolLockTilt": (c_int, [c_int]),
    "digHoloConfigSetPolLockDefocus": (c_int, [c_int, c_int]),
    "digHoloConfigGetPolLockDefocus": (c_int, [c_int]),
    "digHoloConfigSetPolLockBasisWaist": (c_int, [c_int, c_int]),
    "digHoloConfigGetPolLockBasisWaist": (c_int, [c_int]),
    "digHoloConfigSetBasisGroupCount": (c_int, [c_int, c_int]),
    "digHoloConfigGetBasisGroupCount": (c_int, [c_int]),
    "digHoloConfigSetBasisWaist": (c_int, [c_int, c_int, c_float]),
    "digHoloConfigGetBasisWaist": (c_float, [c_int, c_int]),
    "digHoloConfigSetBasisTypeHG": (c_int, [c_int]),
    "digHoloConfigSetBasisTypeLG": (c_int, [c_int]),
    "digHoloConfigSetBasisTypeCustom": (c_int, [c_int, c_int, c_int, c_void_p]),
    "digHoloBasisGetFields": (c_void_p, [c_int, P_int, P_int, C.POINTER(P_float), C.POINTER(P_float), P_int, P_int]),
    "digHoloConfigSetBasisType": (c_int, [c_int, c_int]),
    "digHoloConfigGetBasisType": (c_int, [c_int]),
    "digHoloProcessBatch": (c_void_p, [c_int, P_int, P_int, P_int]),
    "digHoloGetFields": (c_void_p, [c_int, P_int, P_int, PP_float, PP_float, P_int, P_int]),
    "digHoloBasisGetCoefs": (c_void_p, [c_int, P_int, P_int, P_int]),
    "digHoloProcessFFT": (c_int, [c_int]),
    "digHoloProcessIFFT": (c_int, [c_int]),
    "digHoloProcessRemoveTilt": (c_int, [c_int]),
    "digHoloProcessBasisExtractCoefs": (c_int, [c_int]),
    "digHoloConfigSetThreadCount": (c_int, [c_int, c_int]),
    "digHoloConfigGetThreadCount": (c_int, [c_int]),
    "digHoloBenchmark": (c_float, [c_int, c_float, P_float]),
    "digHoloConfigSetVerbosity": (c_int, [c_int, c_int]),
    "digHoloConfigGetVerbosity": (c_int, [c_int]),
    # extensions
    "digHoloConfigSetOverlapPrecision": (c_int, [c_int, c_int]),
    "digHoloConfigGetOverlapPrecision": (c_int, [c_int]),
    "digHoloProcessBatchDevice": (c_void_p, [c_int, P_int, P_int, P_int]),
    "digHoloProcessBatchDeviceNoAlloc": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "digHoloGetFieldsDevice": (c_void_p, [c_int, P_int, P_int, P_int, P_int]),
    "digHoloSynchronize": (c_int, [c_int]),
    "digHoloSetStream": (c_int, [c_int, c_void_p]),
    "digHoloGetStream": (c_void_p, [c_int]),
    "digHoloGetWindowGeometry": (c_int, [c_int, P_int]),
    "digHoloConfigSetFusedWindow": (c_int, [c_int, c_int]),
    "digHoloConfigSetFusedOverlap": (c_int, [c_int, c_int]),
    "digHoloConfigSetWindowRowTransform": (c_int, [c_int, c_int]),
    "digHoloConfigGetWindowRowTransform": (c_int, [c_int]),
    "digHoloGetWindowPath": (c_int, [c_int]),
    "digHoloConfigSetStageTiming": (c_int, [c_int, c_int]),
    "digHoloGetStageTimes": (c_int, [c_int, P_float]),
    "digHoloConfigSetAutoTune": (c_int, [c_int, c_int]),
    "digHoloConfigGetAutoTune": (c_int, [c_int]),
    "digHoloProcessBatchMultiDevice": (c_int, [c_int, P_int]),
    "digHoloAutoAlignGradient": (c_int, [c_int, c_int, c_float, P_float, P_float]),
    "digHoloAutoAlignGradientSearch": (c_float, [c_int, c_int, c_float]),
}

# constants of the header
DIGHOLO_ERROR_SUCCESS = 0
DIGHOLO_ERROR_ERROR = 1
DIGHOLO_ERROR_INVALIDHANDLE = 2
DIGHOLO_ERROR_NULLPOINTER = 3
DIGHOLO_ERROR_INVALIDDIMENSION = 5
DIGHOLO_ERROR_INVALIDPOLARISATION = 6
DIGHOLO_ERROR_INVALIDAXIS = 7
DIGHOLO_ERROR_INVALIDARGUMENT = 8
DIGHOLO_ERROR_MEMORYALLOCATION = 9
DIGHOLO_ERROR_FILENOTCREATED = 10
DIGHOLO_ERROR_FILENOTFOUND = 11
DIGHOLO_AUTOALIGNMODE_FULL = 0
DIGHOLO_AUTOALIGNMODE_TWEAK = 1
DIGHOLO_AUTOALIGNMODE_ESTIMATE = 2
DIGHOLO_METRIC_COUNT = 9
DIGHOLO_ANALYSIS_COUNT = 7
DIGHOLO_ANALYSIS_MAXABS = 3
DIGHOLO_ANALYSIS_MAXABSIDX = 4
DIGHOLO_WAVELENGTHORDER_INPUT = 0
DIGHOLO_WAVELENGTHORDER_OUTPUT = 1
DIGHOLO_WAVELENGTHORDER_FAST = 0
DIGHOLO_WAVELENGTHORDER_SLOW = 1
DIGHOLO_PRECISION_FP32 = 0
DIGHOLO_PRECISION_SPLIT16 = 1

_lib: Optional[C.CDLL] = None


def declared_symbols() -> List[str]:
    """Every function name declared with EXT_C in include/digHolo.h."""
    text = open(HEADER).read()
    return re.findall(r"^EXT_C\s+[\w\s\*]+?\b(digHolo\w+)\s*\(", text, flags=re.M)


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python -m digholo_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# ---------------------------------------------------------------------------
# numpy helpers
# ---------------------------------------------------------------------------

def configure(lib, h, *, frame_width, frame_height, pol_count=1, pixel_size=20e-6, fft_window=(0, 0),
              aperture_deg=0.7, wavelength=1565e-9, beam_centre_x=(0.0, 0.0), beam_centre_y=(0.0, 0.0),
              tilt_x=(1.5, 1.5), tilt_y=(1.5, 1.5), defocus=(0.0, 0.0), waist=200e-6, group_count=10,
              fill_factor=1, resolution_mode=1, pol_lock_tilt=0, pol_lock_defocus=0, pol_lock_waist=1):
    """Apply a configuration through the C API setters; raises on any error code."""
    def ok(rc, what):
        if rc != DIGHOLO_ERROR_SUCCESS:
            raise ValueError(f"{what} returned {rc}")
    ok(lib.digHoloConfigSetFrameDimensions(h, frame_width, frame_height), "SetFrameDimensions")
    ok(lib.digHoloConfigSetFramePixelSize(h, pixel_size), "SetFramePixelSize")
    ok(lib.digHoloConfigSetPolCount(h, pol_count), "SetPolCount")
    ok(lib.digHoloConfigSetfftWindowSize(h, fft_window[0], fft_window[1]), "SetfftWindowSize")
    ok(lib.digHoloConfigSetFourierWindowRadius(h, aperture_deg), "SetFourierWindowRadius")
    ok(lib.digHoloConfigSetWavelengthCentre(h, wavelength), "SetWavelengthCentre")
    ok(lib.digHoloConfigSetIFFTResolutionMode(h, resolution_mode), "SetIFFTResolutionMode")
    ok(lib.digHoloConfigSetFillFactorCorrectionEnabled(h, fill_factor), "SetFillFactor")
    ok(lib.digHoloConfigSetPolLockTilt(h, pol_lock_tilt), "SetPolLockTilt")
    ok(lib.digHoloConfigSetPolLockDefocus(h, pol_lock_defocus), "SetPolLockDefocus")
    ok(lib.digHoloConfigSetPolLockBasisWaist(h, pol_lock_waist), "SetPolLockBasisWaist")
    for p in range(pol_count):
        ok(lib.digHoloConfigSetBeamCentre(h, 0, p, beam_centre_x[p]), "SetBeamCentre")
        ok(lib.digHoloConfigSetBeamCentre(h, 1, p, beam_centre_y[p]), "SetBeamCentre")
        ok(lib.digHoloConfigSetTilt(h, 0, p, tilt_x[p]), "SetTilt")
        ok(lib.digHoloConfigSetTilt(h, 1, p, tilt_y[p]), "SetTilt")
        ok(lib.digHoloConfigSetDefocus(h, p, defocus[p]), "SetDefocus")
    ok(lib.digHoloConfigSetBasisWaist(h, 0, waist), "SetBasisWaist")
    ok(lib.digHoloConfigSetBasisGroupCount(h, group_count), "SetBasisGroupCount")


def window_geometry(lib, h) -> Dict[str, object]:
    out = (C.c_int * 12)()
    rc = lib.digHoloGetWindowGeometry(h, out)
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise ValueError(f"GetWindowGeometry returned {rc}")
    v = list(out)
    return dict(W=v[0], H=v[1], w=v[2], h=v[3], cx=[v[4], v[6]], cy=[v[5], v[7]], SX=[v[8], v[10]],
                SY=[v[9], v[11]])


def get_coefs(lib, h):
    """digHoloBasisGetCoefs as a [B, pol, M] complex64 array (None when there is no basis)."""
    b, m, p = c_int(), c_int(), c_int()
    ptr = lib.digHoloBasisGetCoefs(h, C.byref(b), C.byref(m), C.byref(p))
    if not ptr:
        return None
    out = np.ctypeslib.as_array(C.cast(ptr, P_float), shape=(b.value * p.value * m.value * 2,))
    return out.view(np.complex64).reshape(b.value, p.value, m.value).copy()


def get_fields(lib, h):
    """digHoloGetFields: (fields [B, pol, h, w] complex64, x, y)."""
    fb, fp, w, hh = c_int(), c_int(), c_int(), c_int()
    xp, yp = P_float(), P_float()
    fptr = lib.digHoloGetFields(h, C.byref(fb), C.byref(fp), C.byref(xp), C.byref(yp), C.byref(w), C.byref(hh))
    if not fptr:
        raise RuntimeError("digHoloGetFields failed")
    n = fb.value * fp.value * hh.value * w.value
    fields = np.ctypeslib.as_array(C.cast(fptr, P_float), shape=(2 * n,)).view(np.complex64)
    fields = fields.reshape(fb.value, fp.value, hh.value, w.value).copy()
    x = np.ctypeslib.as_array(xp, shape=(w.value,)).copy()
    y = np.ctypeslib.as_array(yp, shape=(hh.value,)).copy()
    return fields, x, y


def process_batch(lib, h, frames: np.ndarray):
    """SetBatch(Uint16) + ProcessBatch + GetFields on host numpy frames.

    Returns (coefs [B, pol, M] complex64, fields [B, pol, h, w] complex64, x, y).
    """
    frames = np.ascontiguousarray(frames)
    B = frames.shape[0]
    if frames.dtype == np.uint16:
        rc = lib.digHoloSetBatchUint16(h, B, frames.ctypes.data, 0)
    else:
        frames = np.ascontiguousarray(frames, dtype=np.float32)
        rc = lib.digHoloSetBatch(h, B, frames.ctypes.data)
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise ValueError(f"SetBatch returned {rc}")
    b, m, p = c_int(), c_int(), c_int()
    ptr = lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p))
    if b.value == 0:
        raise RuntimeError("digHoloProcessBatch failed")
    coefs = None
    if ptr:
        coefs = np.ctypeslib.as_array(C.cast(ptr, P_float), shape=(b.value * p.value * m.value * 2,))
        coefs = coefs.view(np.complex64).reshape(b.value, p.value, m.value).copy()
    fb, fp, w, hh = c_int(), c_int(), c_int(), c_int()
    xp, yp = P_float(), P_float()
    fptr = lib.digHoloGetFields(h, C.byref(fb), C.byref(fp), C.byref(xp), C.byref(yp), C.byref(w), C.byref(hh))
    if not fptr:
        raise RuntimeError("digHoloGetFields failed")
    n = fb.value * fp.value * hh.value * w.value
    fields = np.ctypeslib.as_array(C.cast(fptr, P_float), shape=(2 * n,)).view(np.complex64)
    fields = fields.reshape(fb.value, fp.value, hh.value, w.value).copy()
    x = np.ctypeslib.as_array(xp, shape=(w.value,)).copy()
    y = np.ctypeslib.as_array(yp, shape=(hh.value,)).copy()
    return coefs, fields, x, y


def set_wavelengths(lib, h, wavelengths) -> None:
    """digHoloConfigSetWavelengths from a sequence of wavelengths (m)."""
    lam = np.ascontiguousarray(wavelengths, dtype=np.float32)
    rc = lib.digHoloConfigSetWavelengths(h, lam.ctypes.data_as(P_float), lam.size)
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise ValueError(f"digHoloConfigSetWavelengths returned {rc}")


def get_wavelengths(lib, h) -> np.ndarray:
    n = c_int()
    ptr = lib.digHoloConfigGetWavelengths(h, C.byref(n))
    if not ptr or n.value <= 0:
        raise RuntimeError("digHoloConfigGetWavelengths failed")
    return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()


def get_ref_calibration_fields(lib, h):
    """digHoloConfigGetRefCalibrationFields -> ([L, pol, h, w] complex64, x, y) or None."""
    L, P, w, hh = c_int(), c_int(), c_int(), c_int()
    xp, yp = P_float(), P_float()
    ptr = lib.digHoloConfigGetRefCalibrationFields(h, C.byref(L), C.byref(P), C.byref(xp), C.byref(yp), C.byref(w),
                                                   C.byref(hh))
    if not ptr:
        return None
    n = L.value * P.value * hh.value * w.value
    f = np.ctypeslib.as_array(C.cast(ptr, P_float), shape=(2 * n,)).view(np.complex64)
    f = f.reshape(L.value, P.value, hh.value, w.value).copy()
    x = np.ctypeslib.as_array(xp, shape=(w.value,)).copy()
    y = np.ctypeslib.as_array(yp, shape=(hh.value,)).copy()
    return f, x, y


def batch_summary(lib, h, plane: int):
    """digHoloBatchGetSummary as numpy arrays.

    Returns (params [7, pol, B + 1] float32, peak_index [pol, B + 1] int32,
    summed [pol, Ny, Nx] float32, x [Nx], y [Ny]).
    """
    pc, tc, pol, nx, ny = c_int(), c_int(), c_int(), c_int(), c_int()
    par, xp, yp, tot = P_float(), P_float(), P_float(), P_float()
    rc = lib.digHoloBatchGetSummary(h, plane, C.byref(pc), C.byref(tc), C.byref(pol), C.byref(par), C.byref(nx),
                                    C.byref(ny), C.byref(xp), C.byref(yp), C.byref(tot))
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise RuntimeError(f"digHoloBatchGetSummary returned {rc}")
    P, T, Nx, Ny = pol.value, tc.value, nx.value, ny.value
    flat = np.ctypeslib.as_array(par, shape=(pc.value * P * T,)).copy()
    params = flat.reshape(pc.value, P, T)
    peak_index = flat.view(np.int32).reshape(pc.value, P, T)[DIGHOLO_ANALYSIS_MAXABSIDX].copy()
    # the summed maps are [pol][W * H] in the reference's layout; Ny * Nx used per pol
    geo = window_geometry(lib, h)
    stride = geo["W"] * geo["H"]
    maps = np.ctypeslib.as_array(tot, shape=(P * stride,)).copy().reshape(P, stride)[:, :Ny * Nx]
    x = np.ctypeslib.as_array(xp, shape=(Nx,)).copy()
    y = np.ctypeslib.as_array(yp, shape=(Ny,)).copy()
    return params, peak_index, maps.reshape(P, Ny, Nx), x, y


def get_fields16(lib, h):
    """digHoloGetFields16: (re [B, pol, h, w] int16, im, scale [pol, B] float32, x, y)."""
    b, p, w, hh = c_int(), c_int(), c_int(), c_int()
    re, im = C.POINTER(C.c_short)(), C.POINTER(C.c_short)()
    sc, xp, yp = P_float(), P_float(), P_float()
    rc = lib.digHoloGetFields16(h, C.byref(b), C.byref(p), C.byref(re), C.byref(im), C.byref(sc), C.byref(xp),
                                C.byref(yp), C.byref(w), C.byref(hh))
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise RuntimeError(f"digHoloGetFields16 returned {rc}")
    B, P, W, H = b.value, p.value, w.value, hh.value
    n = B * P * H * W
    R = np.ctypeslib.as_array(re, shape=(n,)).copy().reshape(B, P, H, W)
    I = np.ctypeslib.as_array(im, shape=(n,)).copy().reshape(B, P, H, W)
    scale = np.ctypeslib.as_array(sc, shape=(P * (B + 1),)).copy().reshape(P, B + 1)[:, :B]
    x = np.ctypeslib.as_array(xp, shape=(W,)).copy()
    y = np.ctypeslib.as_array(yp, shape=(H,)).copy()
    return R, I, scale, x, y


GRAD_PARAMS = ("waist", "defocus", "tilt_x", "beam_centre_x", "tilt_y", "beam_centre_y")


def set_align_params(lib, h, params) -> None:
    """Write a [pol, 6] array of alignment parameters (the AutoAlign order of
    GRAD_PARAMS, API units) through the C API setters."""
    params = np.asarray(params, dtype=np.float64)
    for p in range(params.shape[0]):
        w, d, tx, cx, ty, cy = (float(v) for v in params[p])
        lib.digHoloConfigSetBasisWaist(h, p, w)
        lib.digHoloConfigSetDefocus(h, p, d)
        lib.digHoloConfigSetTilt(h, 0, p, tx)
        lib.digHoloConfigSetBeamCentre(h, 0, p, cx)
        lib.digHoloConfigSetTilt(h, 1, p, ty)
        lib.digHoloConfigSetBeamCentre(h, 1, p, cy)


def get_align_params(lib, h, P: int) -> np.ndarray:
    return np.array([[lib.digHoloConfigGetBasisWaist(h, p), lib.digHoloConfigGetDefocus(h, p),
                      lib.digHoloConfigGetTilt(h, 0, p), lib.digHoloConfigGetBeamCentre(h, 0, p),
                      lib.digHoloConfigGetTilt(h, 1, p), lib.digHoloConfigGetBeamCentre(h, 1, p)]
                     for p in range(P)], dtype=np.float64)


def alignment_gradient(lib, h, hold: bool = False, mdl_weight: float = 0.0):
    """digHoloAutoAlignGradient: (loss, grad [2, 6] float64)."""
    loss = c_float()
    grad = (c_float * 12)()
    rc = lib.digHoloAutoAlignGradient(h, 1 if hold else 0, mdl_weight, C.byref(loss), grad)
    if rc != DIGHOLO_ERROR_SUCCESS:
        raise RuntimeError(f"digHoloAutoAlignGradient returned {rc}")
    return float(loss.value), np.array(list(grad), dtype=np.float64).reshape(2, 6)
