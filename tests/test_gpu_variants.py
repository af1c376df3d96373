"""Input variants on the GPU path against the FP64 oracle: transposed uint16
frames (digHolo.cpp:1206-1229), in-phase frame averaging in the three frame
layouts (digHolo.cpp:1885-2114, :8661-8672) and batch calibration
(digHolo.cpp:1831-1840, :4469-4530), through the specialised, staged and step
paths.  Bar: 1e-5 relative L2 for fields and coefficients.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O
from test_gpu_parity import TOL, make_case, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def run(lib, h, frames, B, avg=1, mode=0, transpose=0):
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchAvgUint16(h, B, buf.ctypes.data_as(C.c_void_p), avg, mode, transpose) == 0
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    ptr = lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p))
    assert b.value == B, "ProcessBatch failed"
    coefs = api.get_coefs(lib, h) if ptr else None
    fields = api.get_fields(lib, h)[0]
    return coefs, fields, buf


def compare(frames, cfg, coefs, fields, **kw):
    f_ref, c_ref, _ = O.process_batch(frames, oracle_for(cfg), **kw)
    ef, ec = O.relative_l2(fields, f_ref), O.relative_l2(coefs, c_ref)
    print(f"fields {ef:.2e} coefs {ec:.2e}", flush=True)
    assert ef <= TOL and ec <= TOL


@pytest.mark.parametrize("mode", [3, 0])
def test_transposed_uint16_frames(lib, mode):
    # frames stored as the reference's transposed input: the W x H buffer whose
    # conversion gives the 320 x 256 frames the handle is configured for
    frames, cfg = make_case("small", 3, seed=21)
    N, H, W = frames.shape
    raw = np.empty_like(frames)
    for f in range(N):   # inverse of out[i H + j] = in[j W + i]
        raw[f].reshape(-1)[:] = frames[f].reshape(W, H).T.reshape(-1)
    assert np.array_equal(O.transpose_u16(raw), frames)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, mode) == 0
    coefs, fields, _ = run(lib, h, raw, N, transpose=1)
    compare(raw, cfg, coefs, fields, transpose=True)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("avg_mode", [0, 1, 2])
def test_averaging_layouts(lib, avg_mode):
    B, A = 2, 3
    frames, cfg = make_case("dual640", B * A, seed=30 + avg_mode)   # every frame its own noise and coefficients
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    coefs, fields, buf = run(lib, h, frames, B, avg=A, mode=avg_mode)
    assert coefs.shape[0] == B
    compare(buf, cfg, coefs, fields, avg_count=A, avg_mode=avg_mode)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("window_mode,precision", [(3, 0), (3, 1), (0, 0)])
def test_batch_calibration(lib, window_mode, precision):
    B = 3
    frames, cfg = make_case("dual640", B, seed=40)
    cal = np.array([[0.5 + 0.2j, -1.1j], [2.0, 0.3 - 0.4j]], dtype=np.complex64)   # [pol][batch], wraps
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, window_mode) == 0
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    assert lib.digHoloConfigSetBatchCalibration(h, cal.ctypes.data_as(C.c_void_p), 2, 2) == 0
    assert lib.digHoloConfigGetBatchCalibrationEnabled(h) == 1
    coefs, fields, buf = run(lib, h, frames, B)
    assert lib.digHoloGetWindowPath(h) == window_mode
    compare(buf, cfg, coefs, fields, batch_cal=cal)
    # disabling it restores the uncalibrated result
    assert lib.digHoloConfigSetBatchCalibrationEnabled(h, 0) == 0
    coefs0, fields0, _ = run(lib, h, frames, B)
    compare(buf, cfg, coefs0, fields0)
    lib.digHoloDestroy(h)


def test_batch_calibration_from_file(lib):
    cal = np.array([[1.5 - 0.5j, 0.25j, -1.0]], dtype=np.complex64)   # pol-independent, 3 batch elements
    B = 2
    frames, cfg = make_case("small", B, seed=41)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cal.bin")
        cal.tofile(path)
        assert lib.digHoloConfigSetBatchCalibrationFromFile(h, path.encode(), 1, 2) == api.DIGHOLO_ERROR_INVALIDDIMENSION
        assert lib.digHoloConfigGetBatchCalibrationEnabled(h) == 0
        assert lib.digHoloConfigSetBatchCalibrationFromFile(h, path.encode(), 1, 3) == 0
    coefs, fields, buf = run(lib, h, frames, B)
    compare(buf, cfg, coefs, fields, batch_cal=cal)
    lib.digHoloDestroy(h)


def test_steps_with_averaging_and_calibration(lib):
    B, A = 2, 2
    frames, cfg = make_case("small", B * A, seed=50)
    cal = np.array([[0.8 - 0.6j, 1.2 + 0.1j]], dtype=np.complex64)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetBatchCalibration(h, cal.ctypes.data_as(C.c_void_p), 1, 2) == 0
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchAvgUint16(h, B, buf.ctypes.data_as(C.c_void_p), A, 1, 0) == 0
    for step in (lib.digHoloProcessFFT, lib.digHoloProcessIFFT, lib.digHoloProcessRemoveTilt,
                 lib.digHoloProcessBasisExtractCoefs):
        assert step(h) == 0
    compare(buf, cfg, api.get_coefs(lib, h), api.get_fields(lib, h)[0], avg_count=A, avg_mode=1, batch_cal=cal)
    lib.digHoloDestroy(h)
