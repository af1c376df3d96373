"""digHoloBenchmark (digHolo.h:1776, digHolo.cpp:13241-13408) on the GPU path.

The reference returns batches per second end to end (hzTotalMeas,
digHolo.cpp:13345, :13396) and writes seven floats: the FFT, IFFT, ApplyTilt
and overlap stage rates, the HG-basis regeneration rate, the rate the stage
times add up to and the measured end-to-end rate (:13374-13382).  Checked here:
every entry is filled and positive, info[6] is the return value, and the
return value times the batch size agrees with frames/s timed independently
around digHoloProcessBatchDevice (HIP events on the handle's stream), on the
fused window kernel and on the staged kernels.
"""
import ctypes as C

import numpy as np
import pytest

from digholo_amd import api, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


@pytest.mark.parametrize("window_mode,precision", [(3, 1), (3, 0), (0, 1)])
def test_benchmark_rates(lib, window_mode, precision):
    import torch
    B = 200
    spec = synth.spec_dual_640()
    frames, _, _, _ = synth.make_batch(spec, 20, seed=4)
    dev = torch.from_numpy(np.tile(frames, (B // 20, 1, 1)).astype(np.int16)).cuda()
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256),
                      waist=300e-6, group_count=45, beam_centre_x=(-1.6e-3, 1.6e-3))
        assert lib.digHoloConfigSetFusedWindow(h, window_mode) == 0
        assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
        assert lib.digHoloSetBatchUint16(h, B, C.c_void_p(dev.data_ptr()), 0) == 0
        info = (C.c_float * 7)()
        ret = lib.digHoloBenchmark(h, C.c_float(0.5), info)
        vals = list(info)
        print(f"window mode {window_mode} precision {precision}: return {ret:.2f} batches/s, info {vals}, "
              f"path {lib.digHoloGetWindowPath(h)}", flush=True)
        assert ret > 0 and all(v > 0 for v in vals)
        assert vals[6] == pytest.approx(ret, rel=1e-6)
        # the stages add up to about the measured rate (stage events cost a little)
        assert 0.5 * ret <= vals[5] <= 1.5 * ret
        # frames/s timed around the device entry point
        b, m, p = C.c_int(), C.c_int(), C.c_int()
        stream = torch.cuda.ExternalStream(lib.digHoloGetStream(h))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record(stream)
        for _ in range(reps):
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        e1.record(stream)
        e1.synchronize()
        fps = B * reps / (e0.elapsed_time(e1) * 1e-3)
        print(f"frames/s: benchmark {ret * B:.0f}, events {fps:.0f}", flush=True)
        assert ret * B == pytest.approx(fps, rel=0.25)
    finally:
        lib.digHoloDestroy(h)
