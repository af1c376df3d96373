"""digHoloProcessBatchDeviceNoAlloc: the allocation-free batch entry point (no
counterpart in the reference).  Its fields and coefficients equal those of
digHoloProcessBatchDevice on the same frames to 1e-5 relative L2 (the batch
path's floating-point bar; the two run different kernels) and meet the same
bar against the FP64 oracle, on two calls in a row with the second batch
smaller, in both overlap precisions and on uint16 and float32 frames.  A batch
larger than the scratch array is refused without running.
"""
import ctypes as C

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


CASES = {
    "dual640": (synth.spec_dual_640, dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256),
                                          waist=300e-6, group_count=45, beam_centre_x=(-1.6e-3, 1.6e-3))),
    "small": (synth.spec_small, dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128),
                                     waist=200e-6, group_count=10)),
}


def oracle_cfg(cfg):
    P = cfg["pol_count"]
    bx = list(cfg.get("beam_centre_x", (0.0, 0.0)))
    return O.Config(frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                    fft_window_x=cfg["fft_window"][0], fft_window_y=cfg["fft_window"][1], aperture_deg=0.7,
                    tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=cfg["group_count"],
                    waist=[cfg["waist"]] * 2, beam_centre_x=bx)


@pytest.mark.parametrize("name,precision,u16", [("dual640", 1, True), ("dual640", 0, True), ("small", 1, False)])
def test_no_alloc_matches_batch_entry_point(lib, name, precision, u16):
    import torch
    make_spec, cfg = CASES[name]
    frames, _, _, _ = synth.make_batch(make_spec(), 6, seed=17)
    print(f"case {name} precision {precision} uint16 {u16}", flush=True)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
        for B in (6, 3):                      # two calls, the second with fewer frames
            host = np.ascontiguousarray(frames[:B])
            if u16:
                dev = torch.from_numpy(host.astype(np.int16)).cuda()
            else:
                dev = torch.from_numpy(host.astype(np.float32)).cuda()
            torch.cuda.synchronize()
            # the allocating entry point first: it builds the tables and plans
            coefs_ref, fields_ref, _, _ = api.process_batch(lib, h, host if u16 else host.astype(np.float32))
            Pn, M = coefs_ref.shape[1], coefs_ref.shape[2]
            hh, w = fields_ref.shape[2], fields_ref.shape[3]
            fields = torch.full((B, Pn, hh, w, 2), float("nan"), device="cuda")
            coefs = torch.full((B, Pn, M, 2), float("nan"), device="cuda")
            torch.cuda.synchronize()
            rc = lib.digHoloProcessBatchDeviceNoAlloc(h, C.c_void_p(dev.data_ptr()), B, 1 if u16 else 0,
                                                      C.c_void_p(fields.data_ptr()), C.c_void_p(coefs.data_ptr()))
            assert rc == 0
            assert lib.digHoloSynchronize(h) == 0
            f = fields.cpu().numpy().view(np.complex64)[..., 0]
            c = coefs.cpu().numpy().view(np.complex64)[..., 0]
            ef, ec = O.relative_l2(f, fields_ref), O.relative_l2(c, coefs_ref)
            f_or, c_or, _ = O.process_batch(host, oracle_cfg(cfg))
            of, oc = O.relative_l2(f, f_or), O.relative_l2(c, c_or)
            print(f"B={B}: vs ProcessBatch fields {ef:.2e} coefs {ec:.2e}; vs FP64 fields {of:.2e} coefs {oc:.2e}",
                  flush=True)
            assert ef <= TOL and ec <= TOL and of <= TOL and oc <= TOL
    finally:
        lib.digHoloDestroy(h)


def test_no_alloc_refuses_oversized_batch(lib):
    import torch
    make_spec, cfg = CASES["dual640"]
    frames, _, _, _ = synth.make_batch(make_spec(), 2, seed=3)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        dev = torch.from_numpy(frames.astype(np.int16)).cuda()
        out = torch.zeros(16, device="cuda")
        torch.cuda.synchronize()
        # before any batch: no tables yet
        assert lib.digHoloProcessBatchDeviceNoAlloc(h, C.c_void_p(dev.data_ptr()), 2, 1, C.c_void_p(out.data_ptr()),
                                                    C.c_void_p(out.data_ptr())) == api.DIGHOLO_ERROR_NULLPOINTER
        api.process_batch(lib, h, frames)
        # 640x512 dual pol needs 430,088 bytes per frame: 200 frames exceed 64 MiB and are refused unlaunched
        rc = lib.digHoloProcessBatchDeviceNoAlloc(h, C.c_void_p(dev.data_ptr()), 200, 1, C.c_void_p(out.data_ptr()),
                                                  C.c_void_p(out.data_ptr()))
        assert rc == api.DIGHOLO_ERROR_MEMORYALLOCATION
        assert lib.digHoloSynchronize(h) == 0
        assert torch.count_nonzero(out).item() == 0
    finally:
        lib.digHoloDestroy(h)
