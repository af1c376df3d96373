"""Several handles' batches in one launch sequence (digHoloProcessBatchMultiDevice)
and the launch autotuner (digHoloConfigSetAutoTune); no counterpart in the
reference.

Three handle configurations of the uneven-batch workload (320x256 / 128 window
/ 10 groups, 640x512 / 192 / 30, 1024x1024 / 256 / 45; dual pol, uint16 frames
in HBM), at small batch counts (one launch of the multi-batch window kernel):
* every call after the capture reproduces the coefficients and fields each
  handle computes on its own (outputs cleared between calls, eager torch work
  and synchronisations in between): the merged launch runs the same kernel
  body compiled into another kernel, whose floating-point contraction may
  differ from the handle's own kernel by an ulp, so it is held to 1e-6 of the
  largest value, and every replay to the first one bit for bit;
* above the one-launch limit the handles' own kernels run one after the other,
  again bit for bit, ordered against work on the handles' own streams;
* a configuration change re-captures, and the replay again matches;
* with autotuning on, the chosen kernels reproduce the default ones to 1e-5
  relative L2, and the choice lands in the cache file.
"""
import ctypes as C
import os

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [  # frame W, H, window, groups, batch
    (320, 256, 128, 10, 12),
    (640, 512, 192, 30, 6),
    (1024, 1024, 256, 45, 3),
]


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


class DevArray:
    """torch view of library-owned device memory (float32 pairs)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n, 2), "typestr": "<f4", "data": (ptr, False), "version": 3}


def frames_for(W, H, G, B, seed):
    import torch
    spec = synth.HologramSpec(frame_width=W, frame_height=H, pol_count=2, waist=200e-6, group_count=min(G, 10),
                              beam_centre_x=[-W * 5e-6, W * 5e-6])
    f, _, _, _ = synth.make_batch(spec, B, seed=seed)
    return torch.from_numpy(f.astype(np.int16)).cuda()


def make_handle(lib, W, H, win, G, B, frames):
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=W, frame_height=H, pol_count=2, fft_window=(win, win), waist=200e-6,
                  group_count=G, beam_centre_x=(-W * 5e-6, W * 5e-6))
    assert lib.digHoloSetBatchUint16(h, B, C.c_void_p(frames.data_ptr()), 0) == 0
    return h


def outputs(lib, h):
    import torch
    b, p, w, hh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    fptr = lib.digHoloGetFieldsDevice(h, C.byref(b), C.byref(p), C.byref(w), C.byref(hh))
    nf = b.value * p.value * w.value * hh.value
    fields = torch.as_tensor(DevArray(fptr, nf), device="cuda")
    return fields, (b.value, p.value, hh.value, w.value)


def close(a, b, rel=1e-6):
    """a equals b to FP32 rounding: max |a - b| <= rel max |b|."""
    return float((a - b).abs().max()) <= rel * float(b.abs().max())


def individual(lib, h):
    import torch
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    cptr = lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
    assert cptr
    assert lib.digHoloSynchronize(h) == 0
    n = b.value * p.value * m.value
    coefs = torch.as_tensor(DevArray(cptr, n), device="cuda")
    fields, shape = outputs(lib, h)
    return coefs, fields


def test_multi_graph_replays_match_individual_batches(lib):
    import torch
    hs, frames, ref, views = [], [], [], []
    try:
        for i, (W, H, win, G, B) in enumerate(SHAPES):
            fr = frames_for(W, H, G, B, seed=i)
            frames.append(fr)
            h = make_handle(lib, W, H, win, G, B, fr)
            hs.append(h)
            c, f = individual(lib, h)
            views.append((c, f))
            ref.append((c.clone(), f.clone()))
            print(f"handle {h}: {W}x{H} window {win} G {G} B {B}: coefs {tuple(c.shape)} fields {tuple(f.shape)}",
                  flush=True)
        ids = (C.c_int * len(hs))(*hs)
        first = None
        for rep in range(4):
            for c, f in views:          # clear the outputs so a replay that did nothing fails
                c.zero_()
                f.zero_()
            torch.cuda.synchronize()
            x = torch.randn(1 << 20, device="cuda").square().sum()   # eager work between replays
            assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
            assert lib.digHoloSynchronize(hs[0]) == 0
            for (c, f), (cr, fr_) in zip(views, ref):
                assert close(c, cr), f"replay {rep}: coefficients differ"
                assert close(f, fr_), f"replay {rep}: fields differ"
            if first is None:
                first = [(c.clone(), f.clone()) for c, f in views]
            for (c, f), (c0, f0) in zip(views, first):
                assert torch.equal(c, c0) and torch.equal(f, f0), f"replay {rep} differs from replay 0"
            print(f"replay {rep} matches ({x.item():.1f})", flush=True)
        # a changed configuration re-captures
        lib.digHoloConfigSetTilt(hs[1], 0, 0, 1.48)
        c1, f1 = individual(lib, hs[1])
        want = (c1.clone(), f1.clone())
        for c, f in views:
            c.zero_()
            f.zero_()
        torch.cuda.synchronize()
        assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
        assert lib.digHoloSynchronize(hs[0]) == 0
        assert close(views[1][0], want[0]) and close(views[1][1], want[1])
        assert close(views[0][0], ref[0][0]) and close(views[2][1], ref[2][1])
    finally:
        for h in hs:
            lib.digHoloDestroy(h)


@pytest.mark.parametrize("shape", [SHAPES[1], (640, 512, 256, 45, 8)])
def test_autotune_choice_and_cache(lib, tmp_path, monkeypatch, shape):
    """The tuned launch reproduces the default one to 1e-5 and lands in the
    cache (key path fused tpw rows ms); the 256-wide window also has the row
    transform choice (MFMA or registers)."""
    import torch
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("DIGHOLO_TUNE_CACHE", str(cache))
    W, H, win, G, B = shape
    fr = frames_for(W, H, G, 8, seed=5)
    h = make_handle(lib, W, H, win, G, 8, fr)
    try:
        for prec in (api.DIGHOLO_PRECISION_FP32, api.DIGHOLO_PRECISION_SPLIT16):
            lib.digHoloConfigSetOverlapPrecision(h, prec)
            lib.digHoloConfigSetAutoTune(h, 0)
            c0, f0 = individual(lib, h)
            c0, f0 = c0.clone(), f0.clone()
            assert lib.digHoloConfigSetAutoTune(h, 1) == 0
            c1, f1 = individual(lib, h)
            path = lib.digHoloGetWindowPath(h)
            rows = lib.digHoloConfigGetWindowRowTransform(h)
            print(f"precision {prec}: tuned window path {path}, row transform {rows}", flush=True)
            assert rows in (0, 1) and (rows == 0 or win == 256)
            ec = O.relative_l2(c1.cpu().numpy(), c0.cpu().numpy())
            ef = O.relative_l2(f1.cpu().numpy(), f0.cpu().numpy())
            assert ec <= 1e-5 and ef <= 1e-5, (ec, ef)
        text = cache.read_text()
        print(text, flush=True)
        # (the library's tuned choices are process-wide: the file holds every
        # shape tuned so far, two of them this one's)
        lines = [ln for ln in text.splitlines() if ln.strip()]
        assert all(len(ln.split()) == 6 for ln in lines)
        assert len([ln for ln in lines if f"_W{win}x{win}_" in ln and f"_B{8}" in ln]) == 2
        # mode 2 reuses the remembered choice without timing anything new
        assert lib.digHoloConfigSetAutoTune(h, 2) == 0
        individual(lib, h)
        assert cache.read_text() == text
    finally:
        lib.digHoloDestroy(h)


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_graph_outputs_match_oracle(lib, precision):
    """The uneven-batch workload's three handle shapes (320x256 / 128 / 10
    groups, 640x512 / 192 / 30, 1024x1024 / 256 / 45; dual pol) through one graph
    launch: each handle's fields and coefficients, read after the replay, are
    within 1e-5 relative L2 of the FP64 oracle (FP32 and split-fp16 overlap)."""
    hs, frames, host = [], [], []
    try:
        for i, (W, H, win, G, B) in enumerate(SHAPES):
            fr = frames_for(W, H, G, B, seed=30 + i)
            frames.append(fr)
            host.append(fr.cpu().numpy().astype(np.uint16))
            h = make_handle(lib, W, H, win, G, B, fr)
            assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
            hs.append(h)
        ids = (C.c_int * len(hs))(*hs)
        for rep in range(2):        # the first launch captures, the second replays
            assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
        assert lib.digHoloSynchronize(hs[0]) == 0
        for (W, H, win, G, B), h, fr in zip(SHAPES, hs, host):
            coefs = api.get_coefs(lib, h)
            fields = api.get_fields(lib, h)[0]
            oc = O.Config(frame_width=W, frame_height=H, pol_count=2, fft_window_x=win, fft_window_y=win,
                          aperture_deg=0.7, beam_centre_x=[-W * 5e-6, W * 5e-6], tilt_x=[1.5, 1.5],
                          tilt_y=[1.5, 1.5], group_count=G, waist=[200e-6, 200e-6])
            f_ref, c_ref, _ = O.process_batch(fr, oc)
            ef, ec = O.relative_l2(fields, f_ref), O.relative_l2(coefs, c_ref)
            print(f"{W}x{H} window {win} G {G} B {B} path {lib.digHoloGetWindowPath(h)}: "
                  f"fields {ef:.2e} coefs {ec:.2e}", flush=True)
            assert ef <= 1e-5 and ec <= 1e-5
    finally:
        for h in hs:
            lib.digHoloDestroy(h)


def test_multi_large_batches_and_stream_order(lib):
    """Batches above the one-launch limit (1060 (pol, frame) items: 520 frames of
    320x256 / 128 / 10 groups and 10 of 640x512 / 192 / 30) run as the handles'
    own window kernels one after the other on the first handle's stream: every
    call reproduces each handle's own outputs bit for bit.  Ordering across the
    handles' streams: work queued on the second handle's stream before the call
    (a 1 ms spin, then its coefficients cleared) comes first, and a later
    download on that stream (digHoloBasisGetCoefs) waits for a call held back by
    a spin on the first handle's stream."""
    import torch
    shapes = [(320, 256, 128, 10, 520), (640, 512, 192, 30, 10)]
    hs, frames, ref, views = [], [], [], []
    try:
        for i, (W, H, win, G, B) in enumerate(shapes):
            fr = frames_for(W, H, G, B, seed=50 + i)
            frames.append(fr)
            h = make_handle(lib, W, H, win, G, B, fr)
            hs.append(h)
            c, f = individual(lib, h)
            views.append((c, f))
            ref.append((c.clone(), f.clone()))
            print(f"handle {h}: {W}x{H} window {win} G {G} B {B}", flush=True)
        ids = (C.c_int * len(hs))(*hs)
        for rep in range(3):
            for c, f in views:
                c.zero_()
                f.zero_()
            torch.cuda.synchronize()
            assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
            assert lib.digHoloSynchronize(hs[0]) == 0
            for (c, f), (cr, fr_) in zip(views, ref):
                assert torch.equal(c, cr), f"call {rep}: coefficients differ"
                assert torch.equal(f, fr_), f"call {rep}: fields differ"
            print(f"call {rep} matches", flush=True)
        s0 = torch.cuda.ExternalStream(lib.digHoloGetStream(hs[0]))
        s1 = torch.cuda.ExternalStream(lib.digHoloGetStream(hs[1]))
        # work on the second handle's stream before the call comes first
        with torch.cuda.stream(s1):
            torch.cuda._sleep(2_000_000)
            views[1][0].zero_()
        assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
        assert lib.digHoloSynchronize(hs[0]) == 0
        assert torch.equal(views[1][0], ref[1][0]), "the call ran ahead of the second handle's stream"
        # a download on the second handle's stream follows a delayed call
        views[1][0].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s0):
            torch.cuda._sleep(2_000_000)
        assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
        got = api.get_coefs(lib, hs[1])
        want = ref[1][0].cpu().numpy().view(np.complex64).reshape(got.shape)
        assert np.array_equal(got, want), "the download ran ahead of the call"
        assert lib.digHoloSynchronize(hs[0]) == 0
    finally:
        for h in hs:
            lib.digHoloDestroy(h)
