"""Parity of the HIP path (through the C-ABI) with the FP64 oracle.

Bar: relative L2 error of fields and coefficients <= 1e-5 for the FP32 path
(the tolerance the batch path is specified to), on seeded frames of every
window-size family the path meets (radix 2, 3, 5 transforms), single and dual
polarisation, uint16 and float32 input, host and device frame pointers, a
window that wraps around ky, defocus, pol locks, G = 0, and a second call on
the same handle with a smaller batch.
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


def make_case(name, B, seed, quantise=True, **over):
    if name == "small":
        spec = synth.spec_small()
        cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), waist=200e-6,
                   group_count=10)
    elif name == "dual640":
        spec = synth.spec_dual_640()
        cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), waist=300e-6,
                   group_count=45, beam_centre_x=(-1.6e-3, 1.6e-3))
    elif name == "dual640w192":
        spec = synth.spec_dual_640()
        spec.group_count = 30
        cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(192, 192), waist=300e-6,
                   group_count=30, beam_centre_x=(-1.6e-3, 1.6e-3))
    elif name == "radix5":
        spec = synth.HologramSpec(frame_width=320, frame_height=320, pol_count=1, waist=150e-6, group_count=12)
        cfg = dict(frame_width=320, frame_height=320, pol_count=1, fft_window=(160, 160), waist=150e-6,
                   group_count=12)
    elif name == "large1024":
        spec = synth.HologramSpec(frame_width=1024, frame_height=1024, pol_count=2, waist=300e-6,
                                  group_count=80, beam_centre_x=[-2.56e-3, 2.56e-3], group_decay=20.0)
        cfg = dict(frame_width=1024, frame_height=1024, pol_count=2, fft_window=(512, 512), waist=300e-6,
                   group_count=80, beam_centre_x=(-2.56e-3, 2.56e-3))
    elif name == "radix7":
        spec = synth.HologramSpec(frame_width=224, frame_height=224, pol_count=1, waist=120e-6, group_count=8)
        cfg = dict(frame_width=224, frame_height=224, pol_count=1, fft_window=(112, 112), waist=120e-6,
                   group_count=8)
    else:
        raise KeyError(name)
    for k, v in over.pop("spec", {}).items():
        setattr(spec, k, v)
    cfg.update(over)
    frames, _, _, _ = synth.make_batch(spec, B, seed=seed, quantise=quantise)
    if not quantise:
        frames = frames.astype(np.float32)
    return frames, cfg


def oracle_for(cfg):
    P = cfg.get("pol_count", 1)
    two = lambda k, d: list(cfg.get(k, d))
    return O.Config(frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                    pixel_size=cfg.get("pixel_size", 20e-6), fft_window_x=cfg["fft_window"][0],
                    fft_window_y=cfg["fft_window"][1], aperture_deg=cfg.get("aperture_deg", 0.7),
                    wavelength=cfg.get("wavelength", 1565e-9),
                    beam_centre_x=two("beam_centre_x", (0.0, 0.0)), beam_centre_y=two("beam_centre_y", (0.0, 0.0)),
                    tilt_x=two("tilt_x", (1.5, 1.5)), tilt_y=two("tilt_y", (1.5, 1.5)),
                    defocus=two("defocus", (0.0, 0.0)), waist=[cfg["waist"]] * 2,
                    group_count=cfg["group_count"], fill_factor=cfg.get("fill_factor", 1),
                    pol_lock_tilt=cfg.get("pol_lock_tilt", 0), pol_lock_defocus=cfg.get("pol_lock_defocus", 0),
                    pol_lock_waist=cfg.get("pol_lock_waist", 1), resolution_mode=cfg.get("resolution_mode", 1))


def run_gpu(lib, frames, cfg, h=None):
    if h is None:
        h = lib.digHoloCreate()
        api.configure(lib, h, **cfg)
    coefs, fields, x, y = api.process_batch(lib, h, frames)
    return coefs, fields, x, y, h


def check(frames, cfg, coefs, fields, tol=TOL):
    f_ref, c_ref, plan = O.process_batch(frames, oracle_for(cfg))
    assert fields.shape == f_ref.shape
    ef = O.relative_l2(fields, f_ref)
    print(f"fields rel-L2 {ef:.3e}", flush=True)
    assert ef <= tol
    if cfg["group_count"] > 0:
        assert coefs.shape == c_ref.shape
        ec = O.relative_l2(coefs, c_ref)
        print(f"coefs rel-L2 {ec:.3e}", flush=True)
        assert ec <= tol
        # per-frame, per-pol too: no frame may hide behind the batch norm
        for b in range(coefs.shape[0]):
            for p in range(coefs.shape[1]):
                assert O.relative_l2(coefs[b, p], c_ref[b, p]) <= 5 * tol
    else:
        assert coefs is None
    return plan


@pytest.mark.parametrize("fused", [3, 2, 1, 0])
@pytest.mark.parametrize("name,B", [("small", 8), ("dual640", 3), ("dual640w192", 3), ("radix5", 4)])
def test_parity_uint16(lib, name, B, fused):
    frames, cfg = make_case(name, B, seed=11)
    print(f"case {name} B={B} frames {frames.shape} fused={fused}", flush=True)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, fused) == 0
    coefs, fields, x, y, h = run_gpu(lib, frames, cfg, h=h)
    # these shapes fit the fused kernels; the specialised one has no 80-point row transform
    expect = 2 if (fused == 3 and name == "radix5") else fused
    assert lib.digHoloGetWindowPath(h) == expect
    plan = check(frames, cfg, coefs, fields)
    # reconstructed axes: (i - w/2) * px W / w (digHolo.cpp:8184-8189)
    px = np.float32(20e-6)
    assert np.allclose(x, (np.arange(plan.w) - plan.w // 2) * np.float32(px * plan.W / plan.w), rtol=1e-6)


def test_parity_float_frames_wrapped_window_defocus(lib):
    frames, cfg = make_case("dual640", 2, seed=5, quantise=False,
                            tilt_x=(1.5, 1.4), tilt_y=(1.5, -1.6), defocus=(0.3, -0.2),
                            beam_centre_y=(40e-6, -70e-6), fill_factor=0,
                            spec=dict(tilt_x_deg=[1.5, 1.4], tilt_y_deg=[1.5, -1.6], defocus=[0.3, -0.2],
                                      beam_centre_y=[40e-6, -70e-6], group_count=45))
    assert frames.dtype == np.float32
    coefs, fields, _, _, _ = run_gpu(lib, frames, cfg)
    plan = check(frames, cfg, coefs, fields)
    assert plan.SY[1] < 0


def test_parity_pol_locks(lib):
    frames, cfg = make_case("dual640", 2, seed=6, tilt_x=(1.5, 1.2), tilt_y=(1.5, 1.1), defocus=(0.4, -0.3),
                            pol_lock_tilt=1, pol_lock_defocus=1)
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg)
    check(frames, cfg, coefs, fields)
    # the lock writes pol 0's values into pol 1 of the configuration
    assert abs(lib.digHoloConfigGetTilt(h, 0, 1) - 1.5) < 1e-7
    assert abs(lib.digHoloConfigGetDefocus(h, 1) - 0.4) < 1e-7


def test_group_count_zero_gives_fields_only(lib):
    frames, cfg = make_case("small", 3, seed=7, group_count=0)
    coefs, fields, _, _, _ = run_gpu(lib, frames, cfg)
    check(frames, cfg, coefs, fields)


def test_second_call_smaller_batch_and_new_config(lib):
    frames, cfg = make_case("small", 6, seed=8)
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg)
    check(frames, cfg, coefs, fields)
    frames2 = frames[:2].copy()
    cfg2 = dict(cfg, tilt_x=(1.45, 1.45), beam_centre_x=(30e-6, 0.0))
    api.configure(lib, h, **cfg2)
    coefs2, fields2, _, _, _ = run_gpu(lib, frames2, cfg2, h=h)
    check(frames2, cfg2, coefs2, fields2)


def test_device_pointer_frames_and_device_outputs(lib):
    import torch
    frames, cfg = make_case("dual640", 3, seed=9)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    dev = torch.from_numpy(frames.astype(np.int32)).to(torch.int16).cuda()   # same bits as uint16 < 32768
    torch.cuda.synchronize()
    assert lib.digHoloSetBatchUint16(h, frames.shape[0], C.c_void_p(dev.data_ptr()), 0) == 0
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    ptr = lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
    assert ptr and (b.value, p.value, m.value) == (3, 2, 1035)
    assert lib.digHoloSynchronize(h) == 0
    nbytes = b.value * p.value * m.value * 8
    host = np.empty(nbytes // 8, dtype=np.complex64)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(host.ctypes.data, ptr, nbytes, 2) == 0          # hipMemcpyDeviceToHost
    # the host API path on the same handle returns the same numbers
    hb, hm, hp = C.c_int(), C.c_int(), C.c_int()
    hptr = lib.digHoloBasisGetCoefs(h, C.byref(hb), C.byref(hm), C.byref(hp))
    again = np.ctypeslib.as_array(C.cast(hptr, C.POINTER(C.c_float)), shape=(nbytes // 4,)).view(np.complex64)
    assert np.array_equal(again, host)
    _, c_ref, _ = O.process_batch(frames, oracle_for(cfg))
    assert O.relative_l2(host.reshape(c_ref.shape), c_ref) <= TOL


@pytest.mark.parametrize("name,B", [("small", 6), ("dual640", 3), ("radix5", 3), ("large1024", 1)])
def test_split16_overlap_parity(lib, name, B):
    # split-FP16 MFMA overlap (hi+lo operands, FP32 accumulate) meets the same bar
    frames, cfg = make_case(name, B, seed=21)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetOverlapPrecision(h, api.DIGHOLO_PRECISION_SPLIT16) == 0
    coefs16, fields16, _, _, _ = run_gpu(lib, frames, cfg, h=h)
    check(frames, cfg, coefs16, fields16)
    assert lib.digHoloConfigSetOverlapPrecision(h, api.DIGHOLO_PRECISION_FP32) == 0
    coefs32, _, _, _, _ = run_gpu(lib, frames, cfg, h=h)
    _, c_ref, _ = O.process_batch(frames, oracle_for(cfg))
    e16, e32 = O.relative_l2(coefs16, c_ref), O.relative_l2(coefs32, c_ref)
    print(f"{name}: coef rel-L2 vs FP64  split16 {e16:.2e}  fp32 {e32:.2e}", flush=True)


@pytest.mark.parametrize("name,B", [("small", 6), ("dual640", 3)])
def test_fused_overlap_modes(lib, name, B):
    # the overlap inside the specialised window kernel, both precisions, against the
    # separate overlap kernel and the oracle
    frames, cfg = make_case(name, B, seed=23)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    out = {}
    for prec in (api.DIGHOLO_PRECISION_FP32, api.DIGHOLO_PRECISION_SPLIT16):
        assert lib.digHoloConfigSetOverlapPrecision(h, prec) == 0
        for mode in (1, 0):
            assert lib.digHoloConfigSetFusedOverlap(h, mode) == 0
            coefs, fields, _, _, _ = run_gpu(lib, frames, cfg, h=h)
            assert lib.digHoloGetWindowPath(h) == 3
            check(frames, cfg, coefs, fields)
            out[prec, mode] = coefs
    for prec in (api.DIGHOLO_PRECISION_FP32, api.DIGHOLO_PRECISION_SPLIT16):
        d = O.relative_l2(out[prec, 1], out[prec, 0])
        print(f"{name} precision {prec}: fused vs separate {d:.2e}", flush=True)
        assert d < 2e-6
    assert lib.digHoloConfigSetFusedOverlap(h, 2) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("precision", [0, 1])
def test_large_window_high_order(lib, precision):
    """1024x1024 sensor, 512x512 windows, 80 groups on the staged kernels; the
    split overlap takes its per-field scale from the field kernel."""
    frames, cfg = make_case("large1024", 2, seed=12)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    assert lib.digHoloGetWindowPath(h) == 0          # 160x160 field: staged kernels
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("precision", [0, 1])
def test_large_sensor_float32_frames(lib, precision):
    """The large-sensor workload as specified: 1024x1024 float32 frames (not
    quantised), dual pol, 512x512 windows, 80 groups with group power
    ~ exp(-g/20), FP32 and split-fp16 overlap, against the oracle."""
    frames, cfg = make_case("large1024", 2, seed=14, quantise=False)
    assert frames.dtype == np.float32
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    assert lib.digHoloGetWindowPath(h) == 0
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("name,B,quantise,over", [
    ("radix7", 3, True, {}), ("small", 3, True, {}), ("small", 3, False, {}),
    ("dual640", 2, False, {"fft_window": (256, 192), "beam_centre_y": (30e-6, -50e-6)}),
    ("dual640w192", 2, True, {"fft_window": (192, 256)}),
    ("large1024", 1, True, {"fft_window": (512, 384)}), ("large1024", 1, False, {"fft_window": (384, 512)})])
def test_staged_kernels_any_radix(lib, name, B, quantise, over):
    """The staged kernels: a 56-point row transform (radix 7) has no in-place
    butterfly and takes the ping-pong kernels, the others the in-place kernels
    (uint16 and float32 frames, rectangular FFT windows, ragged column strips);
    all against the oracle.  The 1024 x 1024 cases put the register row
    transform (512-wide window, 128 x 160 fields) beside the LDS column
    transform, and the LDS row transform beside the register column transform
    (512-row window, 160 x 128 fields)."""
    frames, cfg = make_case(name, B, seed=4, quantise=quantise, **over)
    print(f"case {name} B={B} quantise={quantise} {over}", flush=True)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, 0) == 0
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    assert lib.digHoloGetWindowPath(h) == 0
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("name,B", [("small", 4), ("dual640", 2)])
def test_pipeline_steps_match_process_batch(lib, name, B):
    """ProcessFFT -> ProcessIFFT -> ProcessRemoveTilt -> ProcessBasisExtractCoefs
    (digHolo.h:1670-1706) gives the batch result, and each step refuses to run
    before its input exists."""
    frames, cfg = make_case(name, B, seed=5)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloProcessIFFT(h) == api.DIGHOLO_ERROR_NULLPOINTER
    assert lib.digHoloProcessRemoveTilt(h) == api.DIGHOLO_ERROR_NULLPOINTER
    assert lib.digHoloProcessBasisExtractCoefs(h) == api.DIGHOLO_ERROR_NULLPOINTER
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchUint16(h, B, buf.ctypes.data_as(C.c_void_p), 0) == 0
    assert lib.digHoloProcessFFT(h) == 0
    assert lib.digHoloProcessRemoveTilt(h) == api.DIGHOLO_ERROR_NULLPOINTER
    assert lib.digHoloProcessIFFT(h) == 0
    assert lib.digHoloProcessRemoveTilt(h) == 0
    assert lib.digHoloProcessBasisExtractCoefs(h) == 0
    coefs = api.get_coefs(lib, h)
    fields = api.get_fields(lib, h)[0]
    check(frames, cfg, coefs, fields)
    c2, f2, _, _ = api.process_batch(lib, h, frames)
    assert O.relative_l2(fields, f2) < 1e-6 and O.relative_l2(coefs, c2) < 1e-6
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,B", [("small", 3), ("dual640w192", 2), ("dual640", 2)])
def test_parity_full_resolution(lib, name, B, precision):
    """IFFTResolutionMode 0 (the reference default): the W x H field plane, on the
    staged kernels, FP32 and split-fp16 overlap."""
    frames, cfg = make_case(name, B, seed=12, resolution_mode=0)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    coefs, fields, x, y, h = run_gpu(lib, frames, cfg, h=h)
    geo = api.window_geometry(lib, h)
    assert fields.shape[-2:] == (geo["H"], geo["W"])
    assert lib.digHoloGetWindowPath(h) == 0
    check(frames, cfg, coefs, fields)
    px = np.float32(20e-6)
    assert np.allclose(x, (np.arange(geo["W"]) - geo["W"] // 2) * px, rtol=1e-6)
    lib.digHoloDestroy(h)


def test_full_resolution_steps_and_averaging(lib):
    B, A = 2, 2
    frames, cfg = make_case("small", B * A, seed=13, resolution_mode=0)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchAvgUint16(h, B, buf.ctypes.data_as(C.c_void_p), A, 0, 0) == 0
    for step in (lib.digHoloProcessFFT, lib.digHoloProcessIFFT, lib.digHoloProcessRemoveTilt,
                 lib.digHoloProcessBasisExtractCoefs):
        assert step(h) == 0
    coefs, fields = api.get_coefs(lib, h), api.get_fields(lib, h)[0]
    f_ref, c_ref, _ = O.process_batch(buf, oracle_for(cfg), avg_count=A, avg_mode=0)
    assert O.relative_l2(fields, f_ref) <= TOL and O.relative_l2(coefs, c_ref) <= TOL
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("fused,precision", [(0, api.DIGHOLO_PRECISION_FP32), (0, api.DIGHOLO_PRECISION_SPLIT16),
                                               (1, api.DIGHOLO_PRECISION_SPLIT16), (2, api.DIGHOLO_PRECISION_FP32),
                                               (3, api.DIGHOLO_PRECISION_SPLIT16)])
def test_batch_beyond_grid_frame_extent(lib, fused, precision):
    """More frames than a grid's y / z extent (65535): the kernels that take
    the frame from blockIdx.y / z are launched over the batch in chunks.  Frame
    b of the batch is unique frame b mod 4; the frames on both sides of the
    chunk boundary and the last one match the oracle."""
    frames, cfg = make_case("small", 4, seed=3)
    B = 65536 + 3
    print(f"B={B} fused={fused} precision={precision}", flush=True)
    big = np.ascontiguousarray(frames[np.arange(B) % 4])
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, fused) == 0
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    assert lib.digHoloSetBatchUint16(h, B, big.ctypes.data_as(C.c_void_p), 0) == 0
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    cptr = lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p))
    assert cptr and b.value == B
    assert lib.digHoloGetWindowPath(h) == fused
    nc = m.value * p.value
    coefs = np.ctypeslib.as_array(C.cast(cptr, api.P_float), shape=(2 * B * nc,)).view(np.complex64)
    fb, fp, w, hh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    xp, yp = api.P_float(), api.P_float()
    fptr = lib.digHoloGetFields(h, C.byref(fb), C.byref(fp), C.byref(xp), C.byref(yp), C.byref(w), C.byref(hh))
    assert fptr and fb.value == B
    nf = fp.value * hh.value * w.value
    fields = np.ctypeslib.as_array(C.cast(fptr, api.P_float), shape=(2 * B * nf,)).view(np.complex64)
    f_ref, c_ref, _ = O.process_batch(frames, oracle_for(cfg))
    for fr in (0, 65534, 65535, 65536, B - 1):
        u = fr % 4
        ef = O.relative_l2(fields[fr * nf:(fr + 1) * nf].reshape(f_ref[u].shape), f_ref[u])
        ec = O.relative_l2(coefs[fr * nc:(fr + 1) * nc].reshape(c_ref[u].shape), c_ref[u])
        print(f"frame {fr}: fields {ef:.2e} coefs {ec:.2e}", flush=True)
        assert ef <= 5 * TOL and ec <= 5 * TOL
    lib.digHoloDestroy(h)
    del big


def test_batch_frame_limit(lib):
    """A batch beyond the library's frame limit (2^24) is refused before any
    frame is read, and the handle still processes a valid batch afterwards."""
    frames, cfg = make_case("small", 2, seed=3)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    buf = np.ascontiguousarray(frames)
    assert lib.digHoloSetBatchUint16(h, (1 << 24) + 1, buf.ctypes.data_as(C.c_void_p), 0) == 0
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    assert not lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p))
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


def test_window_kernel_grid_pairs_pols(lib):
    """The window kernel's one-dimensional grid pairs the two pols of frames
    8q .. 8q + 7 sixteen workgroups at a time and maps the frames after the
    last full group plainly: 19 frames cover both parts."""
    frames, cfg = make_case("dual640", 19, seed=21)
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg)
    assert lib.digHoloGetWindowPath(h) == 3
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("precision", [0, 1])
def test_distance_from_reference_emulation(lib, precision):
    """The HIP path against the oracle's emulation of the reference's int16 tail
    (float32 sincos, int16 reference wave, fields and HG modes, float32 block
    accumulation; digHolo.cpp:716-812, :10386-10497, :2701-2837, :1444-1471,
    :2149-2341) at 640x512 dual-pol, G = 45: within the 1e-4 relative L2 that a
    comparison with the reference's own output allows, and as far from it as the
    exact path is (the HIP path sits ~3e-7 from the exact values)."""
    frames, cfg = make_case("dual640", 3, seed=21)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
        coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    finally:
        lib.digHoloDestroy(h)
    oc = oracle_for(cfg)
    f_ex, c_ex, _ = O.process_batch(frames, oc)
    f_rf, c_rf, _ = O.process_batch(frames, oc, emulate_reference=True)
    ef, ec = O.relative_l2(fields, f_rf), O.relative_l2(coefs, c_rf)
    ef0, ec0 = O.relative_l2(f_ex, f_rf), O.relative_l2(c_ex, c_rf)
    print(f"precision {precision}: HIP vs reference emulation fields {ef:.3e} coefs {ec:.3e}; "
          f"exact vs emulation fields {ef0:.3e} coefs {ec0:.3e}", flush=True)
    assert ef <= 1e-4 and ec <= 1e-4
    assert abs(ef - ef0) <= 0.02 * ef0 and abs(ec - ec0) <= 0.02 * ec0


def _tilt_for_c0(lib, cfg, target_c0):
    """The x tilt (degrees) whose Fourier window starts at column target_c0 of
    the r2c half plane (window start SX - W/2), from the host-only plan."""
    h = lib.digHoloCreate()
    try:
        for t in np.arange(0.30, 2.20, 0.001):
            c = dict(cfg, tilt_x=(float(t), float(t)))
            api.configure(lib, h, **c)
            geo = api.window_geometry(lib, h)
            if all(sx - geo["W"] // 2 == target_c0 for sx in geo["SX"]):
                return float(t)
    finally:
        lib.digHoloDestroy(h)
    return None


@pytest.mark.parametrize("case", ["16bit", "c0=0"])
def test_parity_mfma_rows_edge_cases(lib, case):
    """The MFMA row transform of 256-wide windows: samples above 12 bits (an
    exact fp16 remainder product) and a window that starts at column 0 (the
    column-0 sum of the -2048 sample offset restored).  (The window planner
    keeps Fourier windows below the Nyquist column on these geometries: their
    start is even and at most W/2 - w.)"""
    _, cfg = make_case("dual640", 2, seed=3)
    spec_over = {}
    if case != "16bit":
        t = _tilt_for_c0(lib, cfg, 0)
        if t is None:
            pytest.skip("no tilt gives that window start")
        cfg["tilt_x"] = (t, t)
        spec_over = dict(tilt_x_deg=[t, t])
    frames, cfg = make_case("dual640", 2, seed=3, spec=spec_over, **{k: v for k, v in cfg.items() if k == "tilt_x"})
    if case == "16bit":
        frames = (frames.astype(np.uint32) * 16).astype(np.uint16)   # up to 65520: every sample above 4096 possible
        assert frames.max() > 4096
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    assert lib.digHoloConfigSetFusedWindow(h, 3) == 0
    coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
    assert lib.digHoloGetWindowPath(h) == 3
    geo = api.window_geometry(lib, h)
    print(f"case {case}: tilt {cfg.get('tilt_x')} window starts {[sx - geo['W'] // 2 for sx in geo['SX']]}", flush=True)
    check(frames, cfg, coefs, fields)
    lib.digHoloDestroy(h)


@pytest.mark.parametrize("rows", [0, 1])
def test_parity_window_row_transforms(lib, rows):
    """Both row transforms of the size-specialised kernel on the 256-wide
    window (0: MFMA, 1: registers), split and FP32 overlap."""
    frames, cfg = make_case("dual640", 3, seed=21)
    for prec in (api.DIGHOLO_PRECISION_SPLIT16, api.DIGHOLO_PRECISION_FP32):
        h = lib.digHoloCreate()
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetWindowRowTransform(h, rows) == 0
        assert lib.digHoloConfigGetWindowRowTransform(h) == rows
        assert lib.digHoloConfigSetOverlapPrecision(h, prec) == 0
        coefs, fields, _, _, h = run_gpu(lib, frames, cfg, h=h)
        assert lib.digHoloGetWindowPath(h) == 3
        print(f"row transform {rows} precision {prec}", flush=True)
        check(frames, cfg, coefs, fields)
        lib.digHoloDestroy(h)
    h = lib.digHoloCreate()
    assert lib.digHoloConfigSetWindowRowTransform(h, 2) == api.DIGHOLO_ERROR_INVALIDARGUMENT
    lib.digHoloDestroy(h)
