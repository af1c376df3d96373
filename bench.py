"""Benchmark of the digHolo batch path on one MI355X.

Workload: 640x512 uint16 frames, dual polarisation, batch 1000, FFT window
256x256 per pol (IFFT window 80x80), 45 HG groups (1035 modes per pol).  One
step = digHoloProcessBatchDevice over the whole batch with the frames already
resident in HBM, fields and coefficients left in HBM.

Prints one JSON line.  "value" is frames/s over K timed steps with nothing
but the batches on the stream (HIP events around the K steps).  "roofline" is
the dominant kernel (the window kernel): achieved = algorithmic bytes per
launch (window pixels in, fields and coefficients out) / its average launch
time, from HIP events on the handle's stream around that kernel alone over a
second set of K steps; "traffic" is its measured HBM bytes per launch from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE record.  After timing, frames
0, B/2 and B-1 of the timed batch are checked against the FP64 oracle.
"host_inclusive" is the rate with host frames in and coefficients and fields
back on the host (digHoloProcessBatch + digHoloGetFields).  "cpu_baseline" is
the FP64 oracle on a bounded sample of the same workload, single-threaded,
with the reference's own CPU figure (measured in another container) beside it.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import tempfile
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s end-to-end 640x512 dual-pol, 45 HG groups: split-BF16 MFMA vs FP32"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

B = 1000
FRAME_W, FRAME_H, POLS, WIN, GROUPS, WAIST = 640, 512, 2, 256, 45, 300e-6
BEAM_X = (-1.6e-3, 1.6e-3)


def algorithmic_bytes_per_frame(P, W, H, w, h, M):
    """Window input (uint16) + coefficients + fields out, per frame (all pols)."""
    return P * W * H * 2 + P * M * 8 + P * w * h * 8


def fft_stage_bytes_per_frame(P, W, H, w, h):
    """Window input (uint16) + the h x w complex window out, per frame: the FFT
    stage's algorithmic traffic (the fused kernel writes the field, the same size)."""
    return P * W * H * 2 + P * w * h * 8


def overlap_flops_per_frame(P, G, w, h):
    """Separable overlap, real HG x complex field = 4 flop/MAC, triangle in stage 2."""
    M = G * (G + 1) // 2
    return P * 4 * (G * w * h + M * h)


def measured_traffic(kernel, fused_overlap):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a copy kernel of known size), or
    None when no summary for this kernel and mode is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    key = kernel + (":fused_overlap" if fused_overlap else "")
    ent = rec.get(key)
    return ent.get("bytes_per_launch") if isinstance(ent, dict) else None


def make_frames(n_unique=100, seed=1, spec=None, batch=B):
    """n_unique rendered holograms; every one of the batch frames gets its own read noise."""
    import numpy as np
    import torch
    from digholo_amd import synth
    spec = spec or synth.spec_dual_640()
    rng = np.random.default_rng(seed)
    coefs = synth.random_coefs(spec, n_unique, rng)
    clean, _, _ = synth.render(spec, coefs, quantise=False)
    base = torch.from_numpy(clean.astype(np.float32)).cuda()
    out = torch.empty((batch, spec.frame_height, spec.frame_width), dtype=torch.int16, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    for b0 in range(0, batch, n_unique):
        n = min(n_unique, batch - b0)
        noisy = base[:n] + spec.noise_counts * torch.randn(base[:n].shape, generator=g, device="cuda")
        out[b0:b0 + n] = noisy.round().clamp(0, 4095).to(torch.int16)
    torch.cuda.synchronize()
    return out


def cpu_baseline(min_s=10.0, max_s=30.0):
    """FP64 oracle timed on this host on a bounded sample of the same workload."""
    import numpy as np
    from digholo_amd import synth
    from oracle import digholo_oracle as O
    spec = synth.spec_dual_640()
    frames, _, _, _ = synth.make_batch(spec, 20, seed=7)
    cfg = O.Config(frame_width=FRAME_W, frame_height=FRAME_H, pol_count=POLS, fft_window_x=WIN,
                   fft_window_y=WIN, aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5],
                   group_count=GROUPS, waist=[WAIST, WAIST], beam_centre_x=list(BEAM_X))
    done, t0 = 0, time.perf_counter()
    while True:
        O.process_batch(frames, cfg)
        done += frames.shape[0]
        el = time.perf_counter() - t0
        if el >= min_s or el >= max_s:
            break
    return {"value": done / el, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{done} frames of the same workload through the FP64 numpy oracle "
                      f"(single-threaded pocketfft + einsum), {el:.1f} s",
            "reference_cross_box": REFERENCE_CPU}


# The reference's own CPU path on this workload (digHoloBenchmark end-to-end
# rate, 640x512 dual pol, batch 1000, FFT 256 -> IFFT 80, 45 groups), measured
# in another container: it cannot be built where the GPU is.
REFERENCE_CPU = {"value": 7712.0, "unit": "frames/s", "cores": 8, "kind": "reference",
                 "sample": "digHoloBenchmark end-to-end 7.712 batches/s x 1000 frames, float32 frames, "
                           "8 threads, Intel Xeon 2.70 GHz AVX2/FMA, FFTs from MKL's FFTW interface; "
                           "measured in a different container from the GPU host"}


def autoalign_comparison(lib, seed=3, batch=500, mdl_rel=0.05):
    """Alignment workload: 640x512 dual-pol frames, batch 500, 45 groups, with the
    true tilt (+-0.2 deg), defocus (+-0.5 dioptre), waist (+-10 %) and beam
    centre (+-50 um) of each pol drawn from the seed and hidden.  Both searches
    start from the nominal values (tilt 1.5 deg, centres -/+1.6 mm, defocus 0,
    waist 300 um) and aim at the mode-dependent loss: the reference's tweak
    search (digHoloAutoAlign, TWEAK, goal MDL: an insertion-loss pass, then the
    MDL pass) and the gradient search (digHoloAutoAlignGradientSearch) on
    -IL_lin + mu (ln s_max^2 - ln s_min^2) with mu = mdl_rel x |IL_lin| at the
    start, each timed on the host around the call; the insertion loss and MDL
    each ends on are those of its final evaluation.  Also the cost of one
    gradient evaluation with and without the MDL term."""
    import numpy as np
    from digholo_amd import api, synth
    rng = np.random.default_rng(seed)
    spec = synth.spec_dual_640()
    spec.tilt_x_deg = list(1.5 + rng.uniform(-0.2, 0.2, 2))
    spec.tilt_y_deg = list(1.5 + rng.uniform(-0.2, 0.2, 2))
    spec.defocus = list(rng.uniform(-0.5, 0.5, 2))
    spec.waist = WAIST * (1 + rng.uniform(-0.1, 0.1))
    spec.beam_centre_x = [BEAM_X[p] + rng.uniform(-50e-6, 50e-6) for p in range(2)]
    spec.beam_centre_y = list(rng.uniform(-50e-6, 50e-6, 2))
    frames = make_frames(n_unique=50, seed=seed, spec=spec, batch=batch)
    res = {}
    mu = None
    for method in ("gradient_cost", "tweak", "gradient"):
        h = lib.digHoloCreate()
        try:
            api.configure(lib, h, frame_width=FRAME_W, frame_height=FRAME_H, pol_count=POLS,
                          fft_window=(WIN, WIN), waist=WAIST, group_count=GROUPS, beam_centre_x=BEAM_X)
            lib.digHoloConfigSetAutoAlignMode(h, api.DIGHOLO_AUTOALIGNMODE_TWEAK)
            lib.digHoloConfigSetAutoAlignGoalIdx(h, 1)   # DIGHOLO_METRIC_MDL
            assert lib.digHoloSetBatchUint16(h, batch, C.c_void_p(frames.data_ptr()), 0) == 0
            b, m, p = C.c_int(), C.c_int(), C.c_int()
            lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))   # tables, warm-up
            lib.digHoloSynchronize(h)
            if method == "gradient_cost":
                loss0, _ = api.alignment_gradient(lib, h)
                mu = mdl_rel * abs(loss0)
                cost = {}
                for name, w in (("power_only", 0.0), ("with_mdl", mu)):
                    api.alignment_gradient(lib, h, mdl_weight=w)
                    t0 = time.perf_counter()
                    for _ in range(5):
                        api.alignment_gradient(lib, h, mdl_weight=w)
                    cost[name] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
                res["gradient_eval_ms"] = cost
                continue
            t0 = time.perf_counter()
            ret = lib.digHoloAutoAlign(h) if method == "tweak" else lib.digHoloAutoAlignGradientSearch(h, 40, mu)
            dt = time.perf_counter() - t0
            prm = api.get_align_params(lib, h, 2)
            res[method] = {"seconds": round(dt, 4), "goal_mdl_db": round(float(ret), 4),
                           "il_db": round(float(lib.digHoloAutoAlignGetMetric(h, 0)), 4),
                           "mdl_db": round(float(lib.digHoloAutoAlignGetMetric(h, 1)), 4),
                           "tilt_err_deg": round(float(np.abs(prm[:, [2, 4]] - np.array(
                               [spec.tilt_x_deg, spec.tilt_y_deg]).T).max()), 5),
                           "defocus_err": round(float(np.abs(prm[:, 1] - np.array(spec.defocus)).max()), 4)}
        finally:
            lib.digHoloDestroy(h)
    res["mdl_weight"] = mu
    res["workload"] = (f"640x512 uint16 dual-pol, batch {batch}, 45 HG groups, hidden tilt/defocus/waist/centre "
                       f"(seed {seed}), MDL goal (gradient: power + MDL term, weight {mdl_rel} of the power "
                       f"term), both from the nominal start")
    res["gradient_speedup"] = round(res["tweak"]["seconds"] / res["gradient"]["seconds"], 2)
    return res


UNEVEN = [  # frame W, H, FFT window, HG groups, frames
    (320, 256, 128, 10, 3000),
    (640, 512, 192, 30, 1500),
    (1024, 1024, 256, 45, 500),
]


def uneven_batching(lib, reps=20, scale=1.0):
    """Uneven batches: 5000 dual-pol uint16 frames across three handle
    configurations (UNEVEN; `scale` times the frame counts), frames from the
    single-pol generator's settings (200 um waist, 1.5 deg tilt) at each size,
    kernels chosen by the autotuner.  Timed with HIP events: one call for all
    three (digHoloProcessBatchMultiDevice, on the first handle's stream) against
    the sum of the three batches' own times (digHoloProcessBatchDevice each,
    back to back on its handle's stream)."""
    import torch
    from digholo_amd import api, synth
    hs, frames = [], []
    counts = [max(1, int(round(u[4] * scale))) for u in UNEVEN]
    for i, ((W, H, win, G, _), n) in enumerate(zip(UNEVEN, counts)):
        spec = synth.HologramSpec(frame_width=W, frame_height=H, pol_count=2, waist=200e-6, group_count=min(G, 10),
                                  beam_centre_x=[-W * 5e-6, W * 5e-6])
        fr = make_frames(n_unique=10 if W >= 1024 else 25, seed=10 + i, spec=spec, batch=n)
        frames.append(fr)
        h = lib.digHoloCreate()
        api.configure(lib, h, frame_width=W, frame_height=H, pol_count=2, fft_window=(win, win), waist=200e-6,
                      group_count=G, beam_centre_x=(-W * 5e-6, W * 5e-6))
        lib.digHoloConfigSetOverlapPrecision(h, api.DIGHOLO_PRECISION_SPLIT16)
        lib.digHoloConfigSetAutoTune(h, 1)
        assert lib.digHoloSetBatchUint16(h, n, C.c_void_p(fr.data_ptr()), 0) == 0
        hs.append(h)
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    for _ in range(2):          # tunes each shape once, then warm
        for h in hs:
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
            lib.digHoloSynchronize(h)
    seq_ms = 0.0                # each batch on its own, back to back on its stream
    for h in hs:
        st = torch.cuda.ExternalStream(lib.digHoloGetStream(h))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        e1.record(st)
        e1.synchronize()
        seq_ms += e0.elapsed_time(e1) / reps
    ids = (C.c_int * len(hs))(*hs)
    stream = torch.cuda.ExternalStream(lib.digHoloGetStream(hs[0]))   # exists once the handle has run
    for _ in range(2):
        assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
    lib.digHoloSynchronize(hs[0])
    ev[0].record(stream)
    for _ in range(reps):
        assert lib.digHoloProcessBatchMultiDevice(len(hs), ids) == 0
    ev[1].record(stream)
    ev[1].synchronize()
    multi_ms = ev[0].elapsed_time(ev[1]) / reps
    paths = [lib.digHoloGetWindowPath(h) for h in hs]
    for h in hs:
        lib.digHoloDestroy(h)
    total = sum(counts)
    return {"workload": f"{total} dual-pol uint16 frames: {counts[0]} x 320x256/128/10 groups, {counts[1]} x "
                        f"640x512/192/30, {counts[2]} x 1024x1024/256/45, split-fp16 overlap, autotuned kernels",
            "multi_ms": round(multi_ms, 4), "multi_frames_s": round(total / (multi_ms * 1e-3), 1),
            "sequential_ms": round(seq_ms, 4), "sequential_frames_s": round(total / (seq_ms * 1e-3), 1),
            "multi_speedup": round(seq_ms / multi_ms, 3), "window_paths": paths}


LARGE_W, LARGE_WIN, LARGE_G, LARGE_BX = 1024, 512, 80, (-2.56e-3, 2.56e-3)


def large_setup(lib, batch):
    """Handle and resident float32 frames of the 1024x1024, 80-group workload."""
    import torch
    from digholo_amd import api, synth
    spec = synth.HologramSpec(frame_width=LARGE_W, frame_height=LARGE_W, pol_count=2, waist=300e-6,
                              group_count=LARGE_G, beam_centre_x=list(LARGE_BX), group_decay=20.0)
    fr16 = make_frames(n_unique=8, seed=5, spec=spec, batch=batch)
    frames = fr16.to(torch.float32).contiguous()
    del fr16
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=LARGE_W, frame_height=LARGE_W, pol_count=2, fft_window=(LARGE_WIN, LARGE_WIN),
                  waist=300e-6, group_count=LARGE_G, beam_centre_x=LARGE_BX)
    assert lib.digHoloSetBatch(h, batch, C.c_void_p(frames.data_ptr())) == 0
    return h, frames


def traffic_probe_large(lib, batch=256):
    import torch
    from digholo_amd import api
    h, frames = large_setup(lib, batch)
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    for prec in (api.DIGHOLO_PRECISION_FP32, api.DIGHOLO_PRECISION_FP32,
                 api.DIGHOLO_PRECISION_SPLIT16, api.DIGHOLO_PRECISION_SPLIT16):
        assert lib.digHoloConfigSetOverlapPrecision(h, prec) == 0
        assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        assert lib.digHoloSynchronize(h) == 0
    x = torch.empty(128 * 1024 * 1024, dtype=torch.int32, device="cuda").fill_(1)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"traffic probe (large) done, window path {lib.digHoloGetWindowPath(h)}", flush=True)
    lib.digHoloDestroy(h)


def large_sensor(lib, batch=256, steps=10):
    """High-order decomposition: 1024x1024 float32 frames, dual pol, batch 256,
    FFT window 512x512 (IFFT 160x160), 80 HG groups (3240 modes per pol), group
    power ~ exp(-g/20).  Times the FP32 and the split-fp16 overlap (HIP events on
    the handle's stream, stage times from the library) and reports the
    coefficient error of each against the FP64 oracle on two frames."""
    import numpy as np
    import torch
    from digholo_amd import api
    from oracle import digholo_oracle as O
    W, Hh, win, G, bx = LARGE_W, LARGE_W, LARGE_WIN, LARGE_G, LARGE_BX
    h, frames = large_setup(lib, batch)
    b, m, p = C.c_int(), C.c_int(), C.c_int()
    res = {}
    sample = frames[:2].cpu().numpy().astype(np.float64)
    oc = O.Config(frame_width=W, frame_height=Hh, pol_count=2, fft_window_x=win, fft_window_y=win, aperture_deg=0.7,
                  tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=G, waist=[300e-6] * 2, beam_centre_x=list(bx))
    _, c_ref, _ = O.process_batch(sample, oc)
    geo = api.window_geometry(lib, h)
    for name, prec in (("fp32", api.DIGHOLO_PRECISION_FP32), ("split16", api.DIGHOLO_PRECISION_SPLIT16)):
        lib.digHoloConfigSetOverlapPrecision(h, prec)
        for _ in range(2):
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        lib.digHoloSynchronize(h)
        stream = torch.cuda.ExternalStream(lib.digHoloGetStream(h))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # kernel times from the library's stage events, on a second set of batches
        lib.digHoloConfigSetStageTiming(h, 1)
        st = (C.c_float * 5)()
        lib.digHoloGetStageTimes(h, st)
        for _ in range(steps):
            assert lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        lib.digHoloSynchronize(h)
        lib.digHoloGetStageTimes(h, st)
        lib.digHoloConfigSetStageTiming(h, 0)
        coefs = api.get_coefs(lib, h)
        err = O.relative_l2(coefs[:2], c_ref)
        ov_ms = float(st[4])
        flops = overlap_flops_per_frame(2, G, geo["w"], geo["h"]) * batch
        fft_ms = float(st[0] + st[1] + st[2])
        fft_bytes = (2 * win * win * 4 + 2 * geo["w"] * geo["h"] * 8) * batch
        res[name] = {"ms_per_batch": round(ms, 4), "frames_s": round(batch / (ms * 1e-3), 1),
                     "overlap_ms": round(ov_ms, 4),
                     "overlap_tflops": round(flops / (ov_ms * 1e-3) / 1e12, 2) if ov_ms > 0 else None,
                     "overlap_mfma_frac": round(flops / (ov_ms * 1e-3) / 1e12 /
                                                (FP32_PEAK_TFLOPS if prec == 0 else 2500.0), 4) if ov_ms > 0 else None,
                     "fft_stages_ms": round(fft_ms, 4),
                     "fft_stages_hbm_frac": round(fft_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "coef_rel_l2_vs_fp64": float(f"{err:.3e}"), "window_path": lib.digHoloGetWindowPath(h)}
    lib.digHoloDestroy(h)
    res["workload"] = ("1024x1024 float32 dual-pol, batch 256, FFT 512x512 -> IFFT 160x160, 80 HG groups "
                       "(3240 modes/pol), group power ~ exp(-g/20)")
    res["split_vs_fp32_overlap"] = round(res["fp32"]["overlap_ms"] / res["split16"]["overlap_ms"], 3)
    return res


def host_inclusive(lib, frames_dev, reps=3):
    """Frames/s with the frames in host memory and the results on the host:
    digHoloSetBatchUint16 on a host buffer, digHoloProcessBatch (coefficients
    copied back by the call) and digHoloGetFields (fields copied back), timed on
    the host around whole calls.  Pageable memory as a caller's numpy array,
    and page-locked memory allocated for the purpose."""
    import numpy as np
    import torch
    from digholo_amd import api
    res = {}
    for kind in ("pageable", "pinned"):
        if kind == "pageable":
            host = frames_dev.cpu().numpy().view(np.uint16)
            ptr = host.ctypes.data
        else:
            host = torch.empty(frames_dev.shape, dtype=torch.int16, pin_memory=True)
            host.copy_(frames_dev)
            ptr = host.data_ptr()
        h = lib.digHoloCreate()
        try:
            api.configure(lib, h, frame_width=FRAME_W, frame_height=FRAME_H, pol_count=POLS, fft_window=(WIN, WIN),
                          waist=WAIST, group_count=GROUPS, beam_centre_x=BEAM_X)
            lib.digHoloConfigSetOverlapPrecision(h, api.DIGHOLO_PRECISION_SPLIT16)
            b, m, p, w, hh = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
            xp, yp = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()

            def call():
                assert lib.digHoloSetBatchUint16(h, B, C.c_void_p(ptr), 0) == 0
                assert lib.digHoloProcessBatch(h, C.byref(b), C.byref(m), C.byref(p))
                assert lib.digHoloGetFields(h, C.byref(b), C.byref(p), C.byref(xp), C.byref(yp), C.byref(w),
                                            C.byref(hh))
            call()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            dt = (time.perf_counter() - t0) / reps
        finally:
            lib.digHoloDestroy(h)
        res[kind] = {"frames_s": round(B / dt, 1), "ms_per_batch": round(dt * 1e3, 3)}
    res["note"] = ("host uint16 frames (655 MB per batch) copied in, coefficients and fields (119 MB) copied "
                   "out on every call: bound by the host link, not by the kernels")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autoalign", action="store_true",
                    help="skip the gradient-vs-tweak alignment comparison (batch 500)")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the three-configuration graph-launch measurement (5000 frames)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 1024x1024, 80-group decomposition measurement")
    ap.add_argument("--precision", choices=["fp32", "split16"], default="split16")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc: two batches with the fused split overlap, two with the FP32 "
                         "overlap, then a 512 MiB device copy that scales the counters; prints no JSON")
    ap.add_argument("--traffic-probe-large", action="store_true",
                    help="the same for the 1024x1024, 80-group workload (staged kernels): two FP32 batches, "
                         "two split batches, then the 512 MiB copy; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from digholo_amd import api

    lib = api.load()
    torch.cuda.set_device(0)
    if args.traffic_probe_large:
        traffic_probe_large(lib)
        return
    frames = make_frames()
    h = lib.digHoloCreate()
    api.configure(lib, h, frame_width=FRAME_W, frame_height=FRAME_H, pol_count=POLS, fft_window=(WIN, WIN),
                  waist=WAIST, group_count=GROUPS, beam_centre_x=BEAM_X)
    assert lib.digHoloSetBatchUint16(h, B, C.c_void_p(frames.data_ptr()), 0) == 0
    geo = api.window_geometry(lib, h)
    b, m, p = C.c_int(), C.c_int(), C.c_int()

    def step():
        ptr = lib.digHoloProcessBatchDevice(h, C.byref(b), C.byref(m), C.byref(p))
        if not ptr:
            raise RuntimeError("digHoloProcessBatchDevice failed")

    step()
    assert lib.digHoloSynchronize(h) == 0
    if args.traffic_probe:
        for prec in (api.DIGHOLO_PRECISION_SPLIT16, api.DIGHOLO_PRECISION_SPLIT16,
                     api.DIGHOLO_PRECISION_FP32, api.DIGHOLO_PRECISION_FP32):
            assert lib.digHoloConfigSetOverlapPrecision(h, prec) == 0
            step()
            assert lib.digHoloSynchronize(h) == 0
        x = torch.empty(128 * 1024 * 1024, dtype=torch.int32, device="cuda").fill_(1)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"traffic probe done, window path {lib.digHoloGetWindowPath(h)}", flush=True)
        return
    # time on the stream the kernels run on (the handle's own HIP stream)
    stream = torch.cuda.ExternalStream(lib.digHoloGetStream(h))

    def timed(precision, fused=-1):
        """K steps with nothing else on the stream (the rate), then K steps with
        the library's stage events around each kernel (the kernel times)."""
        assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
        assert lib.digHoloConfigSetFusedOverlap(h, fused) == 0
        for _ in range(args.warmup):
            step()
        lib.digHoloSynchronize(h)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            step()
        end.record(stream)
        end.synchronize()
        ms = start.elapsed_time(end) / args.steps
        lib.digHoloConfigSetStageTiming(h, 1)
        stage = (C.c_float * 5)()
        lib.digHoloGetStageTimes(h, stage)          # drop anything collected so far
        for _ in range(args.steps):
            step()
        lib.digHoloSynchronize(h)
        lib.digHoloGetStageTimes(h, stage)
        lib.digHoloConfigSetStageTiming(h, 0)
        return ms, [float(x) for x in stage]

    ms32, st32 = timed(api.DIGHOLO_PRECISION_FP32, 0)            # window kernel + FP32 overlap kernel
    ms16s, st16s = timed(api.DIGHOLO_PRECISION_SPLIT16, 0)       # window kernel + split overlap kernel
    ms16, st16 = timed(api.DIGHOLO_PRECISION_SPLIT16, -1)        # split overlap inside the window kernel
    headline = api.DIGHOLO_PRECISION_SPLIT16 if args.precision == "split16" else api.DIGHOLO_PRECISION_FP32
    if headline == api.DIGHOLO_PRECISION_FP32:
        lib.digHoloConfigSetOverlapPrecision(h, headline)
        lib.digHoloConfigSetFusedOverlap(h, 0)
        step()
    ms_per_step, st = (ms16, st16) if headline == api.DIGHOLO_PRECISION_SPLIT16 else (ms32, st32)
    path = lib.digHoloGetWindowPath(h)

    # the timed batch's own outputs against the FP64 oracle: first, middle, last frame
    lib.digHoloSynchronize(h)
    from oracle import digholo_oracle as O
    picks = [0, B // 2, B - 1]
    sample = frames[picks].cpu().numpy().astype(np.uint16)
    oc = O.Config(frame_width=FRAME_W, frame_height=FRAME_H, pol_count=POLS, fft_window_x=WIN, fft_window_y=WIN,
                  aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], group_count=GROUPS,
                  waist=[WAIST, WAIST], beam_centre_x=list(BEAM_X))
    f_ref, c_ref, _ = O.process_batch(sample, oc)
    coefs_all = api.get_coefs(lib, h)
    fields_all = api.get_fields(lib, h)[0]
    check = {"frames": picks,
             "coef_rel_l2": float(f"{O.relative_l2(coefs_all[picks], c_ref):.3e}"),
             "field_rel_l2": float(f"{O.relative_l2(fields_all[picks], f_ref):.3e}")}
    assert check["coef_rel_l2"] <= 1e-5 and check["field_rel_l2"] <= 1e-5, check

    Wd, Hd, w, hh, M = geo["W"], geo["H"], geo["w"], geo["h"], m.value
    frames_s = B / (ms_per_step * 1e-3)
    fused_overlap = headline == api.DIGHOLO_PRECISION_SPLIT16 and path == 3 and st[4] < 0.05 * st[0]
    # camera window -> fields: the one window kernel when fused (the other two
    # stage intervals are then empty event pairs), else the three staged kernels
    fft_ms = st[0] if path > 0 else st[0] + st[1] + st[2]
    if fused_overlap:   # the window kernel also writes the coefficients
        fft_bytes = algorithmic_bytes_per_frame(POLS, Wd, Hd, w, hh, M) * B
    else:
        fft_bytes = fft_stage_bytes_per_frame(POLS, Wd, Hd, w, hh) * B
    achieved = fft_bytes / (fft_ms * 1e-3) / 1e9
    ov_flops = overlap_flops_per_frame(POLS, GROUPS, w, hh) * B
    kernel = {3: "k_window_v3", 2: "k_window_mfma", 1: "k_fused_window"}.get(
        path, "k_row_fft + k_col_fft_window + k_ifft_rows + k_ifft_cols_tilt")
    traffic = measured_traffic(kernel, fused_overlap)
    ov_bytes = (POLS * w * hh * 8 + POLS * M * 8) * B

    def overlap_stats(ov_ms, peak):
        return {"ms": round(ov_ms, 4), "tflops": round(ov_flops / (ov_ms * 1e-3) / 1e12, 2),
                "mfma_frac": round(ov_flops / (ov_ms * 1e-3) / 1e12 / peak, 4),
                "hbm_frac": round(ov_bytes / (ov_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    out = {
        "metric": METRIC,
        "value": round(frames_s, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+split-f16",
        "dtype_note": "FP32 data and accumulation; the column DFT and the overlap multiply split-fp16 (hi+lo) "
                      "operands on MFMA with FP32 accumulation (22 significant bits per operand)",
        "data": "synthetic (seeded HG holograms, 12-bit uint16 with read noise)",
        "config": {"workload": "640x512 uint16 dual-pol, batch 1000, FFT 256x256 -> IFFT 80x80, "
                               "45 HG groups (1035 modes/pol), frames resident in HBM",
                   "global_batch": B, "overlap_precision": "split16" if headline else "fp32",
                   "fft_window": Wd, "ifft_window": w, "modes_per_pol": M, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm",
                     "kernel": kernel + (" (window + split-fp16 overlap)" if fused_overlap else ""),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": fft_bytes, "avg_launch_ms": round(fft_ms, 4)},
        "timed_batch_vs_fp64": check,
        "fp32_overlap": {"frames_s": round(B / (ms32 * 1e-3), 1), "ms_per_step": round(ms32, 4),
                         "window_ms": round(st32[0], 4), "overlap": overlap_stats(st32[4], FP32_PEAK_TFLOPS)},
        "split16_separate_overlap": {"frames_s": round(B / (ms16s * 1e-3), 1), "ms_per_step": round(ms16s, 4),
                                     "window_ms": round(st16s[0], 4),
                                     "overlap": overlap_stats(st16s[4], 2500.0)},
        "split16_fused_overlap": {"frames_s": round(B / (ms16 * 1e-3), 1), "ms_per_step": round(ms16, 4),
                                  "window_with_overlap_ms": round(st16[0], 4),
                                  "fused_into_window_kernel": fused_overlap},
        "split_vs_fp32_end_to_end": round(ms32 / ms_per_step, 3),
        "split_vs_fp32_overlap_kernel": round(st32[4] / st16s[4], 3) if st16s[4] > 0 else None,
        "stage_ms": {"window": st[0], "col_fft_window": st[1], "ifft_tilt": st[2], "tables": st[3],
                     "overlap": st[4]},
        "end_to_end_hbm_frac": round(algorithmic_bytes_per_frame(POLS, Wd, Hd, w, hh, M) * frames_s / 1e9
                                     / HBM_PEAK_GBS, 4),
    }
    out["host_inclusive"] = host_inclusive(lib, frames)
    lib.digHoloDestroy(h)
    del frames
    torch.cuda.empty_cache()
    if not args.no_autoalign:
        out["autoalign"] = autoalign_comparison(lib)
    if not args.no_uneven:
        # the autotuner remembers its choices for the three shapes across runs here
        os.environ.setdefault("DIGHOLO_TUNE_CACHE", os.path.join(tempfile.gettempdir(), "digholo_tune_cache.txt"))
        out["uneven_batching"] = uneven_batching(lib)
        out["uneven_batching_small"] = uneven_batching(lib, reps=50, scale=0.01)
    if not args.no_large:
        torch.cuda.empty_cache()
        out["large_sensor"] = large_sensor(lib)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
