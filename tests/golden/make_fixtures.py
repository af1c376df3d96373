"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no test vectors and cannot be built in this image, so these
fixtures are the oracle's outputs on seeded synthetic holograms: they pin the
oracle against regressions (tests/test_golden.py, CPU) and give the GPU path a
fixed target independent of the oracle code at test time (GPU).

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from digholo_amd import synth  # noqa: E402
from oracle import digholo_oracle as O  # noqa: E402

CASES = {
    # C0 geometry (320x256 single pol, FFT 128 -> IFFT 48, 10 HG groups), two frames
    "c0_single": dict(
        spec=dict(frame_width=320, frame_height=256, pol_count=1, waist=200e-6, group_count=10),
        cfg=dict(frame_width=320, frame_height=256, pol_count=1, fft_window_x=128, fft_window_y=128,
                 aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], waist=[200e-6, 200e-6],
                 group_count=10, resolution_mode=1, fill_factor=0),
        batch=2, seed=101),
    # dual pol with locks off, defocus, off-centre beams, 256x128 frames (FFT 128 -> IFFT 48)
    "dual_small": dict(
        spec=dict(frame_width=256, frame_height=128, pol_count=2, waist=150e-6, group_count=8,
                  beam_centre_x=[30e-6, -40e-6], beam_centre_y=[20e-6, -10e-6], tilt_x_deg=[1.5, 1.4],
                  tilt_y_deg=[1.5, -1.6], defocus=[0.2, -0.3]),
        cfg=dict(frame_width=256, frame_height=128, pol_count=2, fft_window_x=128, fft_window_y=128,
                 aperture_deg=0.7, tilt_x=[1.5, 1.4], tilt_y=[1.5, -1.6], defocus=[0.2, -0.3],
                 waist=[150e-6, 150e-6], group_count=8, beam_centre_x=[-1.28e-3 + 30e-6, 1.28e-3 - 40e-6],
                 beam_centre_y=[20e-6, -10e-6], resolution_mode=1, pol_lock_tilt=0, pol_lock_defocus=0,
                 fill_factor=0),
        batch=3, seed=102),
    # full-resolution fields of the C0 geometry, pixel-integrated frames (fill-factor correction on)
    "c0_fullres": dict(
        spec=dict(frame_width=320, frame_height=256, pol_count=1, waist=200e-6, group_count=6, pixel_quadrature=4),
        cfg=dict(frame_width=320, frame_height=256, pol_count=1, fft_window_x=128, fft_window_y=128,
                 aperture_deg=0.7, tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], waist=[200e-6, 200e-6],
                 group_count=6, resolution_mode=0, fill_factor=1),
        batch=1, seed=103),
}


def render(case):
    spec = synth.HologramSpec(**case["spec"])
    if spec.pol_count == 2:
        # beams in the middle of each pol's half of the frame
        half = spec.frame_width * spec.pixel_size / 2
        spec.beam_centre_x = [-half / 2 + case["spec"]["beam_centre_x"][0], half / 2 + case["spec"]["beam_centre_x"][1]]
    frames, _, _, _ = synth.make_batch(spec, case["batch"], seed=case["seed"])
    return frames


def oracle_cfg(cfg):
    return O.Config(**cfg)


def main():
    for name, case in CASES.items():
        if case["spec"]["pol_count"] == 2:   # the configured centres are the rendered ones
            half = case["spec"]["frame_width"] * 20e-6 / 2
            assert all(abs(a - b) < 1e-12 for a, b in zip(
                case["cfg"]["beam_centre_x"],
                [-half / 2 + case["spec"]["beam_centre_x"][0], half / 2 + case["spec"]["beam_centre_x"][1]]))
        frames = render(case)
        cfg = oracle_cfg(case["cfg"])
        fields, coefs, plan = O.process_batch(frames, cfg)
        metrics = O.power_meter(coefs) if coefs.shape[0] > 1 else np.zeros(9)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, frames=frames, coefs=coefs, fields=fields.astype(np.complex64), metrics=metrics,
                            plan=np.array([plan.W, plan.H, plan.w, plan.h] + plan.cx + plan.cy + plan.SX + plan.SY),
                            cfg=np.array(json.dumps(case["cfg"])))
        print(f"{name}: frames {frames.shape} coefs {coefs.shape} fields {fields.shape} -> "
              f"{os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
