"""The HIP path against the committed golden fixtures (tests/golden/*.npz): fields
and coefficients within 1e-5 relative L2, the window plan exactly, the power-meter
metrics within 1e-3 dB."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from digholo_amd import api
from oracle import digholo_oracle as O
from test_golden import NAMES, load

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def configure(lib, cfg):
    h = lib.digHoloCreate()
    P = cfg["pol_count"]
    api.configure(lib, h, frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                  fft_window=(cfg["fft_window_x"], cfg["fft_window_y"]), aperture_deg=cfg["aperture_deg"],
                  beam_centre_x=tuple(cfg.get("beam_centre_x", (0.0, 0.0))),
                  beam_centre_y=tuple(cfg.get("beam_centre_y", (0.0, 0.0))),
                  tilt_x=tuple(cfg["tilt_x"]), tilt_y=tuple(cfg["tilt_y"]),
                  defocus=tuple(cfg.get("defocus", (0.0, 0.0))), waist=cfg["waist"][0],
                  group_count=cfg["group_count"], fill_factor=cfg["fill_factor"],
                  resolution_mode=cfg["resolution_mode"], pol_lock_tilt=cfg.get("pol_lock_tilt", 0),
                  pol_lock_defocus=cfg.get("pol_lock_defocus", 0))
    return h


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(lib, name, precision):
    d, cfg = load(name)
    h = configure(lib, cfg)
    assert lib.digHoloConfigSetOverlapPrecision(h, precision) == 0
    out = (C.c_int * 12)()
    assert lib.digHoloGetWindowGeometry(h, out) == 0
    P = cfg["pol_count"]
    plan = d["plan"]
    assert (out[0], out[1], out[2], out[3]) == tuple(plan[:4])
    coefs, fields, _, _ = api.process_batch(lib, h, d["frames"])
    ef, ec = O.relative_l2(fields, d["fields"]), O.relative_l2(coefs, d["coefs"])
    print(f"{name} precision {precision} path {lib.digHoloGetWindowPath(h)}: fields {ef:.2e} coefs {ec:.2e}",
          flush=True)
    assert ef <= TOL and ec <= TOL
    if d["coefs"].shape[0] > 1:
        assert lib.digHoloAutoAlignCalcMetrics(h) == 0
        for m in (O.METRIC_IL, O.METRIC_MDL, O.METRIC_DIAG, O.METRIC_SNRAVG):
            got = lib.digHoloAutoAlignGetMetric(h, m)
            ref = 10 * np.log10(d["metrics"][m])
            assert abs(got - ref) < 1e-3, (m, got, ref)
    lib.digHoloDestroy(h)
