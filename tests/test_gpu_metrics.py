"""AutoAlign power-meter metrics on the GPU path vs the oracle power meter.

The library computes the row reductions and the FP64 Gram matrix on the GPU and
the extreme eigenvalues on the host; the oracle uses numpy's SVD on the same
coefficients (read back through digHoloBasisGetCoefs), so this isolates the
metric arithmetic.  Bar: 1e-3 dB on every metric.
"""
import ctypes as C

import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def gpu_metrics(lib, h):
    assert lib.digHoloAutoAlignCalcMetrics(h) == 0
    return np.array([lib.digHoloAutoAlignGetMetric(h, m) for m in range(9)], dtype=np.float64)


def run_case(lib, spec, cfg, B, independent=0, mul_conj=0, seed=3):
    frames, _, _, _ = synth.make_batch(spec, B, seed=seed)
    h = lib.digHoloCreate()
    api.configure(lib, h, **cfg)
    lib.digHoloConfigSetAutoAlignPolIndependence(h, independent)
    lib.digHoloConfigSetAutoAlignBasisMulConjTrans(h, mul_conj)
    coefs, _, _, _ = api.process_batch(lib, h, frames)
    got = gpu_metrics(lib, h)
    want = O.metrics_db(O.power_meter(coefs, independent=bool(independent), mul_conj=bool(mul_conj)))
    print("gpu ", np.round(got, 4), "\noracle", np.round(want, 4), flush=True)
    finite = np.isfinite(want)
    assert np.array_equal(finite, np.isfinite(got))
    assert np.allclose(got[finite], want[finite], atol=1e-3, rtol=0)
    return got


def small():
    spec = synth.spec_small()
    cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), waist=200e-6,
               group_count=10)
    return spec, cfg


def dual(G=10):
    spec = synth.spec_dual_640()
    spec.group_count = G
    cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), waist=300e-6,
               group_count=G, beam_centre_x=(-1.6e-3, 1.6e-3))
    return spec, cfg


def test_metrics_batch_smaller_than_modes(lib):
    spec, cfg = dual()
    run_case(lib, spec, cfg, B=40)                        # 2 x 55 modes > 40 frames: 40 x 40 Gram


def test_metrics_batch_larger_than_modes(lib):
    spec, cfg = small()
    run_case(lib, spec, cfg, B=80)                        # 55 modes < 80 frames: 55 x 55 Gram


def test_metrics_independent_pols_and_mulconj(lib):
    spec, cfg = dual()
    run_case(lib, spec, cfg, B=30, independent=1)
    run_case(lib, spec, cfg, B=30, mul_conj=1)


def test_metrics_single_frame(lib):
    spec, cfg = dual()
    got = run_case(lib, spec, cfg, B=1)
    assert got[O.METRIC_MDL] == -np.inf


def test_metrics_large_gram(lib):
    # 2 x 1035 modes, 300 frames: the host eigen-solve on a 300 x 300 Gram
    spec, cfg = dual(G=45)
    run_case(lib, spec, cfg, B=300)
