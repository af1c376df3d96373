"""LG and custom bases (digHolo.cpp:9469-9605, :9819-10031, :4750-4765) and the
basis fields (:5233-5296) on the GPU path, against the oracle restatement:

* LG coefficients (block-diagonal per HG group) and custom-matrix coefficients
  (the min(M, in) x min(M, out) block of a column-major matrix) within 1e-5
  relative L2 of the oracle's transform of its FP64 HG coefficients;
* the basis fields of the HG, LG and custom bases within 1e-5 of the oracle's;
* the power-meter metrics in the LG basis equal the oracle's power meter on
  its LG coefficients (1e-3 dB; IL and MDL are those of the HG basis, the
  unitary transform keeps the singular values);
* SetBasisType returns the value stored and an unsupported type processes as HG.
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


def case(dual):
    if dual:
        spec = synth.spec_dual_640()
        spec.group_count = 20
        cfg = dict(frame_width=640, frame_height=512, pol_count=2, fft_window=(256, 256), group_count=20,
                   waist=300e-6, beam_centre_x=(-1.6e-3, 1.6e-3))
    else:
        spec = synth.spec_small()
        cfg = dict(frame_width=320, frame_height=256, pol_count=1, fft_window=(128, 128), group_count=10)
    return spec, cfg


def oracle_cfg(cfg):
    P = cfg["pol_count"]
    bx = list(cfg.get("beam_centre_x", (0.0, 0.0)))
    return O.Config(frame_width=cfg["frame_width"], frame_height=cfg["frame_height"], pol_count=P,
                    fft_window_x=cfg["fft_window"][0], fft_window_y=cfg["fft_window"][1], aperture_deg=0.7,
                    tilt_x=[1.5, 1.5], tilt_y=[1.5, 1.5], beam_centre_x=bx, group_count=cfg["group_count"],
                    waist=[cfg.get("waist", 200e-6)] * 2)


def basis_fields(lib, h):
    m, p, w, hh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    xp, yp = api.P_float(), api.P_float()
    ptr = lib.digHoloBasisGetFields(h, C.byref(m), C.byref(p), C.byref(xp), C.byref(yp), C.byref(w), C.byref(hh))
    assert ptr
    n = m.value * p.value * hh.value * w.value
    out = np.ctypeslib.as_array(C.cast(ptr, api.P_float), shape=(2 * n,)).view(np.complex64)
    return out.reshape(m.value, p.value, hh.value, w.value).copy()


@pytest.mark.parametrize("dual", [False, True])
def test_lg_basis(lib, dual):
    spec, cfg = case(dual)
    B = 6
    frames = synth.make_batch(spec, B, seed=4)[0]
    oc = oracle_cfg(cfg)
    _, hg, _ = O.process_batch(frames, oc)
    want = O.lg_transform(hg, cfg["group_count"])
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetBasisTypeLG(h) == 0
        assert lib.digHoloConfigGetBasisType(h) == 1
        coefs, _, _, _ = api.process_batch(lib, h, frames)
        e = O.relative_l2(coefs, want)
        print(f"LG coefficients rel-L2 {e:.2e}", flush=True)
        assert coefs.shape == want.shape and e <= 1e-5
        again = api.get_coefs(lib, h)
        assert np.array_equal(again, coefs)
        F = basis_fields(lib, h)
        Fw = O.basis_fields(oc, O.basis_field_weights(cfg["group_count"], 1))
        ef = O.relative_l2(F, Fw)
        print(f"LG basis fields rel-L2 {ef:.2e}", flush=True)
        assert F.shape == Fw.shape and ef <= 1e-5
        assert lib.digHoloAutoAlignCalcMetrics(h) == 0
        got = np.array([lib.digHoloAutoAlignGetMetric(h, k) for k in range(api.DIGHOLO_METRIC_COUNT)])
        ref = O.metrics_db(O.power_meter(want))
        fin = np.isfinite(ref)
        assert np.allclose(got[fin], ref[fin], atol=1e-3, rtol=0), (got, ref)
    finally:
        lib.digHoloDestroy(h)


@pytest.mark.parametrize("n_in,n_out", [(60, 12), (7, 30)])
def test_custom_basis(lib, n_in, n_out):
    spec, cfg = case(False)
    B = 5
    frames = synth.make_batch(spec, B, seed=2)[0]
    oc = oracle_cfg(cfg)
    _, hg, _ = O.process_batch(frames, oc)
    rng = np.random.default_rng(n_in)
    T = (rng.standard_normal(n_in * n_out) + 1j * rng.standard_normal(n_in * n_out)).astype(np.complex64)
    want = O.custom_transform(hg, T, n_in, n_out)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetBasisTypeCustom(h, n_in, n_out, T.ctypes.data) == 0
        T0 = T.copy()
        T[:] = 0   # the matrix was copied
        coefs, _, _, _ = api.process_batch(lib, h, frames)
        e = O.relative_l2(coefs, want)
        print(f"custom {n_in}x{n_out}: coefficients {coefs.shape} rel-L2 {e:.2e}", flush=True)
        assert coefs.shape == want.shape and e <= 1e-5
        F = basis_fields(lib, h)
        Fw = O.basis_fields(oc, O.basis_field_weights(cfg["group_count"], 2, T0, n_in, n_out))
        assert F.shape == Fw.shape and O.relative_l2(F, Fw) <= 1e-5
    finally:
        lib.digHoloDestroy(h)


def test_hg_basis_fields_and_basis_type_codes(lib):
    spec, cfg = case(False)
    frames = synth.make_batch(spec, 3, seed=1)[0]
    oc = oracle_cfg(cfg)
    _, hg, _ = O.process_batch(frames, oc)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        assert lib.digHoloConfigSetBasisType(h, 2) == 2     # returns the value stored (digHolo.cpp:14736-14742)
        assert lib.digHoloConfigSetBasisType(h, 7) == 7     # unsupported: processed as HG
        coefs, _, _, _ = api.process_batch(lib, h, frames)
        assert O.relative_l2(coefs, hg) <= 1e-5
        F = basis_fields(lib, h)
        Fw = O.basis_fields(oc, O.basis_field_weights(cfg["group_count"], 0))
        assert F.shape == Fw.shape and O.relative_l2(F, Fw) <= 1e-5
    finally:
        lib.digHoloDestroy(h)
