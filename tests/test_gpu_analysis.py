"""Plane analysis (BatchGetSummary) and int16 field export (GetFields16) on the
GPU path vs the oracle.

The oracle analyses its own FP64 Fourier planes and fields of the same frames.
Bars: power, centre of mass and effective area within 1e-4 relative (centres
of mass within 1e-4 of the axis extent), the int16 scale MAXABS within 1e-5
relative, the peak index exact, the batch-summed maps within 1e-4 of their
peak; int16 fields within one count of round(V * MAXABS) on the oracle's V.
"""
import numpy as np
import pytest

from digholo_amd import api, synth
from oracle import digholo_oracle as O

from test_gpu_parity import make_case, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return api.load()


def compare(params, idx, maps, x, y, want, want_maps, xw, yw, label):
    np.testing.assert_allclose(x, xw, rtol=1e-6, atol=0)
    np.testing.assert_allclose(y, yw, rtol=1e-6, atol=0)
    ext_x = float(np.abs(xw).max())
    ext_y = float(np.abs(yw).max())
    tol = {O.ANALYSIS_TOTALPOWER: ("rel", 1e-4), O.ANALYSIS_AEFF: ("rel", 1e-4), O.ANALYSIS_MAXABS: ("rel", 1e-5),
           O.ANALYSIS_COMX: ("abs", 1e-4 * ext_x), O.ANALYSIS_COMY: ("abs", 1e-4 * ext_y),
           O.ANALYSIS_COMYWRAP: ("abs", 2e-4 * ext_y)}
    for k, (kind, t) in tol.items():
        got, ref = params[k].astype(np.float64), want[k]
        err = np.abs(got - ref) / (np.abs(ref) if kind == "rel" else 1.0)
        print(f"{label} param {k}: max err {err.max():.3g} (bar {t:.3g})", flush=True)
        assert err.max() <= t, (label, k, got, ref)
    assert np.array_equal(idx, want[O.ANALYSIS_MAXABSIDX].astype(np.int32)), label
    scale = np.abs(want_maps).max()
    assert np.abs(maps - want_maps).max() <= 1e-4 * scale, label


@pytest.mark.parametrize("case", ["small", "dual640", "radix5"])
def test_summary_planes_match_oracle(lib, case):
    frames, cfg = make_case(case, B=5, seed=21)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        _, fields, _, _ = api.process_batch(lib, h, frames)
        oc = oracle_for(cfg)
        f_ref, _, plan = O.process_batch(frames, oc)

        params, idx, maps, x, y = api.batch_summary(lib, h, 1)
        want, want_maps = O.field_plane_analysis(f_ref, oc, plan)
        xw, yw = O.field_axes(oc, plan)
        compare(params, idx, maps, x, y, want, want_maps, xw, yw, f"{case} field plane")

        params, idx, maps, x, y = api.batch_summary(lib, h, 0)
        want, want_maps = O.fourier_plane_analysis(frames, oc)
        Nx = plan.W // 2 + 1
        compare(params, idx, maps, x, y, want, want_maps, O.k_axis_f32(plan.W, oc.pixel_size)[:Nx],
                O.k_axis_f32(plan.H, oc.pixel_size), f"{case} Fourier plane")

        # int16 export: scale = MAXABS of the field plane, values round(V * scale)
        R, I, scale, fx, fy = api.get_fields16(lib, h)
        field_params = api.batch_summary(lib, h, 1)[0]
        np.testing.assert_array_equal(scale, field_params[O.ANALYSIS_MAXABS][:, :-1])
        Rw, Iw = O.fields_int16(f_ref, scale.T)                 # scale.T: [B, pol]
        assert np.abs(R.astype(np.int32) - Rw).max() <= 1 and np.abs(I.astype(np.int32) - Iw).max() <= 1
        back = (R + 1j * I) / scale.T[:, :, None, None]
        assert np.abs(back - fields).max() <= 0.75 / scale.min()
        peak = np.abs(R + 1j * I.astype(np.float64)).max(axis=(2, 3))
        assert np.all(np.abs(peak - 32767) <= 1.5)               # the full int16 range is used
    finally:
        lib.digHoloDestroy(h)


def test_summary_follows_the_latest_batch(lib):
    # a second, smaller batch replaces the analysis of the first
    frames, cfg = make_case("small", B=6, seed=3)
    h = lib.digHoloCreate()
    try:
        api.configure(lib, h, **cfg)
        api.process_batch(lib, h, frames)
        p6 = api.batch_summary(lib, h, 1)[0]
        api.process_batch(lib, h, frames[:2])
        p2 = api.batch_summary(lib, h, 1)[0]
        assert p6.shape[-1] == 7 and p2.shape[-1] == 3
        np.testing.assert_allclose(p2[:, :, :2], p6[:, :, :2], rtol=1e-6)
    finally:
        lib.digHoloDestroy(h)
