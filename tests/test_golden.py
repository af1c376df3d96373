"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py
from the FP64 oracle on seeded synthetic holograms): the oracle still produces
them (regression pin), and the HIP path meets the parity bar against them
(tests/test_gpu_golden.py)."""
import json
import os

import numpy as np
import pytest

from oracle import digholo_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["c0_single", "dual_small", "c0_fullres"]


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}, json.loads(str(d["cfg"]))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    d, cfg = load(name)
    fields, coefs, plan = O.process_batch(d["frames"], O.Config(**cfg))
    got_plan = np.array([plan.W, plan.H, plan.w, plan.h] + plan.cx + plan.cy + plan.SX + plan.SY)
    assert np.array_equal(got_plan, d["plan"])
    assert O.relative_l2(coefs, d["coefs"]) < 1e-12
    assert O.relative_l2(fields.astype(np.complex64), d["fields"]) < 1e-6
    if coefs.shape[0] > 1:
        assert np.allclose(O.power_meter(coefs), d["metrics"], rtol=1e-9)


def test_golden_recovers_the_rendered_modes():
    # the fixtures are meaningful: coefficients are the rendered ones up to one
    # complex factor per (frame, pol), at the level of 12-bit quantisation
    from digholo_amd import synth
    import sys
    sys.path.insert(0, GOLDEN)
    import make_fixtures as MF
    for name, case in MF.CASES.items():
        d, _ = load(name)
        spec = synth.HologramSpec(**case["spec"])
        c = synth.random_coefs(spec, case["batch"], np.random.default_rng(case["seed"]))
        for b in range(case["batch"]):
            for p in range(spec.pol_count):
                ct, cr = c[b, p], d["coefs"][b, p]
                a = np.vdot(ct, cr) / np.vdot(ct, ct)
                assert np.linalg.norm(cr - a * ct) / np.linalg.norm(cr) < 3e-3, (name, b, p)
