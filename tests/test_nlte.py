"""Non-LTE source function (calc_nlte_t.f, used by cirsradg_wave.f:931-944).

Pinned by the reference's own cirsradg_wave on the nlte_k1 and nlte_p golden
cases (test_oracle.py runs every case_*.npz), and here by:
* the defining identity: Planck's 1e4/v form at the non-LTE temperature is
  f times its value at T (the invplanck_wave.f inverse), with f from the
  collisional (k1 N / (A + k1 N)) or pressure (p 1e6 / (1 + p 1e6)) form;
* the LTE limits (f -> 1) and flag 0;
* the stop for k1 outside [1e-20, 1e-3].
"""
import copy
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.golden_io import case_cia, load_case

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KB = 1.380658e-23


@pytest.mark.parametrize("t,p,v", [(150.0, 1e-4, 600.0), (110.0, 0.3, 1300.0), (90.0, 1e-6, 900.0)])
def test_nlte_temperature_scales_the_planck_source(t, p, v):
    k1, a = 1e-11, 1e5
    n = p * 101325.0 / (KB * t) * 1e-6
    for flag, f in ((1, k1 * n / (a + k1 * n)), (2, p * 1e6 / (1 + p * 1e6))):
        tn = oracle.nlte_t(flag, k1, a, t, p, v)
        assert tn <= t
        # planck_wave's iwave != 0 form at the same v (the wavenumber index is passed as iwave)
        assert oracle.planck(1, v, tn) == pytest.approx(f * oracle.planck(1, v, t), rel=1e-12)


def test_nlte_lte_limits():
    assert oracle.nlte_t(0, 1e-11, 1e5, 123.4, 1e-3, 700.0) == 123.4
    assert oracle.nlte_t(2, 0.0, 0.0, 150.0, 1e3, 700.0) == pytest.approx(150.0, rel=1e-8)
    assert oracle.nlte_t(1, 1e-3, 1.0, 150.0, 10.0, 700.0) == pytest.approx(150.0, rel=1e-12)


def test_nlte_k1_out_of_range_stops():
    tables, req, _, _ = load_case(os.path.join(HERE, "case_nlte_k1.npz"))
    bad = copy.copy(req)
    bad.nlte_k1 = 1e-2
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, case_cia(os.path.join(HERE, "case_nlte_k1.npz")), bad)


def test_nlte_cases_differ_from_lte():
    for tag, floor in (("nlte_k1", 1e-2), ("nlte_p", 1e-4)):
        path = os.path.join(HERE, f"case_{tag}.npz")
        tables, req, outs, _ = load_case(path)
        lte = copy.copy(req)
        lte.inlte = 0
        o, _, _ = oracle.spectrum(tables, case_cia(path), lte)
        assert np.max(np.abs(o / outs["ref64_output"] - 1)) > floor
