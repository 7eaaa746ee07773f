"""The optimal-estimation oracle (oracle/oe.py) against the reference's own
routines (golden vectors from tests/golden/make_golden_oe.py) and against
what the mathematics fixes."""
import os

import numpy as np
import pytest

from oracle import oe
from tests.golden.make_golden_oe import map_inputs, oe_inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name", ["oe_small", "oe_wide", "oe_medium"])
def test_oe_step_matches_reference(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    r = oe.oe_step(z["kk"], z["sa"], z["se_var"], z["xa"], z["xn"], z["y"], z["yn"])
    assert _rel(r["x_out"], z["x_out"]) < 1e-12
    assert _rel(r["dd"], z["dd"]) < 1e-12
    assert _rel(r["aa"], z["aa"]) < 1e-12
    assert _rel(r["sai"], z["sai"]) < 1e-12
    assert abs(r["phi"] / float(z["phi"]) - 1) < 1e-13
    assert abs(r["chisq"] / float(z["chisq"]) - 1) < 1e-13


def test_oe_step_matches_reference_at_retrieval_shape():
    z = np.load(os.path.join(HERE, "oe_cfg4.npz"), allow_pickle=False)
    inp = oe_inputs(int(z["nx"]), int(z["ny"]), int(z["seed"]))
    r = oe.oe_step(**inp)
    assert _rel(r["x_out"], z["x_out"]) < 1e-11
    assert _rel(r["dd"].sum(axis=1), z["dd_rowsum"]) < 1e-11
    assert _rel(np.diag(r["aa"]), z["aa_diag"]) < 1e-11
    assert abs(r["phi"] / float(z["phi"]) - 1) < 1e-12


def test_map_matches_reference():
    z = np.load(os.path.join(HERE, "map_small.npz"), allow_pickle=False)
    m = map_inputs()
    dx = oe.map2xvec(oe.map2pro(m["dq"], m["nlayin"], m["layinc"], m["dam"], m["dte"], 2),
                     m["xmap"])
    assert _rel(dx, z["doutdx"]) < 1e-13


def test_gain_branches_agree():
    """calc_gain_matrix's two branches are the two sides of the identity
    (Sa^-1 + K^T Se^-1 K)^-1 K^T Se^-1 = Sa K^T (K Sa K^T + Se)^-1."""
    inp = oe_inputs(30, 50, 9)
    kk, sa, se = inp["kk"], inp["sa"], inp["se_var"]
    sai = oe.dinvertm(sa)
    d1, a1 = oe.calc_gain_matrix(kk, sa, sai, np.diag(se), np.diag(1 / se))  # nx < ny
    d2, a2 = oe.calc_gain_matrix(kk[:25], sa, sai, np.diag(se[:25]), np.diag(1 / se[:25]))
    # the nx >= ny branch on the first 25 rows, against the other form built by hand
    kt = kk[:25].T
    d3 = np.linalg.solve(sai + kt @ np.diag(1 / se[:25]) @ kk[:25], kt @ np.diag(1 / se[:25]))
    assert _rel(d2, d3) < 1e-10
    d4 = sa @ kk.T @ np.linalg.inv(kk @ sa @ kk.T + np.diag(se))
    assert _rel(d1, d4) < 1e-10


def test_linear_model_one_step_is_map_estimate():
    """For y = K x (a linear forward model, yn = K xn) one step from any xn
    lands on x^ = xa + (Sa^-1 + K^T Se^-1 K)^-1 K^T Se^-1 (y - K xa)."""
    inp = oe_inputs(20, 60, 5)
    kk, sa, se, xa, y = inp["kk"], inp["sa"], inp["se_var"], inp["xa"], inp["y"]
    sai = np.linalg.inv(sa)
    xhat = xa + np.linalg.solve(sai + kk.T @ (kk / se[:, None]), kk.T @ ((y - kk @ xa) / se))
    for xn in (xa, xa + 1.0, xa - 3.0 * np.arange(20) / 20):
        r = oe.oe_step(kk, sa, se, xa, xn, y, kk @ xn)
        assert _rel(r["x_out"], xhat) < 1e-10


def test_cholesky_factor_and_inverse():
    inp = oe_inputs(50, 10, 6)
    sa = inp["sa"]
    low, p = oe.dcholdc(sa)
    assert _rel(low @ low.T, sa) < 1e-13
    assert np.allclose(np.diag(low), p)
    assert _rel(oe.dinvertm(sa) @ sa, np.eye(50)) < 1e-12


def test_brake_schedule():
    """coreret.f:979-1068: accept and brake x0.3 on a decrease, converge on
    a small decrease with lambda < 1, brake x10 on an increase."""
    assert oe.lm_decide(9.0, 10.0, 1.0, 1, 0.1) == (True, False, 9.0, 0.3)
    assert oe.lm_decide(9.999, 10.0, 0.3, 2, 0.1) == (True, True, 9.999, 0.3)
    assert oe.lm_decide(11.0, 10.0, 0.3, 2, 0.1) == (False, False, 10.0, 3.0)
    assert oe.lm_decide(10.005, 10.0, 0.3, 6, 0.1) == (True, True, 10.005, 0.3)
    assert oe.lm_decide(10.005, 10.0, 0.01, 6, 0.1) == (False, False, 10.0, 0.1)
