"""Gas continua (ngascon.f) in the oracle, pinned against the reference.

* ``hmin_bf``/``hmin_ff`` (H- bound-free and free-free) equal the reference's
  own HMIN_BF/HMIN_FF, each called from a zero sum, and ch4cont/nh3cont its
  CH4CONT/NH3CONT on raddata's tables (``gascon_funcs.npz``);
* the four continuum cases equal the reference (with HMIN's static sum
  emulated, ``test_oracle.py``), and move far from it without the terms, so
  a dropped or misplaced term fails;
* where the absorption is proportional to the gas's amount (H- bound-free,
  CH4, NH3, H2O at fixed partial pressure), central differences in that
  amount equal the Jacobian column absorb/amount of cirsradg_wave.f:607-609;
* the reference's sum carry-over is real: a bound-free-only column gives
  layer j the absorption of j fresh calls;
* branches not restated are refused (IH2O = 2, ozone, UV, electron order).
"""
import copy
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.golden_io import HERE, case_cia, load_case, load_gascon
from tests.tolerance import jac_err, rad_err

GC_CASES = ["hminus_ir", "hminus_nir", "h2o_ckd", "ch4_nh3_vis"]


def _case(name):
    path = os.path.join(HERE, f"case_{name}.npz")
    tables, req, outs, note = load_case(path)
    return tables, case_cia(path), req, outs


def test_hmin_functions_equal_the_reference():
    z = np.load(os.path.join(HERE, "gascon_funcs.npz"))
    bf = np.array([oracle.hmin_bf(v) for v in z["v"]])
    ff = np.array([oracle.hmin_ff(v, t) for v, t in zip(z["v"], z["t"])])
    assert np.count_nonzero(z["ref64_bf"]) > 10 and np.count_nonzero(z["ref64_ff"]) > 10
    np.testing.assert_allclose(bf, z["ref64_bf"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(ff, z["ref64_ff"], rtol=1e-13, atol=0)


def test_ch4_nh3_tables_equal_the_reference():
    z = np.load(os.path.join(HERE, "gascon_funcs.npz"))
    g = load_gascon()
    ch4 = np.array([oracle.tabcont(g, "ch4", v) for v in z["vt"]])
    nh3 = np.array([oracle.tabcont(g, "nh3", v) for v in z["vt"]])
    assert np.count_nonzero(z["ref64_ch4"]) > 10 and np.count_nonzero(z["ref64_nh3"]) > 3
    np.testing.assert_allclose(ch4, z["ref64_ch4"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(nh3, z["ref64_nh3"], rtol=1e-13, atol=0)


def _without_continua(cia, req, name):
    """The same inputs with the gas continua switched off: flags cleared,
    H-, e- and H relabelled to ids no branch of ngascon.f takes."""
    c2, r2 = copy.copy(cia), copy.copy(req)
    c2.gascon = None
    c2.hmin_static = 0
    r2.idgas = req.idgas.copy()
    for a, b in ((91, 92), (80, 81), (48, 49)):
        r2.idgas[r2.idgas == a] = b
    return c2, r2


@pytest.mark.parametrize("name", GC_CASES)
def test_cases_need_the_continua(name):
    tables, cia, req, outs = _case(name)
    c2, r2 = _without_continua(cia, req, name)
    o, g, d = oracle.spectrum(tables, c2, r2)
    assert rad_err(o, outs["ref64_output"]) > 1e-3 or jac_err(d, outs["ref64_doutputdq"]) > 1e-2


@pytest.mark.parametrize("name,gas", [("hminus_nir", 91), ("h2o_ckd", 1), ("ch4_nh3_vis", 6),
                                      ("ch4_nh3_vis", 11)])
def test_continuum_jacobian_is_the_amount_derivative(name, gas):
    tables, cia, req, outs = _case(name)
    # the free-free term reads the H- amount but is credited to H (48) alone
    # (cirsradg_wave.f:607-609 divides each gas's own absorption by its own
    # amount): without e- it is off and bound-free is all H- adds
    req = copy.deepcopy(req)
    req.idgas[req.idgas == 80] = 81
    k = int(np.nonzero(req.idgas == gas)[0][0])
    o, g, d = oracle.spectrum(tables, cia, req)
    for j in (3, 12, 21):
        h = 1e-4
        r1, r2 = copy.deepcopy(req), copy.deepcopy(req)
        a = req.amount[j, k]
        r1.amount[j, k] = a * (1 + h)
        r2.amount[j, k] = a * (1 - h)
        o1, _, _ = oracle.spectrum(tables, cia, r1)
        o2, _, _ = oracle.spectrum(tables, cia, r2)
        fd = (o1 - o2) / (2 * h * a)
        # the path's layer position of layer j (top to bottom)
        m = req.nlayer - 1 - j
        jac = d[:, :, m, k]
        np.testing.assert_allclose(fd, jac, rtol=2e-5, atol=1e-6 * np.abs(jac).max())


def test_reference_sum_carries_over_the_layers():
    """With only bound-free active (no e-) and the H- amount in one layer at
    a time, the reference (static sum) gives layer j (1-based) j times the
    fresh-sum absorption: the emulated carry-over against the plain form."""
    tables, cia, req, outs = _case("hminus_nir")
    req = copy.deepcopy(req)
    req.idgas[req.idgas == 80] = 81
    req.idgas[req.idgas == 48] = 49
    k = int(np.nonzero(req.idgas == 91)[0][0])
    keep = req.amount[:, k].copy()
    for j in (0, 4, 17):
        req.amount[:, k] = 0.0
        req.amount[j, k] = keep[j]
        fresh = copy.copy(cia)
        fresh.hmin_static = 0
        stat = copy.copy(cia)
        stat.hmin_static = 1
        scaled = copy.deepcopy(req)
        scaled.amount[j, k] = keep[j] * (j + 1)
        o_s, _, _ = oracle.spectrum(tables, stat, req)
        o_f, _, _ = oracle.spectrum(tables, fresh, scaled)
        np.testing.assert_allclose(o_s, o_f, rtol=1e-12)


@pytest.mark.parametrize("what", ["ih2o2", "io3", "uv", "order", "no91"])
def test_unrestated_branches_are_refused(what):
    tables, cia, req, outs = _case("h2o_ckd" if what in ("ih2o2", "io3", "uv") else "hminus_ir")
    cia, req = copy.copy(cia), copy.deepcopy(req)
    cia.gascon = copy.copy(cia.gascon) if cia.gascon is not None else load_gascon()
    if what == "ih2o2":
        cia.gascon.ih2o = 2
    elif what == "io3":
        cia.gascon.io3 = 1
        req.idgas[0] = 3
    elif what == "uv":
        cia.gascon.iuvscat = 256  # H2O's UV cross-sections
    elif what == "order":  # H (48) before e- (80)
        ids = req.idgas.copy()
        i80, i48 = int(np.nonzero(ids == 80)[0][0]), int(np.nonzero(ids == 48)[0][0])
        ids[i80], ids[i48] = 48, 80
        req.idgas = ids
    else:  # e- and H without H-
        req.idgas[req.idgas == 91] = 92
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, cia, req)
