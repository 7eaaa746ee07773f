"""Rayleigh extinction (cirsradg_wave.f:663-691) in the oracle, pinned by:

* the reference's own RAYLEIGHJ/V/A/LS (radtran/scatter/) built as REAL(8),
  on 64 random (v, P, T, mixture) points each (tests/golden/rayleigh.npz);
* the whole forward model on the rayleigh_* golden cases (test_oracle.py
  runs every case_*.npz);
* the lambda^-4 law in the long-wavelength limit, and the published air
  cross-section at 550 nm for the Earth-air form;
* the gate: iray > 0 with itype = 0 is the same as iray = 0;
* central differences of the model in the H2 and He amounts, whose Jacobian
  columns are Rayleigh alone above the CIA table's 1500 cm-1.
"""
import copy
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.golden_io import load_case, load_cia
from tests.tolerance import jac_err, rad_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _vectors():
    return np.load(os.path.join(HERE, "rayleigh.npz"), allow_pickle=False)


@pytest.mark.parametrize("iray", [1, 2, 3, 4])
def test_rayleigh_functions_match_reference(iray):
    z = _vectors()
    got = np.array([oracle.rayleigh(iray, z["v"][i], z["p"][i], z["t"][i], *z["mix"][i])
                    for i in range(len(z["v"]))])
    ref64 = z[f"ref64_iray{iray}"]
    assert np.all(ref64 > 0)
    assert np.max(np.abs(got / ref64 - 1)) < 1e-13
    # The reference as written (REAL = 4 bytes) agrees to single precision,
    # except that rayleighls.f forms nr**2 - 1 with nr - 1 ~ 1e-4 in REAL(4):
    # a cancellation that leaves about 4e-4.
    tol32 = 1e-3 if iray == 4 else 1e-5
    assert np.max(np.abs(z[f"ref32_iray{iray}"] / ref64 - 1)) < tol32


@pytest.mark.parametrize("iray", [1, 2, 3, 4])
def test_rayleigh_lambda4_law(iray):
    """sigma(2v) / sigma(v) = 16 once dispersion vanishes (v -> 0); the
    dispersive forms rise faster than v^4 in the visible."""
    lo = oracle.rayleigh(iray, 20.0, 1.0, 273.15, 0.16, 0.002, 0.0007)
    hi = oracle.rayleigh(iray, 40.0, 1.0, 273.15, 0.16, 0.002, 0.0007)
    assert abs(hi / lo - 16.0) < (1e-12 if iray == 2 else 1e-4)
    if iray != 2:
        a = oracle.rayleigh(iray, 10000.0, 1.0, 273.15, 0.16, 0.002, 0.0007)
        b = oracle.rayleigh(iray, 20000.0, 1.0, 273.15, 0.16, 0.002, 0.0007)
        assert b / a > 16.0


def test_rayleigh_air_cross_section_at_550nm():
    """Air at 1 atm and 288.15 K, 550 nm: about 4.5e-27 cm2 per molecule
    (Bucholtz 1995, Appl. Opt. 34, 2765; Bodhaine et al. 1999)."""
    s = oracle.rayleigh(3, 1e4 / 0.55, 1.0, 288.15)
    assert abs(s / 4.5e-27 - 1) < 0.05


@pytest.mark.parametrize("tag", ["j", "ls"])
def test_rayleigh_gated_by_itype(tag):
    tables, req, _, _ = load_case(os.path.join(HERE, f"case_rayleigh_{tag}.npz"))
    cia = load_cia()
    off = copy.copy(req)
    off.iray = 0
    gate = copy.copy(req)
    gate.itype = 0
    o0, g0, d0 = oracle.spectrum(tables, cia, off)
    o1, g1, d1 = oracle.spectrum(tables, cia, gate)
    o2, _, d2 = oracle.spectrum(tables, cia, req)
    assert np.array_equal(o0, o1) and np.array_equal(g0, g1) and np.array_equal(d0, d1)
    # and the extinction is visible: radiance falls, H2/He columns appear
    assert np.all(o2 < o0)
    ngas = req.ngas
    assert np.all(d0[..., ngas - 2:ngas] == 0.0)
    assert np.all(np.max(np.abs(d2[..., ngas - 2:ngas]), axis=-2) > 0)


@pytest.mark.parametrize("tag", ["j", "v"])
def test_rayleigh_jacobian_matches_finite_differences(tag):
    """iray 1 and 2 depend on the wavenumber only, so d(tau)/d(amount) =
    xray * 1e-20 is the exact derivative; no CIA above 1500 cm-1."""
    tables, req, _, _ = load_case(os.path.join(HERE, f"case_rayleigh_{tag}.npz"))
    cia = load_cia()
    o, g, d = oracle.spectrum(tables, cia, req)
    ngas, ip = req.ngas, 1
    # Layers from mid-column down: near the top a layer's Rayleigh share of
    # the radiance is ~1e-7, and the difference quotient is roundoff there.
    for j in (11, 18, 25):
        lay = req.layinc[ip, j] - 1
        for k in (ngas - 2, ngas - 1):  # H2, He
            h = 1e-2 * req.amount[lay, k]
            rp, rm = copy.copy(req), copy.copy(req)
            rp.amount = req.amount.copy()
            rm.amount = req.amount.copy()
            rp.amount[lay, k] += h
            rm.amount[lay, k] -= h
            fd = (oracle.spectrum(tables, cia, rp)[0] - oracle.spectrum(tables, cia, rm)[0]) / (2 * h)
            scale = np.max(np.abs(d[:, ip, :, k]))
            assert scale > 0
            assert np.max(np.abs(fd[:, ip] - d[:, ip, j, k])) < 1e-6 * scale


def test_rayleigh_cases_differ_from_no_rayleigh():
    """The golden cases are sensitive to the term: without it the H2/He
    columns vanish (no CIA above 1500 cm-1) and the radiance moves."""
    cia = load_cia()
    for tag in ("j", "v", "a", "ls"):
        tables, req, outs, note = load_case(os.path.join(HERE, f"case_rayleigh_{tag}.npz"))
        off = copy.copy(req)
        off.iray = 0
        o0, _, d0 = oracle.spectrum(tables, cia, off)
        assert rad_err(o0, outs["ref64_output"]) > 1e-3, note
        assert jac_err(d0, outs["ref64_doutputdq"]) == pytest.approx(1.0), note
