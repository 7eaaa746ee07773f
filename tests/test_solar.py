"""The stellar-spectrum forms (/SOLARDAT/ with iread = 999,
get_solar_wave.f, cirsradg_wave.f:862-880, 912-925, 1144-1149) in the oracle.

Pinned by the reference's own cirsradg_wave on the solar_ratio and
solar_trans golden cases (test_oracle.py runs every case_*.npz), and here by
closed forms with the spectrum interpolated by numpy:
* iform 1: every imod 3 output and Jacobian entry is the one without the
  star divided by the star's total power at that wavenumber; imod 0 and 1
  paths are untouched;
* iform 4: every imod 0 output and Jacobian entry is the one without the
  star times its flux at dist AU, y / (4 pi (dist AU)^2), or y for
  dist <= 0, and zero off the spectrum; other paths are untouched.
"""
import copy
import os

import numpy as np

from oracle import oracle
from radtrancode_amd import synth
from tests.golden.golden_io import case_cia, load_case
from tests.tolerance import jac_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AU, PI = 1.49598e13, 3.1415927  # constdef.f


def _case(tag):
    path = os.path.join(HERE, f"case_{tag}.npz")
    t, r, _, _ = load_case(path)
    return t, r, case_cia(path)


def _without_star(c):
    c0 = copy.copy(c)
    c0.solar = None
    return c0


def test_iform1_divides_emission_by_the_star():
    t, r, c = _case("solar_ratio")
    o, g, d = oracle.spectrum(t, c, r)
    o0, g0, d0 = oracle.spectrum(t, _without_star(c), r)
    y = np.interp(r.vwave, c.solar.wave, c.solar.rad)
    for p in range(r.npath):
        f = 1.0 / y if r.imod[p] == 3 else np.ones_like(y)
        assert np.allclose(o[:, p], o0[:, p] * f, rtol=1e-13, atol=0)
        assert jac_err(d[:, p], d0[:, p] * f[:, None, None]) < 1e-13


def test_iform4_scales_transmission_by_the_flux():
    t, r, c = _case("solar_trans")
    for dist in (r.dist, 0.0):
        rr = copy.copy(r)
        rr.dist = dist
        o, g, d = oracle.spectrum(t, c, rr)
        o0, g0, d0 = oracle.spectrum(t, _without_star(c), rr)
        y = np.interp(r.vwave, c.solar.wave, c.solar.rad)
        if dist > 0:
            y = y / (4 * PI * (dist * AU) ** 2)
        for p in range(r.npath):
            f = y if r.imod[p] == 0 else np.ones_like(y)
            assert np.allclose(o[:, p], o0[:, p] * f, rtol=1e-13, atol=0)
            assert jac_err(d[:, p], d0[:, p] * f[:, None, None]) < 1e-13
    # off the spectrum the flux is zero
    c2 = copy.copy(c)
    c2.solar = synth.SolarSpectrum(wave=c.solar.wave + 2000.0, rad=c.solar.rad)
    o, _, d = oracle.spectrum(t, c2, r)
    assert np.all(o[:, r.imod == 0] == 0.0) and np.all(d[:, r.imod == 0] == 0.0)
    assert np.all(o[:, r.imod == 3] > 0)
