"""Dust extinction (cirsradg_wave.f:550-577) and its state map in the oracle.

Pinned by the reference's own cirsradg_wave on the dust golden cases
(test_oracle.py runs every case_*.npz) and its map2pro/map2xvec with DCO
(map_dust.npz), and here by:
* interp.f's restatement reproducing linear functions exactly, inside the
  grid and (as the reference extrapolates from the end intervals) outside;
* central differences of the whole model in each dust column, which tau
  depends on linearly, against the Jacobian's dust columns;
* nsec = 1 meaning one cross-section at every wavenumber;
* the reference's stop for a wavenumber off the cross-section grid.
"""
import copy
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oe, oracle
from tests.golden.golden_io import load_case, load_cia
from tests.golden.make_golden_oe import map_inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _interp(x, y, xin):
    f = oracle.lib().orc_interp
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return f(x.ctypes.data, y.ctypes.data, len(x), float(xin))


def test_interp_reproduces_linear_functions():
    rng = np.random.default_rng(3)
    x = np.sort(rng.uniform(500, 1500, 17))
    y = 2.5e-8 * x - 3e-6
    for xin in [x[0], x[5], 0.5 * (x[3] + x[4]), x[-1], 480.0, 1600.0]:
        assert _interp(x, y, xin) == pytest.approx(2.5e-8 * xin - 3e-6, rel=1e-12)
    # a grid value is returned at a grid point, from the interval below it
    z = rng.normal(size=17)
    for i in range(1, 17):
        assert _interp(x, z, x[i]) == pytest.approx(z[i], rel=1e-13, abs=1e-15)


def _dust_case(tag="dust"):
    return load_case(os.path.join(HERE, f"case_{tag}.npz"))


def test_dust_jacobian_matches_finite_differences():
    tables, req, _, _ = _dust_case()
    cia = load_cia()
    o, g, d = oracle.spectrum(tables, cia, req)
    ip, ngas = 2, req.ngas
    for j in (3, 12, 20):
        lay = req.layinc[ip, j] - 1
        for k in range(req.ncont):
            h = 1e-3 * max(req.cont[lay, k], 1e-3 * req.cont[:, k].max())
            rp, rm = copy.copy(req), copy.copy(req)
            rp.cont = req.cont.copy()
            rm.cont = req.cont.copy()
            rp.cont[lay, k] += h
            rm.cont[lay, k] -= h
            fd = (oracle.spectrum(tables, cia, rp)[0] - oracle.spectrum(tables, cia, rm)[0]) / (2 * h)
            col = d[:, ip, :, ngas + 1 + k]
            assert np.max(np.abs(col)) > 0
            assert np.max(np.abs(fd[:, ip] - col[:, j])) < 1e-6 * np.max(np.abs(col))


def test_single_cross_section_is_used_everywhere():
    """nsec = 1: the one value at every wavenumber, so moving the grid point
    changes nothing and the dust columns equal dR/dtau times that value."""
    tables, req, outs, _ = _dust_case("dust_const")
    cia = load_cia()
    moved = copy.copy(req)
    moved.vsec = req.vsec + 123.0
    a = oracle.spectrum(tables, cia, req)
    b = oracle.spectrum(tables, cia, moved)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    # doubling the cross-section and halving the columns keeps tau, and the
    # dust Jacobian column (d/d cont) doubles
    r2 = copy.copy(req)
    r2.xsec = 2 * req.xsec
    r2.cont = 0.5 * req.cont
    o2, _, d2 = oracle.spectrum(tables, cia, r2)
    assert np.allclose(o2, a[0], rtol=1e-13)
    k = req.ngas + 1
    assert np.allclose(d2[..., k], 2 * a[2][..., k], rtol=1e-12, atol=0)


def test_wavenumber_off_the_dust_grid_is_an_error():
    """cirsradg_wave.f:446-452 stops when vwave is outside vsec."""
    tables, req, _, _ = _dust_case()
    cia = load_cia()
    bad = copy.copy(req)
    bad.vsec = req.vsec + (req.vwave[-1] - req.vsec[0]) + 1.0
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, cia, bad)


def test_dust_map_matches_reference():
    """map2pro.f:62-76: dust columns map through DCO (map_dust.npz holds the
    reference's map2pro + map2xvec output)."""
    z = np.load(os.path.join(HERE, "map_dust.npz"), allow_pickle=False)
    m = map_inputs(seed=12, ncont=2)
    for k in ("dq", "dam", "dte", "dco", "xmap"):
        assert np.array_equal(m[k], z[k])
    dx = oe.map2xvec(oe.map2pro(m["dq"], m["nlayin"], m["layinc"], m["dam"], m["dte"], 2,
                                dco=m["dco"]), m["xmap"])
    assert np.max(np.abs(dx - z["doutdx"])) / np.max(np.abs(z["doutdx"])) < 1e-12
