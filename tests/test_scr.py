"""SCR radiometer paths (imod 8, 13, 14; cirsradg_wave.f:1037-1102) in the
oracle.

Pinned by the reference's own cirsradg_wave on the scr, scr_trans and
scr_rows golden cases (test_oracle.py runs every case_*.npz; scr_rows holds
the nlayin(1) row count of cirsradg_wave.f:1251-1252), and here by:
* a cell of zero optical depth: the wideband cell (1 + e^0)/2 = 1 makes the
  imod 8 product equal its atmosphere path in output, gtsurf and Jacobian,
  and the sideband cell 1 - e^0 = 0 makes it zero;
* a cell's own output against 1 - e^-tau and (1 + e^-tau)/2 of its layer,
  with tau taken from an imod 0 path through the same layer;
* the refusals where the reference would read values it never wrote.
"""
import copy

import numpy as np
import pytest

from oracle import oracle
from radtrancode_amd import synth
from tests.golden.golden_io import load_cia


def _scr(atm=3, cells=(14, 13), cell_scale=(0.0, 0.0)):
    t, c, r = synth.config1(nwave=4, nlayer=20, nkgas=2, ng=8)
    r.amount, r.pp = r.amount * 0.01, r.pp * 0.01
    r.tsurf = 150.0
    r.vem = np.array([500.0, 2000.0])
    r.emissivity = np.array([0.9, 0.95])
    r.imod = np.array([atm, cells[0], 8, cells[1], 8], dtype=np.int32)
    r.nlayin = np.array([20, 1, 2, 1, 2], dtype=np.int32)
    lay, sc = r.layinc.copy(), r.scale.copy()
    lay[1, 0], sc[1, 0] = 3, cell_scale[0]
    lay[3, 0], sc[3, 0] = 5, cell_scale[1]
    lay[2, :2] = [2, 1]
    lay[4, :2] = [4, 1]
    r.layinc, r.scale = lay, sc
    return t, load_cia(), r


def test_transparent_cells():
    t, c, r = _scr()
    o, g, d = oracle.spectrum(t, c, r)
    # wideband cell of tau 0: the product is the atmosphere itself
    assert np.allclose(o[:, 1], 1.0, rtol=1e-14)  # the g-weights' sum
    assert np.allclose(o[:, 2], o[:, 0], rtol=1e-15, atol=0)
    assert np.allclose(g[:, 2], g[:, 0], rtol=1e-15, atol=0)
    assert np.abs(g[:, 0]).max() > 0
    assert np.allclose(d[:, 2], d[:, 0], rtol=1e-15, atol=0)
    # sideband cell of tau 0: nothing passes
    assert np.all(o[:, 3] == 0.0) and np.all(o[:, 4] == 0.0)
    assert np.all(g[:, 4] == 0.0) and np.all(d[:, 4] == 0.0)
    # cells carry no Jacobian
    assert np.all(d[:, [1, 3]] == 0.0)


def test_cell_transmissions():
    """A cell's output is the g-sum of 1 - e^-tau (sideband) or
    (1 + e^-tau)/2 (wideband); an imod 0 path through the same scaled layer
    gives the g-sum of e^-tau."""
    t, c, r = _scr(cell_scale=(1.0, 1.0))
    r2 = copy.deepcopy(r)
    r2.imod = np.array([r.imod[0], 0, 8, 0, 8], dtype=np.int32)
    r2.imod[2] = r2.imod[4] = 0  # the products' slots made plain transmissions
    r2.nlayin = np.array([20, 1, 1, 1, 1], dtype=np.int32)
    r2.layinc = r.layinc.copy()
    r2.layinc[2, 0], r2.layinc[4, 0] = r.layinc[1, 0], r.layinc[3, 0]
    r2.scale = r.scale.copy()
    r2.scale[2, 0], r2.scale[4, 0] = r.scale[1, 0], r.scale[3, 0]
    o, _, _ = oracle.spectrum(t, c, r)
    tr, _, _ = oracle.spectrum(t, c, r2)
    assert np.all((tr[:, 1] > 0.05) & (tr[:, 1] < 0.99))
    assert np.allclose(o[:, 1], 0.5 * (1.0 + tr[:, 1]), rtol=1e-14)  # wideband
    assert np.allclose(o[:, 3], 1.0 - tr[:, 3], rtol=1e-12)          # sideband


@pytest.mark.parametrize("what", ["later_cell", "not_a_cell", "nested", "short_atm"])
def test_reads_of_unwritten_values_are_refused(what):
    t, c, r = _scr()
    lay, nl, im = r.layinc.copy(), r.nlayin.copy(), r.imod.copy()
    if what == "later_cell":
        lay[2, :2] = [4, 1]   # path 4 comes after path 3
    elif what == "not_a_cell":
        lay[2, :2] = [1, 1]   # path 1 is the atmosphere: the reference stops
    elif what == "nested":
        lay[4, :2] = [4, 3]   # path 3 is itself a product
    else:
        im[1], nl[1] = 0, 5   # path 2 as a 5-layer atmosphere of the product...
        lay[2, :2] = [4, 2]   # ...with cell 4 after it: also out of order
        im[3] = 14
    r.layinc, r.nlayin, r.imod = lay, nl, im
    with pytest.raises(RuntimeError):
        oracle.spectrum(t, c, r)


def test_scr_map_matches_reference():
    """map2pro.f:52-57: an imod 8 path is mapped over nlayin(1) rows on path
    1's layers (map_scr.npz holds the reference's map2pro + map2xvec)."""
    import os

    from oracle import oe
    from tests.golden.make_golden_oe import map_inputs
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "map_scr.npz"),
                allow_pickle=False)
    m = map_inputs(seed=14, scr=True)
    for key in ("dq", "dam", "dte", "xmap", "imod", "nlayin", "layinc"):
        assert np.array_equal(m[key], z[key])
    dx = oe.map2xvec(oe.map2pro(m["dq"], m["nlayin"], m["layinc"], m["dam"], m["dte"], 2,
                                imod=m["imod"]), m["xmap"])
    assert np.max(np.abs(dx - z["doutdx"])) / np.max(np.abs(z["doutdx"])) < 1e-12
