"""Para-H2 CIA (FLAGH2P = 1: nparacon_all.f, fpread_all.f) in the oracle.

Pinned by the reference's own cirsradg_wave on the parah2 and parah2_dust
golden cases (test_oracle.py runs every case_*.npz) and its map2pro with DFP
(map_para.npz), and here by:
* a para table whose every fraction holds the normal CIA table's H2-H2 and
  H2-He pairs (and whose five other pairs are that table's): the result must
  equal the FLAGH2P = 0 (equilibrium, INORMAL = 0) one, with a zero fpara
  column;
* central differences in each layer's para fraction, inside the fraction
  grid, against the fpara Jacobian column (the interpolation is linear in
  fpara there, so the difference quotient is exact up to roundoff).
"""
import copy
import os

import numpy as np
import pytest

from oracle import oe, oracle
from radtrancode_amd import synth
from tests.golden.golden_io import case_cia, load_case
from tests.golden.make_golden_oe import map_inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case(tag):
    path = os.path.join(HERE, f"case_{tag}.npz")
    tables, req, outs, note = load_case(path)
    return tables, req, outs, case_cia(path)


def test_flat_para_table_reduces_to_nciacon():
    tables, req, _, cia = _case("parah2")
    k = cia.kcia
    npara = 8
    flat = synth.ParaCIATable(
        dnu=cia.dnu, temps=cia.temps, fp=np.linspace(0.25, 1.0, npara),
        kh2h2=np.repeat(k[:, :, 0:1], npara, axis=2), kh2he=np.repeat(k[:, :, 1:2], npara, axis=2),
        kcia5=k[:, :, 4:9].copy())
    c2 = synth.CIATable(dnu=cia.dnu, temps=cia.temps, kcia=k, para=flat)
    o1, g1, d1 = oracle.spectrum(tables, c2, req)
    plain = copy.copy(req)
    plain.flagh2p = 0
    plain.inormal = 0
    plain.isw = req.isw[req.isw <= req.ngas + 1]
    o0, g0, d0 = oracle.spectrum(tables, c2, plain)
    assert np.allclose(o1, o0, rtol=1e-13, atol=0)
    assert np.allclose(g1, g0, rtol=1e-13, atol=0)
    n = req.ngas + 1
    assert np.allclose(d1[..., :n], d0, rtol=1e-11, atol=1e-13 * np.abs(d0).max())
    assert np.all(d1[..., n] == 0.0)


def test_para_fraction_column_matches_finite_differences():
    tables, req, _, cia = _case("parah2")
    o, g, d = oracle.spectrum(tables, cia, req)
    ip, k = 1, req.nparam - 1
    fp = cia.para.fp
    for j in (3, 10, 15, 25):  # layers away from the case's edge values
        lay = req.layinc[ip, j] - 1
        x = req.hfp[lay]
        assert fp[0] < x < fp[-2]  # inside the grid's interpolated range
        # stay inside the layer's fraction interval
        i = int((x - fp[0]) / (fp[1] - fp[0]))
        h = 0.25 * min(x - fp[i], fp[i + 1] - x)
        rp, rm = copy.copy(req), copy.copy(req)
        rp.hfp = req.hfp.copy()
        rm.hfp = req.hfp.copy()
        rp.hfp[lay] += h
        rm.hfp[lay] -= h
        fd = (oracle.spectrum(tables, cia, rp)[0] - oracle.spectrum(tables, cia, rm)[0]) / (2 * h)
        col = d[:, ip, :, k]
        assert np.max(np.abs(col)) > 0
        assert np.max(np.abs(fd[:, ip] - col[:, j])) < 1e-6 * np.max(np.abs(col))


def test_para_case_depends_on_the_fraction():
    tables, req, outs, cia = _case("parah2")
    moved = copy.copy(req)
    moved.hfp = np.full_like(req.hfp, 0.5)
    o, _, _ = oracle.spectrum(tables, cia, moved)
    assert np.max(np.abs(o / outs["ref64_output"] - 1)) > 1e-4


def test_para_map_matches_reference():
    """map2pro.f:77-80: the para-H2 column maps through DFP."""
    z = np.load(os.path.join(HERE, "map_para.npz"), allow_pickle=False)
    m = map_inputs(seed=13, ncont=1, para=True)
    for key in ("dq", "dam", "dte", "dco", "dfp", "xmap"):
        assert np.array_equal(m[key], z[key])
    dx = oe.map2xvec(oe.map2pro(m["dq"], m["nlayin"], m["layinc"], m["dam"], m["dte"], 2,
                                dco=m["dco"], dfp=m["dfp"], ncont=1), m["xmap"])
    assert np.max(np.abs(dx - z["doutdx"])) / np.max(np.abs(z["doutdx"])) < 1e-12
