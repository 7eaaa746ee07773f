"""k-tables with per-pressure temperature grids (NT < 0) and irregular
wavenumber grids (DELV <= 0) in the oracle.

The golden cases ``kta_*`` (test_oracle.py) hold the oracle to the
reference's own get_kg/read_klist; these pin the new branches by the cases
where they must reduce to the regular ones:
* per-pressure grids that are all the same grid give the one-grid result;
* DELV < 0 on a regular grid at the grid's wavenumbers gives the coincident
  result, and DELV = 0 between its points the two-slice merge of DELV > 0;
* the .kta writer and reader round-trip both forms.
"""
import copy

import numpy as np

from oracle import oracle
from radtrancode_amd import synth
from radtrancode_amd.kta import read_kta, write_kta
from tests.golden.golden_io import load_cia


def _base(nwave=6):
    t = synth.make_ktables(synth.KGAS_IDS[:3], 800.0, 3.0, 12, ng=10, np_=10, nt=6, seed=51)
    r = synth.make_request(nlayer=30, nkgas=3, nwave=nwave, seed=51, table_vmin=800.0,
                           table_delv=3.0, wave_offset=2)
    return t, r


def test_identical_per_pressure_grids_equal_one_grid():
    t, r = _base()
    o0, g0, d0 = oracle.spectrum(t, load_cia(), r)
    t2 = copy.deepcopy(t)
    for kt in t2:
        kt.temp2 = np.tile(kt.temp, (kt.np_, 1))
    o1, g1, d1 = oracle.spectrum(t2, load_cia(), r)
    np.testing.assert_allclose(o1, o0, rtol=1e-13)
    np.testing.assert_allclose(d1, d0, rtol=1e-11, atol=1e-13 * np.abs(d0).max())


def test_irregular_grids_reduce_to_regular():
    t, r = _base()
    cia = load_cia()
    o0, _, d0 = oracle.spectrum(t, cia, r)
    near = copy.deepcopy(t)
    for kt in near:
        kt.vcen = kt.vmin + kt.delv * np.arange(kt.npoint)
        kt.delv = -1.0
    o1, _, d1 = oracle.spectrum(near, cia, r)
    np.testing.assert_array_equal(o1, o0)
    # between table points: DELV = 0 weights the slice below by frack, as
    # DELV > 0 weights it by 1 - (v - v_n)/delv
    rb = copy.deepcopy(r)
    rb.vwave = r.vwave + 0.41 * 3.0
    o2, _, d2 = oracle.spectrum(t, cia, rb)
    zero = copy.deepcopy(near)
    for kt in zero:
        kt.delv = 0.0
    o3, _, d3 = oracle.spectrum(zero, cia, rb)
    np.testing.assert_allclose(o3, o2, rtol=1e-12)
    np.testing.assert_allclose(d3, d2, rtol=1e-10, atol=1e-12 * np.abs(d2).max())


def test_kta_round_trip_of_both_forms(tmp_path):
    t = synth.make_ktables([11], 600.0, 1.0, 5, ng=6, np_=4, nt=3, seed=3)[0]
    t.temp2 = np.array([[50.0, 90.0, 140.0], [60.0, 95.0, 150.0], [70.0, 100.0, 160.0],
                        [80.0, 110.0, 170.0]])
    t.vcen, t.delv = np.array([600.0, 600.7, 602.0, 602.1, 605.0]), 0.0
    for recl in (4, 8):
        p = str(tmp_path / f"t{recl}.kta")
        write_kta(p, t, recl=recl)
        b = read_kta(p, recl=recl)
        rtol = 1e-7 if recl == 4 else 0
        assert b.nt_signed == -3 and b.delv == 0.0
        np.testing.assert_allclose(b.temp2, t.temp2, rtol=rtol)
        np.testing.assert_allclose(b.vcen, t.vcen, rtol=rtol)
        np.testing.assert_allclose(b.k, t.k, rtol=rtol)
