"""The line-by-line path (/LBLTABLE/ ILBL = 2): the oracle pinned against the
reference and against the mathematics, the .lta format, the host checks.

* every lbl_* golden case: the restatement (get_klblg + noverlapg1) equals
  the reference's own cirsradg_wave built with REAL promoted to 8 bytes and
  run on .lta files through read_klbllist, in radiance, gtsurf and every gas
  column of the Jacobian.  The T column is left out: the reference's is
  undefined (noverlapg1.f:108 stores the first pair's dK/dT into grad, later
  gases read the first pair's derivatives);
* the restatement's T column against central differences of its radiance;
* noverlapg1 by hand: its zero branches, including the zero running sum that
  zeroes earlier gas columns (noverlapg1.f:123-128);
* one gas (where the reference's loop never runs) against the same gas
  beside a second of zero amount;
* .lta write/read round trip, and the host's refusals.
"""
import copy
import os

import numpy as np
import pytest

from oracle import oracle
from radtrancode_amd import kta, synth
from tests.golden.golden_io import case_cia, lbl_case_paths, load_case
from tests.tolerance import jac_err, rad_err

LBL = lbl_case_paths()


def _gas_cols(d, ngas):
    keep = np.ones(d.shape[-1], bool)
    keep[ngas] = False
    return d[..., keep]


def test_lbl_cases_present():
    names = {os.path.basename(p) for p in LBL}
    assert {"lbl_basic.npz", "lbl_t2.npz", "lbl_between.npz", "lbl_zeros.npz",
            "lbl_last_table.npz"} <= names


@pytest.mark.parametrize("path", LBL, ids=[os.path.basename(p)[4:-4] for p in LBL])
def test_oracle_lbl_matches_reference_fp64(path):
    tables, req, outs, note = load_case(path)
    o, g, d = oracle.spectrum(tables, case_cia(path), req, ilbl=2)
    assert rad_err(o, outs["ref64_output"]) < 1e-12, note
    assert rad_err(g, outs["ref64_gtsurf"]) < 1e-12, note
    assert jac_err(_gas_cols(d, req.ngas), _gas_cols(outs["ref64_doutputdq"], req.ngas)) < 1e-12, note


def test_lbl_between_case_reaches_the_grid_edges():
    """The between case holds a wavenumber just below vmin (INT truncates
    (v + 0.02 delv - vmin) / delv toward zero: slice 1) and one past the last
    point by less than delv."""
    path = [p for p in LBL if p.endswith("lbl_between.npz")][0]
    tables, req, _, _ = load_case(path)
    t = tables[0]
    x = (req.vwave + 0.02 * t.delv - t.vmin) / t.delv
    assert (x < 0).any() and (x > t.npoint - 1).any()
    assert np.all((x > -1) & (x < t.npoint))


def _copy(req):
    r = copy.copy(req)
    for a in ("amount", "temp", "pp", "emtemp"):
        setattr(r, a, getattr(req, a).copy())
    return r


def test_lbl_temperature_column_matches_finite_differences():
    """d(radiance)/d(T) of the restatement (k through get_klblg's dk/dT summed
    by noverlapg1's branches, the continuum and the Planck terms) against
    central differences, layer temperature and emission temperature moving
    together; thin and warm so every layer shows."""
    tables = synth.make_lbltables(synth.KGAS_IDS[:3], 800.0, 0.5, 40, np_=10, nt=6, seed=41)
    req = synth.make_request(nlayer=12, nkgas=3, nwave=4, seed=41, table_vmin=800.0,
                             table_delv=2.5, wave_offset=1)
    req.amount = req.amount * 1e-3
    req.pp = req.pp * 1e-3
    cia = synth.make_cia()
    o, g, d = oracle.spectrum(tables, cia, req, ilbl=2)
    ip, ngas = 2, req.ngas
    for j in (0, 4, 9):
        lay = req.layinc[ip, j] - 1
        h = 1e-3
        rp, rm = _copy(req), _copy(req)
        for rr, s in ((rp, h), (rm, -h)):
            rr.temp[lay] += s
            rr.emtemp = rr.temp[rr.layinc - 1]
        fd = (oracle.spectrum(tables, cia, rp, ilbl=2)[0] - oracle.spectrum(tables, cia, rm, ilbl=2)[0]) / (2 * h)
        scale = np.max(np.abs(d[:, ip, :, ngas]))
        assert np.max(np.abs(fd[:, ip] - d[:, ip, j, ngas])) < 2e-3 * scale
        for k in range(3):  # and the gas columns (steps well above the radiance's ulp)
            ha = 1e-2 * req.amount[lay, k]
            rp, rm = _copy(req), _copy(req)
            rp.amount[lay, k] += ha
            rm.amount[lay, k] -= ha
            fd = (oracle.spectrum(tables, cia, rp, ilbl=2)[0]
                  - oracle.spectrum(tables, cia, rm, ilbl=2)[0]) / (2 * ha)
            sc = np.max(np.abs(d[:, ip, :, k]))
            assert np.max(np.abs(fd[:, ip] - d[:, ip, j, k])) < 1e-3 * sc


def test_noverlapg1_by_hand():
    a = np.array([2.0, 3.0, 5.0])
    k = np.array([1.5, 0.25, 4.0])
    dk = np.array([0.1, -0.2, 0.3])
    kg, dq = oracle.noverlapg1(a, k, dk)
    assert kg == 1.5 * 2.0 + 0.25 * 3.0 + 4.0 * 5.0
    assert np.array_equal(dq[:3], k)
    assert np.isclose(dq[3], 0.1 * 2.0 - 0.2 * 3.0 + 0.3 * 5.0, rtol=0, atol=1e-15)
    # first gas zero (:85-90): the sum starts at gas 2
    kg, dq = oracle.noverlapg1(a, [0.0, 0.25, 4.0], dk)
    assert kg == 0.25 * 3.0 + 4.0 * 5.0 and dq[0] == 0.0
    assert np.isclose(dq[3], -0.2 * 3.0 + 0.3 * 5.0, rtol=0, atol=1e-15)
    # second gas zero (:94-99): its column 0, its dk/dT not added
    kg, dq = oracle.noverlapg1(a, [1.5, 0.0, 4.0], dk)
    assert kg == 1.5 * 2.0 + 4.0 * 5.0 and dq[1] == 0.0
    assert np.isclose(dq[3], 0.1 * 2.0 + 0.3 * 5.0, rtol=0, atol=1e-15)
    # later gas zero (:134-138): the sum and its derivative unchanged
    kg, dq = oracle.noverlapg1(a, [1.5, 0.25, 0.0], dk)
    assert kg == 1.5 * 2.0 + 0.25 * 3.0 and dq[2] == 0.0
    assert np.isclose(dq[3], 0.1 * 2.0 - 0.2 * 3.0, rtol=0, atol=1e-15)
    # zero running sum before gas 3 (first gas of zero amount, second with
    # no k): the columns before gas 3 are zeroed, gas 1's k included
    kg, dq = oracle.noverlapg1([0.0, 3.0, 5.0], [1.5, 0.0, 4.0], dk)
    assert kg == 4.0 * 5.0
    assert np.array_equal(dq[:3], [0.0, 0.0, 4.0])
    assert np.isclose(dq[3], 0.3 * 5.0, rtol=0, atol=1e-15)


def test_lbl_one_gas_equals_two_with_zero_amount():
    """With one gas noverlapg1's loop never runs (k_g keeps an earlier
    call's value); the restatement takes amount * k, which is what the
    two-gas form (pinned by the reference) gives when the second gas has no
    amount."""
    tables = synth.make_lbltables(synth.KGAS_IDS[:2], 800.0, 0.5, 30, np_=8, nt=5, seed=42)
    req = synth.make_request(nlayer=20, nkgas=2, nwave=3, seed=42, table_vmin=800.0,
                             table_delv=2.0, wave_offset=1)
    cia = synth.make_cia()
    one = _copy(req)
    keep = [0, 2, 3]  # the first k-gas, H2, He
    one.idgas, one.isogas = req.idgas[keep], req.isogas[keep]
    one.amount, one.pp = req.amount[:, keep].copy(), req.pp[:, keep].copy()
    one.isw = np.arange(1, len(keep) + 2, dtype=np.int32)
    two = _copy(req)
    two.amount[:, 1] = 0.0
    o1, g1, d1 = oracle.spectrum(tables, cia, one, ilbl=2)
    o2, g2, d2 = oracle.spectrum(tables, cia, two, ilbl=2)
    assert np.array_equal(o1, o2)
    assert np.array_equal(d1[..., 0], d2[..., 0])
    assert np.array_equal(d1[..., 3], d2[..., 4])  # T


def test_oracle_refuses_ilbl1():
    tables = synth.make_lbltables(synth.KGAS_IDS[:1], 800.0, 0.5, 10, np_=4, nt=3)
    req = synth.make_request(nlayer=5, nkgas=1, nwave=2, table_vmin=800.0, table_delv=0.5)
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, synth.make_cia(), req, ilbl=1)
    req.vwave = np.array([800.0 - 0.6])  # below slice 1's reach
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, synth.make_cia(), req, ilbl=2)


@pytest.mark.parametrize("recl", [4, 8])
@pytest.mark.parametrize("t2", [False, True])
def test_lta_round_trip(tmp_path, recl, t2):
    rng = np.random.default_rng(3)
    temp2 = np.sort(rng.uniform(50, 300, (4, 5)), axis=1) if t2 else None
    t = kta.lbl_table(23, 1, 500.0, 0.01, np.geomspace(1e-4, 10, 4), np.linspace(60, 250, 5),
                      rng.uniform(0, 1, (7, 4, 5)), temp2=temp2)
    p = str(tmp_path / "x.lta")
    kta.write_lta(p, t, recl)
    u = kta.read_lta(p, recl)
    assert (u.idgas, u.isogas, u.npoint, u.np_, u.nt_signed) == (23, 1, 7, 4, -5 if t2 else 5)
    tol = 0 if recl == 8 else 1e-7
    assert np.allclose(u.k, t.k, rtol=tol, atol=0)
    assert np.allclose(u.press, t.press, rtol=tol, atol=0)
    if t2:
        assert np.allclose(u.temp2, temp2, rtol=tol, atol=0)
    # IREC0 = 10 + NP + NT + 2 (calc_lbltable.f:232-236) and the file ends
    # with the table
    raw = np.fromfile(p, dtype=np.uint8).reshape(-1, recl)
    ntemp = 20 if t2 else 5
    assert raw[0, :4].view(np.int32)[0] == 12 + 4 + ntemp
    assert raw.shape[0] == 11 + 4 + ntemp + 7 * 20
