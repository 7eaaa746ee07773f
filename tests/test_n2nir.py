"""N2's near-infrared CIA (n2n2cia.f, n2h2cia.f, nciacon.f:287-296,
nparacon_all.f:312-320) in the oracle.

Pinned by the reference's own cirsradg_wave on the titan_nir and
titan_nir_um golden cases (test_oracle.py runs every case_*.npz), and here
by:
* the reference's N2N2CIA and N2H2CIA at the mid-points of their knots and
  just outside their ranges (n2nir.npz) against verint.f's restatement on
  the knot values;
* verint reproducing linear functions on ascending and descending grids;
* the N2-H2 term acting exactly as the H2-N2 CIA pair (nciacon's alpha(5))
  and the N2-N2 term as the N2-N2 pair (alpha(7)), whose formulas the
  other cases pin: a wrong DABSORB slot or factor fails this;
* the refusal of N2 at 3995-4995 cm-1 without the tables.
"""
import copy
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle
from radtrancode_amd import synth
from tests.golden.golden_io import load_case, load_n2nir

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _verint(x, y, xin):
    f = oracle.lib().orc_verint
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return f(x.ctypes.data, y.ctypes.data, len(x), float(xin))


def _coef(wn, k, v):
    """n2n2cia.f / n2h2cia.f on a knot table: zero outside, verint inside."""
    if v < wn[0] or v > wn[-1]:
        return 0.0
    return _verint(wn, k, v)


def test_tables_reproduce_the_reference_functions():
    z = np.load(os.path.join(HERE, "n2nir.npz"), allow_pickle=False)
    for which in ("n2n2", "n2h2"):
        wn, k = z[f"{which}_wn"], z[f"{which}_k"]
        assert np.all(np.diff(wn) == 5.0) and k.max() > 0
        got = np.array([_coef(wn, k, v) for v in z[f"{which}_probe_wn"]])
        ref = z[f"{which}_probe_k"]
        assert np.max(np.abs(got - ref)) <= 1e-14 * k.max()
        assert np.all(ref[-4:-2] == 0.0)  # just outside the range


def test_verint_is_linear_interpolation():
    rng = np.random.default_rng(5)
    x = np.sort(rng.uniform(0, 10, 12))
    y = 3.0 * x - 2.0
    for xin in (x[0], 0.5 * (x[4] + x[5]), x[-1], -1.0, 12.0):
        assert _verint(x, y, xin) == pytest.approx(3.0 * xin - 2.0, rel=1e-13, abs=1e-12)
        assert _verint(x[::-1], y[::-1], xin) == pytest.approx(3.0 * xin - 2.0, rel=1e-13,
                                                               abs=1e-12)
    # a knot's value is returned at the knot
    z = rng.normal(size=12)
    for i in range(12):
        assert _verint(x, z, x[i]) == pytest.approx(z[i], rel=1e-13, abs=1e-15)


def _nciacon(cia, v, p, t, idgas, amount, pp, xlen=1e5):
    """orc_nciacon for one layer: (absorb, dabsorb[7], iabs[5])."""
    pk = oracle._Pack([], cia, _one_layer_request(idgas))
    f = oracle.lib().orc_nciacon
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    ids = np.asarray(idgas, dtype=np.int32)
    iso = np.zeros_like(ids)
    am = np.asarray(amount, dtype=np.float64)
    ppa = np.asarray(pp, dtype=np.float64)
    ab = np.zeros(1)
    dab = np.zeros(7)
    iabs = np.zeros(5, dtype=np.int32)
    f(C.addressof(pk.cia), v, p, t, 0, len(ids), ids.ctypes.data, iso.ctypes.data,
      am.ctypes.data, ppa.ctypes.data, xlen, ab.ctypes.data, iabs.ctypes.data, dab.ctypes.data)
    return ab[0], dab, iabs


def _one_layer_request(idgas):
    _, _, r = synth.config1(nwave=2, nlayer=2, nkgas=1, ng=4)
    return r


@pytest.mark.parametrize("v", [4100.0, 4512.5, 4700.0])
def test_nir_terms_act_as_the_matching_cia_pairs(v):
    """an2h2 in the N2-H2 slot equals a constant H2-N2 pair coefficient of the
    same value, an2n2 a constant N2-N2 pair coefficient: same absorption and
    the same d(tau)/dq for every gas."""
    c1, c2 = 2.5e-7, 4.0e-8  # N2-H2, N2-N2 (cm-1 amagat-2)
    base = load_n2nir()
    # NIR tables holding constants over the probed wavenumbers
    nir = synth.N2NirTable(wn_n2n2=base.wn_n2h2, k_n2n2=np.full(len(base.wn_n2h2), c2),
                           wn_n2h2=base.wn_n2h2, k_n2h2=np.full(len(base.wn_n2h2), c1))
    dnu = 4.0  # the CIA table reaches 6000 cm-1
    temps = np.linspace(40.0, 400.0, 25)
    zero = np.zeros((1501, 25, 9))
    a_nir = synth.CIATable(dnu=dnu, temps=temps, kcia=zero, n2nir=nir)
    k = zero.copy()
    k[:, :, 4] = c1  # H2-N2
    k[:, :, 6] = c2  # N2-N2
    a_pair = synth.CIATable(dnu=dnu, temps=temps, kcia=k)
    idgas = [6, synth.H2_ID, 22]
    amount = np.array([1e18, 5e21, 4e24])
    pp = np.array([0.001, 0.1, 0.8])
    r1 = _nciacon(a_nir, v, 1.0, 90.0, idgas, amount, pp)
    r2 = _nciacon(a_pair, v, 1.0, 90.0, idgas, amount, pp)
    assert r1[0] > 0
    assert r1[0] == pytest.approx(r2[0], rel=1e-13)
    assert np.allclose(r1[1], r2[1], rtol=1e-13, atol=0)
    assert np.array_equal(r1[2], r2[2])
    # without the tables the terms are absent
    r0 = _nciacon(synth.CIATable(dnu=dnu, temps=temps, kcia=zero), v, 1.0, 90.0, idgas, amount, pp)
    assert r0[0] == 0.0 and np.all(r0[1] == 0.0)


def test_n2_in_the_window_needs_the_tables():
    tables, req, _, _ = load_case(os.path.join(HERE, "case_titan_nir.npz"))
    cia = synth.CIATable(dnu=1.0, temps=np.linspace(40, 400, 25), kcia=np.zeros((1501, 25, 9)))
    with pytest.raises(RuntimeError):
        oracle.spectrum(tables, cia, req)
    # outside 3995-4995 cm-1 N2 runs without them
    out = copy.copy(req)
    out.vwave = req.vwave[-1:]
    oracle.spectrum(tables, cia, out)
