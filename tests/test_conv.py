"""cirsconv.f in the oracle, pinned against the reference and mathematics.

* the restatement equals the reference's own CIRSCONV (REAL(8) build) on
  boxcars inside and past the grid's ends, a wide one, an FWHM table,
  filter functions and fwhm = 0 (``conv.npz``);
* a natural cubic spline reproduces linear functions, so a boxcar over a
  linear spectrum gives its value at the centre, times w/fwhm with an FWHM
  table; filter functions over a constant give the constant; fwhm = 0 copies.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.golden_io import HERE

Z = np.load(os.path.join(HERE, "conv.npz"))


def _filters(z):
    n, v, f, vc = z["filter_n"], z["filter_v"], z["filter_f"], z["filter_vconv"]
    out, o = [], 0
    for k in range(len(n)):
        out.append((vc[k], v[o:o + n[k]], f[o:o + n[k]]))
        o += n[k]
    return out


@pytest.mark.parametrize("name", ["box_inside", "box_ends", "box_wide", "box_table", "identity",
                                  "filter"])
def test_cirsconv_equals_the_reference(name):
    v, y = Z["v"], Z["y"]
    if name == "filter":
        got = oracle.cirsconv(-1.0, v, y, Z["filter_vconv"], filters=_filters(Z))
    else:
        vf = Z[f"{name}_vfwhm"] if f"{name}_vfwhm" in Z.files else None
        xf = Z[f"{name}_xfwhm"] if f"{name}_xfwhm" in Z.files else None
        got = oracle.cirsconv(float(Z[f"{name}_fwhm"]), v, y, Z[f"{name}_vconv"], vf, xf)
    np.testing.assert_allclose(got, Z[f"{name}_ref64"], rtol=1e-13, atol=0)


def test_linear_spectra_and_constant_filters():
    v = np.unique(np.sort(600.0 + 100.0 * np.random.default_rng(4).random(80)))
    a, b = 3.0e-7, -1.7e-9
    y = a + b * v
    vc = np.linspace(608.0, 692.0, 11)
    np.testing.assert_allclose(oracle.cirsconv(6.0, v, y, vc), a + b * vc, rtol=1e-12)
    vf, xf = np.array([500.0, 800.0]), np.array([3.0, 9.0])
    w = np.interp(vc, vf, xf)
    np.testing.assert_allclose(oracle.cirsconv(6.0, v, y, vc, vf, xf), (a + b * vc) * w / 6.0,
                               rtol=1e-12)
    fil = [(c, np.linspace(c - 5, c + 5, 7), np.full(7, 0.3)) for c in vc]
    np.testing.assert_allclose(oracle.cirsconv(-1.0, v, np.full(v.size, 2.5), vc, filters=fil), 2.5,
                               rtol=1e-13)
    np.testing.assert_array_equal(oracle.cirsconv(0.0, v, y, v[:30]), y[:30])
    with pytest.raises(RuntimeError):  # a conv point no filter is centred on
        oracle.cirsconv(-1.0, v, y, np.append(vc, 650.5), filters=fil)
