"""The overlap rebin's two forms on the GPU.

Each random-overlap round (noverlapg.f, rankg.f:104-152) is rebinned by the
row-count form wherever the new gas's g-ordinates are non-decreasing and no
element's weight reaches past two bin edges, and by the record form
otherwise (``RADG_REBIN=records`` keeps every round on the record form).
Both are held to the FP64 oracle and to each other at g-ordinate counts
that take every slots-per-lane layout of the sort (E = 4, 8 and 16), with
descending g-ordinates (the record form); weight sets whose heaviest
element could span two bins are refused at registration.
"""
import os

import numpy as np
import pytest

from radtrancode_amd import synth
from tests.golden.golden_io import load_cia
from tests.tolerance import (TOL_JAC_FP64, TOL_JAC_MIXED, TOL_RAD_FP64, TOL_RAD_MIXED, jac_err,
                             rad_err)

pytestmark = pytest.mark.gpu


def _spectrum(rad, tables, req, mode, records):
    old = os.environ.get("RADG_REBIN")
    if records:
        os.environ["RADG_REBIN"] = "records"
    else:
        os.environ.pop("RADG_REBIN", None)
    try:
        rad.register_ktables(tables)
        return rad.spectrum(req, mode=mode)
    finally:
        if old is None:
            os.environ.pop("RADG_REBIN", None)
        else:
            os.environ["RADG_REBIN"] = old


def _check(rad, tables, req, tight):
    from oracle import oracle
    o_ref, g_ref, d_ref = oracle.spectrum(tables, load_cia(), req, nthreads=8)
    for mode in (0, 1):
        tr, tj = (TOL_RAD_FP64, TOL_JAC_FP64) if mode == 0 else (TOL_RAD_MIXED, TOL_JAC_MIXED)
        oc, gc, dc = _spectrum(rad, tables, req, mode, records=False)
        orr, grr, drr = _spectrum(rad, tables, req, mode, records=True)
        ec, ejc = rad_err(oc, o_ref), jac_err(dc, d_ref)
        er, ejr = rad_err(orr, o_ref), jac_err(drr, d_ref)
        ex, ejx = rad_err(oc, orr), jac_err(dc, drr)
        print(f"  mode {mode}: counts {ec:.2e}/{ejc:.2e} records {er:.2e}/{ejr:.2e} "
              f"counts vs records {ex:.2e}/{ejx:.2e}", flush=True)
        assert ec < tr and ejc < tj
        assert er < tr and ejr < tj
        # the two forms sum the same terms in another order
        lim_r, lim_j = tight if mode == 0 else (1e-5, 1e-4)
        assert ex < lim_r and ejx < lim_j


# (g-ordinates, k-gases): NG 12 and 16 sort 4 slots per lane, 17-21 take 8,
# 24 and 32 take 16
@pytest.mark.parametrize("shape", [(12, 3), (16, 2), (20, 3), (20, 6), (21, 4), (24, 3), (32, 2)],
                         ids=lambda s: "g%d_k%d" % s)
def test_counts_and_records_against_oracle(gpu_lib, shape):
    ng, nk = shape
    tables, cia, req = synth.config1(nwave=5, nlayer=48, nkgas=nk, ng=ng)
    print(f"rebin shape ng={ng} nkgas={nk}", flush=True)
    _check(gpu_lib, tables, req, tight=(1e-12, 1e-9))


def test_descending_ordinates_take_records(gpu_lib):
    """One gas with its g-ordinates stored in descending order of k: its
    rounds fall back to the record form (an ascending-row layout no longer
    holds), which still matches the oracle."""
    tables, cia, req = synth.config1(nwave=5, nlayer=48, nkgas=3, ng=20)
    tables[1].k = np.ascontiguousarray(tables[1].k[..., ::-1])
    print("descending ordinates in gas 1", flush=True)
    _check(gpu_lib, tables, req, tight=(1e-12, 1e-9))


def test_heavy_weights_refused(gpu_lib):
    """g-weights with max(delg)^2 >= min(delg) are refused at registration:
    the heaviest element could then reach past two bin edges, where rankg.f
    closes one bin per element.  Below that bound no element spans two edges
    (the row-count form still checks and hands such a round back)."""
    from radtrancode_amd._native import RadgError
    tables, cia, req = synth.config1(nwave=4, nlayer=32, nkgas=3, ng=12)
    delg = np.full(12, 0.4 / 11)
    delg[0] = 0.6
    for t in tables:
        t.delg = delg.copy()
        t.g_ord = np.cumsum(delg) - 0.5 * delg
    with pytest.raises(RadgError):
        gpu_lib.register_ktables(tables)
