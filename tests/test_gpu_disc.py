"""GPU: disc-integrated emission (imod 18, cirsradg_wave.f:1116-1218).

The golden cases disc, disc_iform3 and disc_nlte_dust (30 and 24 layers,
one layer per lane) run in test_gpu_parity.py against the reference.  Here
the deeper layouts, two, four and eight layers per lane, where the E3 differences
and the dR/dtau suffix sum cross lane boundaries, are compared with the FP64
oracle, with paths that stop part way down; and the refusals.
"""
import copy

import numpy as np
import pytest

from radtrancode_amd import synth
from tests.golden.golden_io import load_cia, load_e3
from tests.tolerance import (TOL_JAC_FP64, TOL_JAC_MIXED, TOL_RAD_FP64, TOL_RAD_MIXED,
                             jac_err, rad_err)

pytestmark = pytest.mark.gpu


def _disc(L, nk, ng, nw):
    tables, cia, req = synth.config1(nwave=nw, nlayer=L, nkgas=nk, ng=ng)
    cia = load_cia()
    cia.e3 = load_e3()
    # semi-transparent: the surface carries 1e-4 to 0.3 of the radiance and the
    # deepest layers' Jacobian rows are a quarter of the largest
    req.amount, req.pp = req.amount * 0.01, req.pp * 0.01
    req.tsurf = 165.0
    req.imod = np.array([18, 3, 18, 0, 18], dtype=np.int32)[:req.npath]
    req.nlayin = req.nlayin.copy()
    req.nlayin[2] = L // 2 + 3  # an imod 18 path that stops part way down
    return tables, cia, req


@pytest.mark.parametrize("shape", [(120, 3, 20, 6), (200, 2, 16, 4), (65, 2, 12, 5),
                                   (300, 1, 8, 2)],
                         ids=lambda s: "L%d_k%d_g%d_w%d" % s)
def test_disc_seeded_against_oracle(gpu_lib, shape):
    from oracle import oracle
    L, nk, ng, nw = shape
    tables, cia, req = _disc(L, nk, ng, nw)
    print(f"disc seeded shape {shape}", flush=True)
    o_ref, g_ref, d_ref = oracle.spectrum(tables, cia, req, jac_mode=1, nthreads=8)
    assert np.all(g_ref[:, [0, 2, 4]] > 0)  # the surface shows on every imod 18 path
    gpu_lib.register_cia(cia)
    try:
        gpu_lib.register_ktables(tables)
        for mode in (0, 1):
            o, g, d = gpu_lib.spectrum(req, mode=mode)
            tr, tj = (TOL_RAD_FP64, TOL_JAC_FP64) if mode == 0 else (TOL_RAD_MIXED, TOL_JAC_MIXED)
            er, eg, ej = rad_err(o, o_ref), rad_err(g, g_ref), jac_err(d, d_ref)
            print(f"  mode {mode}: radiance {er:.2e} gtsurf {eg:.2e} jacobian {ej:.2e}", flush=True)
            assert er < tr and eg < tr and ej < tj
    finally:
        gpu_lib.register_cia(load_cia())


def test_disc_needs_the_e3_table_and_a_surface(gpu_lib):
    """Without /E3TABLE/ the reference would read e3tab.dat; here imod 18 is
    refused with RADG_ENOTABLES.  Its surface term is always added at tsurf,
    so tsurf <= 0 is refused with RADG_EUNSUPPORTED."""
    from radtrancode_amd._native import RadgError
    tables, cia, req = _disc(30, 2, 8, 3)
    gpu_lib.register_ktables(tables)
    gpu_lib.register_cia(load_cia())
    with pytest.raises(RadgError) as e:
        gpu_lib.spectrum(req, mode=0)
    assert e.value.code == 3 and "E3" in str(e.value)
    gpu_lib.register_cia(cia)
    try:
        cold = copy.copy(req)
        cold.tsurf = 0.0
        with pytest.raises(RadgError) as e:
            gpu_lib.spectrum(cold, mode=0)
        assert e.value.code == 2 and "tsurf" in str(e.value)
    finally:
        gpu_lib.register_cia(load_cia())
