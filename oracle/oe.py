"""Optimal-estimation step and layer -> state maps of NEMESIS (TEST
INFRASTRUCTURE: the oracle).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU leg use it, as the checker; the product runs on the
GPU (``radtrancode_amd/csrc/radg_oe.hip``).

FP64 numpy restatement of:

* ``radtran/matrices/dcholdc.f`` / ``dcholsl.f`` / ``dinvertm.f`` (JMOD=2):
  Cholesky inversion as in Numerical Recipes choldc/cholsl;
* ``nemesis/calc_gain_matrix.f``: gain matrix and averaging kernels, both
  branches (nx >= ny and nx < ny);
* ``nemesis/calcnextxn.f``: xn+1 = xa + dd (y - yn) - aa (xa - xn);
* ``nemesis/calc_phiret.f``: the cost function and chi^2;
* ``nemesis/coreret.f:591-1068``: the damped step xn1 = xn + (x_out - xn) /
  (1 + lambda) and the Marquardt brake schedule;
* ``radtran/cirsradg/map2pro.f:49-84`` and ``map2xvec.f:41-54``.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular


def dcholdc(a):
    """dcholdc.f: choldc on the upper triangle of ``a``.  Returns the lower
    factor ``l`` (strictly below the diagonal, as the routine stores it in
    a(j,i), j > i) with its diagonal ``p`` filled in; a non-positive pivot
    is replaced by 1 and the factorisation continues (dcholdc.f:33-48)."""
    a = np.array(a, dtype=np.float64, copy=True)
    n = a.shape[0]
    low = np.zeros_like(a)
    p = np.zeros(n)
    for i in range(n):
        s = a[i, i:] - low[i:, :i] @ low[i, :i]
        p[i] = np.sqrt(s[0]) if s[0] > 0.0 else 1.0
        low[i, i] = p[i]
        low[i + 1:, i] = s[1:] / p[i]
    return low, p


def dcholsl(low, b):
    """dcholsl.f: solve L L^T x = b with the factor of :func:`dcholdc`."""
    y = solve_triangular(low, b, lower=True)
    return solve_triangular(low.T, y, lower=False)


def dinvertm(a):
    """dinvertm.f JMOD=2: inverse by one Cholesky factorisation and a solve
    per column of the identity (dinvertm.f:111-142)."""
    low, _ = dcholdc(a)
    return dcholsl(low, np.eye(a.shape[0]))


def calc_gain_matrix(kk, sa, sai, se, sei):
    """calc_gain_matrix.f:60-150.  ``kk`` [ny, nx]; ``sa``/``sai`` [nx, nx];
    ``se``/``sei`` [ny, ny].  Returns (dd [nx, ny], aa [nx, nx])."""
    kk = np.asarray(kk, dtype=np.float64)
    ny, nx = kk.shape
    kt = kk.T
    if nx >= ny:
        m = sa @ kt                         # :93
        a = kk @ m + se                     # :96-103
        dd = m @ dinvertm(a)                # :111-119
    else:
        m = kt @ sei                        # :126
        c = m @ kk + sai                    # :129-136
        dd = dinvertm(c) @ m                # :143-150
    aa = dd @ kk                            # :155
    return dd, aa


def calcnextxn(xa, xn, y, yn, dd, aa):
    """calcnextxn.f:52-105: x_out = xa + dd (y - yn) - aa (xa - xn)."""
    return xa + (dd @ (y - yn) - aa @ (xa - xn))


def calc_phiret(y, yn, sei, xn, xa, sai):
    """calc_phiret.f:20-80: returns (phi, chisq) with chisq the measurement
    term (yn - y)^T Se^-1 (yn - y) and phi that plus (xn - xa)^T Sa^-1 (xn - xa)."""
    b = yn - y
    phi1 = float(b @ (sei @ b))
    d = xn - xa
    phi2 = float(d @ (sai @ d))
    return phi1 + phi2, phi1


def damped_step(xn, x_out, alambda):
    """coreret.f:631-632: xn1 = xn + (x_out - xn) / (1 + lambda)."""
    return xn + (x_out - xn) / (1.0 + alambda)


def lm_decide(phi, ophi, alambda, it, phlimit):
    """The brake schedule after a trial state (coreret.f:960-1068), without
    the reduced-wavelength and abort-file branches.  Returns (accept,
    converged, new_ophi, new_lambda)."""
    tphi = 100.0 * (ophi - phi) / ophi
    if phi <= ophi:
        if 0.0 <= tphi <= phlimit and alambda < 1.0:
            return True, True, phi, alambda
        return True, False, phi, alambda * 0.3
    if it >= 5 and abs(tphi) <= phlimit:
        if alambda < 0.1:
            return False, False, ophi, min(alambda * 10.0, 1e10)
        return True, True, phi, alambda
    return False, False, ophi, min(alambda * 10.0, 1e10)


def oe_step(kk, sa, se_var, xa, xn, y, yn):
    """One proposal as coreret.f performs it from (xn, yn, kk): Sa^-1 by
    dinvertm, diagonal Se, gain matrix, phi at (xn, yn) and x_out.  Returns
    dict(x_out, dd, aa, sai, phi, chisq)."""
    sai = dinvertm(sa)
    se = np.diag(se_var)
    sei = np.diag(1.0 / se_var)
    dd, aa = calc_gain_matrix(kk, sa, sai, se, sei)
    phi, chisq = calc_phiret(y, yn, sei, xn, xa, sai)
    return dict(x_out=calcnextxn(xa, xn, y, yn, dd, aa), dd=dd, aa=aa, sai=sai, phi=phi,
                chisq=chisq)


def map2pro(dq, nlayin, layinc, dam, dte, ngas, dco=None, dfp=None, ncont=None, imod=None):
    """map2pro.f:49-84.  An imod 8 path (``imod`` [npath], default none) is
    mapped over nlayin(1) rows on path 1's layers (:52-57).  ``dq`` [npath, L, nparam] (layer
    Jacobian, the cirsradg_wave doutputdq order: ngas gases, T, then dust
    types, then the para-H2 fraction), ``layinc`` [npath, L] 1-based,
    ``dam``/``dte``/``dco``/``dfp`` [L, npro]: DAM for gases, DTE for T, DCO
    for the ``ncont`` dust types, DFP after them (map2pro.f:62-80; ``ncont``
    defaults to every remaining column).  Returns doutmoddq [npath, nparam,
    npro]."""
    npath, _, npar = dq.shape
    nc = npar - ngas - 1 if ncont is None else ncont
    out = np.zeros((npath, npar, dam.shape[1]))
    for p in range(npath):
        p1 = 0 if imod is not None and imod[p] == 8 else p
        idx = np.asarray(layinc[p1][:nlayin[p1]]) - 1
        for k in range(npar):
            d = dam if k < ngas else dte if k == ngas else dco if k - ngas - 1 < nc else dfp
            out[p, k] = dq[p, :nlayin[p1], k] @ d[idx]
    return out


def map2xvec(doutmoddq, xmap):
    """map2xvec.f:41-54: doutdx [npath, nv] = sum over (param, level) of
    doutmoddq * xmap, with ``xmap`` [nv, nparam, npro]."""
    return np.einsum("pkj,vkj->pv", doutmoddq, xmap)
