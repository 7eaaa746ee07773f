"""Driver for the reference's own optimal-estimation step and layer -> state
maps (TEST INFRASTRUCTURE).

``oracle/Makefile`` (target ``ref``) compiles ``nemesis/calc_gain_matrix.f``,
``calcnextxn.f``, ``calc_phiret.f``, the ``radtran/matrices`` routines they
call and ``radtran/cirsradg/map2pro.f`` / ``map2xvec.f``, where they lie and
with REAL promoted to 8 bytes, together with this build's harness
``oracle/oe_harness.f`` into ``oracle/_ref/liboeref64.so``.  Used only to
write the golden fixtures under ``tests/golden/``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_ref", "liboeref64.so")


def available() -> bool:
    return os.path.exists(LIB)


_lib = None


def _load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB, mode=os.RTLD_LAZY)
    return _lib


def _f(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _i(v):
    return C.byref(C.c_int(int(v)))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def oe_step(kk, sa, se_var, xa, xn, y, yn):
    """HOESTEP: Sa^-1 (dinvertm), gain matrix, phi at (xn, yn), x_out.
    ``kk`` [ny, nx]."""
    ny, nx = kk.shape
    if nx > 401 or ny > 2048:
        raise ValueError("reference limits: nx <= MAXV = 401, ny <= IDIM = 2048")
    kkf, saf = _f(kk), _f(sa)
    sef, xaf, xnf, yf, ynf = (_f(v) for v in (se_var, xa, xn, y, yn))
    xout = np.zeros(nx, order="F")
    dd = np.zeros((nx, ny), order="F")
    aa = np.zeros((nx, nx), order="F")
    sai = np.zeros((nx, nx), order="F")
    phi, chisq, ichk = C.c_double(0), C.c_double(0), C.c_int(0)
    _load().hoestep_(_i(nx), _i(ny), _p(kkf), _p(saf), _p(sef), _p(xaf), _p(xnf), _p(yf),
                     _p(ynf), _p(xout), _p(dd), _p(aa), _p(sai), C.byref(phi),
                     C.byref(chisq), C.byref(ichk))
    return dict(x_out=np.array(xout), dd=np.array(dd), aa=np.array(aa), sai=np.array(sai),
                phi=phi.value, chisq=chisq.value, icheck=ichk.value)


def map_state(dq, nlayin, layinc, dam, dte, xmap, dco=None, ncont=0, dfp=None, imod=None):
    """HMAP: map2pro then map2xvec for one wavenumber.  ``dq`` [npath, L,
    nparam] with nparam = ngas + 1 + ncont (+ 1 with ``dfp``), ``layinc``
    [npath, L] 1-based, ``dam``/``dte``/``dco``/``dfp`` [L, npro], ``xmap``
    [nv, nparam, npro], ``imod`` [npath] (default 3; 8 maps over path 1's
    layers).  Returns doutdx [npath, nv]."""
    npath, L, npar = dq.shape
    npro = dam.shape[1]
    nv = xmap.shape[0]
    nl = _f(nlayin, np.int32)
    lay = _f(np.asarray(layinc).T, np.int32)          # LAYC(NLAY, NPTH)
    im = _f(np.full(npath, 3) if imod is None else imod, np.int32)
    nfp = 0 if dfp is None else 1
    dco = np.zeros_like(dam) if dco is None else dco
    dfp = np.zeros_like(dam) if dfp is None else dfp
    damf, dtef, dcof, dfpf = _f(dam), _f(dte), _f(dco), _f(dfp)
    xmf, dqf = _f(xmap), _f(dq)
    dx = np.zeros((npath, nv), order="F")
    _load().hmap_(_i(npath), _i(L), _i(npar - 1 - ncont - nfp), _i(ncont), _i(nfp), _i(npro),
                  _i(nv), _p(nl), _p(lay), _p(im), _p(damf), _p(dtef), _p(dcof), _p(dfpf), _p(xmf),
                  _p(dqf), _p(dx))
    return np.array(dx)
