"""ctypes driver for the FP64 CPU restatement (TEST INFRASTRUCTURE).

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module, as the checker.  ``spectrum`` runs the reference algorithm
(``cirsrtfg_wave.f:361-448`` looping ``cirsradg_wave``) over a request.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libradg_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "libradg_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        _lib.orc_cirsradg_wave.restype = C.c_int
        _lib.orc_spectrum.restype = C.c_int
        _lib.orc_planck.restype = C.c_double
        _lib.orc_planck.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_planckg.restype = C.c_double
        _lib.orc_planckg.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_rayleigh.restype = C.c_double
        _lib.orc_rayleigh.argtypes = [C.c_int] + [C.c_double] * 6
        _lib.orc_locate.restype = C.c_int
        _lib.orc_locate.argtypes = [dp, C.c_int, C.c_double]
    return _lib


class _KT(C.Structure):
    _fields_ = [("idgas", C.c_int), ("isogas", C.c_int), ("npoint", C.c_int),
                ("np", C.c_int), ("nt", C.c_int), ("ng", C.c_int),
                ("vmin", C.c_double), ("delv", C.c_double),
                ("lnp", C.c_void_p), ("tk", C.c_void_p), ("g_ord", C.c_void_p),
                ("delg", C.c_void_p), ("k", C.c_void_p), ("t2", C.c_void_p),
                ("vcen", C.c_void_p)]


class _CIA(C.Structure):
    _fields_ = [("dnu", C.c_double), ("temps", C.c_void_p), ("kcia", C.c_void_p),
                ("npara", C.c_int), ("dnu2", C.c_double), ("temps2", C.c_void_p),
                ("fp", C.c_void_p), ("kh2h2", C.c_void_p), ("kh2he", C.c_void_p),
                ("kcia5", C.c_void_p), ("e3tran", C.c_void_p), ("e3val", C.c_void_p),
                ("n1", C.c_int), ("n2", C.c_int), ("wn1", C.c_void_p), ("k1", C.c_void_p),
                ("wn2", C.c_void_p), ("k2", C.c_void_p), ("nsol", C.c_int),
                ("swave", C.c_void_p), ("srad", C.c_void_p),
                ("ih2o", C.c_int), ("ich4", C.c_int), ("io3", C.c_int), ("inh3", C.c_int),
                ("iuvscat", C.c_int), ("nh2o", C.c_int), ("h2o_v0", C.c_double),
                ("h2o_dv", C.c_double), ("h2o_c2", C.c_double), ("h2o_ts", C.c_double),
                ("h2o_ps", C.c_double), ("h2o_s296", C.c_void_p), ("h2o_s260", C.c_void_p),
                ("h2o_f296", C.c_void_p), ("nch4", C.c_int), ("nnh3", C.c_int),
                ("ch4_wl", C.c_void_p), ("ch4_k", C.c_void_p), ("nh3_wl", C.c_void_p),
                ("nh3_k", C.c_void_p), ("hmin_static", C.c_int)]


class _REQ(C.Structure):
    _fields_ = [("nlayer", C.c_int), ("npath", C.c_int), ("ngas", C.c_int),
                ("nsw", C.c_int), ("nem", C.c_int), ("inormal", C.c_int),
                ("iray", C.c_int), ("ispace", C.c_int), ("itype", C.c_int),
                ("flagh2p", C.c_int), ("amform", C.c_int),
                ("dist", C.c_double), ("tsurf", C.c_double),
                ("radius1", C.c_double), ("radextra", C.c_double),
                ("press", C.c_void_p), ("temp", C.c_void_p), ("delh", C.c_void_p),
                ("hfp", C.c_void_p), ("pp", C.c_void_p), ("amount", C.c_void_p),
                ("idgas", C.c_void_p), ("isogas", C.c_void_p),
                ("nlayin", C.c_void_p), ("imod", C.c_void_p), ("layinc", C.c_void_p),
                ("scale", C.c_void_p), ("emtemp", C.c_void_p), ("isw", C.c_void_p),
                ("vem", C.c_void_p), ("emissivity", C.c_void_p),
                ("ncont", C.c_int), ("nsec", C.c_int),
                ("cont", C.c_void_p), ("vsec", C.c_void_p), ("xsec", C.c_void_p),
                ("inlte", C.c_int), ("nlte_k1", C.c_double), ("nlte_a", C.c_double),
                ("iform", C.c_int), ("ilbl", C.c_int)]


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class _Pack:
    """Keeps the contiguous copies alive while the C structs point at them."""

    def __init__(self, tables, cia, req, ilbl=0):
        self.keep = []
        kts = (_KT * max(1, len(tables)))()
        for t, kt in enumerate(tables):
            arrs = [_d(np.log(kt.press)), _d(kt.temp), _d(kt.g_ord), _d(kt.delg), _d(kt.k)]
            extra = [None if x is None else _d(x) for x in (getattr(kt, "temp2", None),
                                                             getattr(kt, "vcen", None))]
            self.keep += arrs + [x for x in extra if x is not None]
            kts[t] = _KT(kt.idgas, kt.isogas, kt.npoint, kt.np_,
                         kt.nt_signed if hasattr(kt, "nt_signed") else kt.nt, kt.ng,
                         kt.vmin, kt.delv, *[a.ctypes.data for a in arrs],
                         *[None if x is None else x.ctypes.data for x in extra])
        self.kts, self.ntab = kts, len(tables)
        ct, kc = _d(cia.temps), _d(cia.kcia)
        self.keep += [ct, kc]
        pa = getattr(cia, "para", None)
        if pa is None:
            self.cia = _CIA(cia.dnu, ct.ctypes.data, kc.ctypes.data, 0, 0.0)
        else:
            arrs = [_d(pa.temps), _d(pa.fp), _d(pa.kh2h2), _d(pa.kh2he), _d(pa.kcia5)]
            self.keep += arrs
            self.cia = _CIA(cia.dnu, ct.ctypes.data, kc.ctypes.data, len(pa.fp), pa.dnu,
                            *[a.ctypes.data for a in arrs])
        e3 = getattr(cia, "e3", None)
        if e3 is not None:
            t3, v3 = _d(e3.trans), _d(e3.e3)
            self.keep += [t3, v3]
            self.cia.e3tran, self.cia.e3val = t3.ctypes.data, v3.ctypes.data
        nir = getattr(cia, "n2nir", None)
        if nir is not None:
            arrs = [_d(nir.wn_n2n2), _d(nir.k_n2n2), _d(nir.wn_n2h2), _d(nir.k_n2h2)]
            self.keep += arrs
            self.cia.n1, self.cia.n2 = len(arrs[0]), len(arrs[2])
            self.cia.wn1, self.cia.k1, self.cia.wn2, self.cia.k2 = [a.ctypes.data for a in arrs]
        sol = getattr(cia, "solar", None)
        if sol is not None:
            sw, sr = _d(sol.wave), _d(sol.rad)
            self.keep += [sw, sr]
            self.cia.nsol, self.cia.swave, self.cia.srad = len(sw), sw.ctypes.data, sr.ctypes.data
        gc = getattr(cia, "gascon", None)
        if gc is not None:
            c = self.cia
            c.ih2o, c.ich4, c.io3, c.inh3, c.iuvscat = (int(gc.ih2o), int(gc.ich4), int(gc.io3),
                                                         int(gc.inh3), int(gc.iuvscat))
            c.h2o_v0, c.h2o_dv, c.h2o_c2, c.h2o_ts, c.h2o_ps = (gc.h2o_v0, gc.h2o_dv, gc.h2o_c2,
                                                                gc.h2o_ts, gc.h2o_ps)
            if gc.h2o_s296 is not None:
                arrs = [_d(gc.h2o_s296), _d(gc.h2o_s260), _d(gc.h2o_f296)]
                self.keep += arrs
                c.nh2o = len(arrs[0])
                c.h2o_s296, c.h2o_s260, c.h2o_f296 = [a.ctypes.data for a in arrs]
            if gc.ch4_wl is not None:
                arrs = [_d(gc.ch4_wl), _d(gc.ch4_k)]
                self.keep += arrs
                c.nch4, (c.ch4_wl, c.ch4_k) = len(arrs[0]), [a.ctypes.data for a in arrs]
            if gc.nh3_wl is not None:
                arrs = [_d(gc.nh3_wl), _d(gc.nh3_k)]
                self.keep += arrs
                c.nnh3, (c.nh3_wl, c.nh3_k) = len(arrs[0]), [a.ctypes.data for a in arrs]
        self.cia.hmin_static = int(getattr(cia, "hmin_static", 0))
        L = req.nlayer
        hfp = req.hfp if req.hfp is not None else np.zeros(L)
        f = [_d(req.press), _d(req.temp), _d(req.delh), _d(hfp), _d(req.pp),
             _d(req.amount), _i(req.idgas), _i(req.isogas), _i(req.nlayin),
             _i(req.imod), _i(req.layinc), _d(req.scale), _d(req.emtemp),
             _i(req.isw), _d(req.vem), _d(req.emissivity)]
        self.keep += f
        nc = req.ncont
        dust = [_d(req.cont), _d(req.vsec), _d(req.xsec)] if nc else []
        self.keep += dust
        self.req = _REQ(L, req.npath, req.ngas, len(req.isw), len(req.vem),
                        req.inormal, req.iray, req.ispace, req.itype, req.flagh2p,
                        req.amform, req.dist, req.tsurf, req.radius1, req.radextra,
                        *[a.ctypes.data for a in f],
                        nc, len(req.vsec) if nc else 0,
                        *([a.ctypes.data for a in dust] if nc else [None, None, None]),
                        int(req.inlte), float(req.nlte_k1), float(req.nlte_a), int(req.iform),
                        int(ilbl))


def spectrum(tables, cia, req, vwave=None, jac_mode=0, nthreads=1, ilbl=0):
    """Run the restatement over ``vwave`` (default: the request's grid).

    ``ilbl=2``: ``tables`` are line-by-line tables (one g-ordinate each, all
    on one wavenumber and P/T grid), looked up by ``get_klblg`` and summed by
    ``noverlapg1`` (``cirsradg_wave.f:336-339,502-506,735-741``).

    Returns ``(output[nwave, npath], gtsurf[nwave, npath],
    doutputdq[nwave, npath, nlayer, nparam])`` where ``doutputdq`` is
    dR/d(amount) for parameters 1..ngas, dR/dT for ngas+1 and dR/d(cont) for
    the ncont dust types after it, indexed by the path's layer position
    (``cirsradg_wave.f:1240-1286``).
    """
    lb = lib()
    pk = _Pack(tables, cia, req, ilbl)
    vw = _d(req.vwave if vwave is None else vwave)
    nw, npath, L, npar = vw.shape[0], req.npath, req.nlayer, req.nparam
    out = np.zeros((nw, npath))
    gts = np.zeros((nw, npath))
    dq = np.zeros((nw, npath, L, npar))
    rc = lb.orc_spectrum(pk.kts, pk.ntab, C.byref(pk.cia), C.byref(pk.req),
                         vw.ctypes.data_as(C.c_void_p), C.c_int(nw), C.c_int(jac_mode),
                         C.c_int(nthreads), out.ctypes.data_as(C.c_void_p),
                         gts.ctypes.data_as(C.c_void_p), dq.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"oracle returned status {rc}")
    return out, gts, dq


def noverlapg(delg, amount, k_gn, dkgndt):
    """``noverlapg.f`` on one layer: returns ``k_g[ng]``, ``dkdq[ngas+1, ng]``."""
    lb = lib()
    k_gn = _d(k_gn)
    dkgndt = _d(dkgndt)
    ngas, ng = k_gn.shape
    kg = np.zeros(ng)
    dk = np.zeros((ngas + 1, ng))
    lb.orc_noverlapg(_d(delg).ctypes.data_as(C.c_void_p), C.c_int(ng), C.c_int(ngas),
                     _d(amount).ctypes.data_as(C.c_void_p), k_gn.ctypes.data_as(C.c_void_p),
                     dkgndt.ctypes.data_as(C.c_void_p), kg.ctypes.data_as(C.c_void_p),
                     dk.ctypes.data_as(C.c_void_p))
    return kg, dk


def noverlapg1(amount, k_gn, dkgndt):
    """``noverlapg1.f`` on one layer (NG = 1): ``k``, ``dkdq[ngas + 1]``."""
    lb = lib()
    k_gn, dkgndt, amount = _d(k_gn), _d(dkgndt), _d(amount)
    ngas = k_gn.shape[0]
    kg = np.zeros(1)
    dk = np.zeros(ngas + 1)
    lb.orc_noverlapg1(C.c_int(ngas), amount.ctypes.data_as(C.c_void_p),
                      k_gn.ctypes.data_as(C.c_void_p), dkgndt.ctypes.data_as(C.c_void_p),
                      kg.ctypes.data_as(C.c_void_p), dk.ctypes.data_as(C.c_void_p))
    return float(kg[0]), dk


def planck(ispace, v, t):
    return lib().orc_planck(int(ispace), float(v), float(t))


def nlte_t(inlte, k1, a, t, p, v):
    """calc_nlte_t.f: the non-LTE source temperature."""
    f = lib().orc_nlte_t
    f.restype = C.c_double
    f.argtypes = [C.c_int] + [C.c_double] * 5
    return f(int(inlte), float(k1), float(a), float(t), float(p), float(v))


def rayleigh(iray, v, p=1.0, t=273.15, fheh2=0.0, fch4h2=0.0, fnh3=0.0):
    """Rayleigh cross-section (cm2) of rayleighj/v/a/ls.f for iray 1..4."""
    return lib().orc_rayleigh(int(iray), float(v), float(p), float(t), float(fheh2),
                              float(fch4h2), float(fnh3))


def planckg(ispace, v, t):
    return lib().orc_planckg(int(ispace), float(v), float(t))


def sort2g(ra):
    """Heapsort of ``ra`` (sort2g.f); returns sorted keys and 1-based order."""
    n = len(ra)
    a = np.zeros(n + 1)
    a[1:] = ra
    ib = np.zeros(n + 1, dtype=np.int32)
    ib[1:] = np.arange(1, n + 1)
    lib().orc_sort2g(C.c_int(n), a.ctypes.data_as(C.c_void_p), ib.ctypes.data_as(C.c_void_p))
    return a[1:], ib[1:]


def hmin_ff(v0, temp):
    """hmin_ff.f with a zero starting sum: H- free-free, cm4 dyne-1."""
    f = lib().orc_hmin_ff
    f.restype = C.c_double
    f.argtypes = [C.c_double, C.c_double]
    return f(float(v0), float(temp))


def hmin_bf(v0):
    """hmin_bf.f with a zero starting sum: H- bound-free cross-section."""
    f = lib().orc_hmin_bf
    f.restype = C.c_double
    f.argtypes = [C.c_double]
    return f(float(v0))


def tabcont(gascon, which, nu):
    """ch4cont.f or nh3cont.f (``which`` 'ch4'/'nh3') on ``gascon``'s
    table, 1e20 cm2/molecule."""
    wl, k = (gascon.ch4_wl, gascon.ch4_k) if which == "ch4" else (gascon.nh3_wl, gascon.nh3_k)
    wl, k = _d(wl), _d(k)
    c = _CIA()
    if which == "ch4":
        c.nch4, c.ch4_wl, c.ch4_k = len(wl), wl.ctypes.data, k.ctypes.data
    else:
        c.nnh3, c.nh3_wl, c.nh3_k = len(wl), wl.ctypes.data, k.ctypes.data
    f = getattr(lib(), f"orc_{which}cont")
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_double]
    return f(C.addressof(c), float(nu))


def cirsconv(fwhm, vwave, y, vconv, vfwhm=None, xfwhm=None, filters=None):
    """cirsconv.f on one spectrum ``y`` at ``vwave`` -> the values at
    ``vconv``; ``vfwhm``/``xfwhm``: the FWHM table (FWHMEXIST); ``filters``
    (fwhm < 0): a list of (vcentral, v[], f[])."""
    vw, yy, vc = _d(vwave), _d(y), _d(vconv)
    out = np.zeros(vc.shape[0])
    nf = 0 if vfwhm is None else len(vfwhm)
    vf = _d(vfwhm if nf else [0.0])
    xf = _d(xfwhm if nf else [0.0])
    filters = filters or []
    fn = np.array([len(f[1]) for f in filters] or [0], dtype=np.int32)
    fc = _d([f[0] for f in filters] or [0.0])
    fv = _d(np.concatenate([f[1] for f in filters]) if filters else [0.0])
    ff = _d(np.concatenate([f[2] for f in filters]) if filters else [0.0])
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    f = lib().orc_cirsconv
    f.restype = C.c_int
    rc = f(C.c_double(fwhm), C.c_int(vw.shape[0]), p(vw), p(yy), C.c_int(vc.shape[0]), p(vc),
           p(out), C.c_int(nf), p(vf), p(xf), C.c_int(len(filters)), p(fn), p(fc), p(fv), p(ff))
    if rc != 0:
        raise RuntimeError(f"oracle cirsconv returned status {rc}")
    return out


def ciaread(cia, vv, temp):
    lb = lib()
    ct, kc = _d(cia.temps), _d(cia.kcia)
    c = _CIA(cia.dnu, ct.ctypes.data, kc.ctypes.data)
    a = np.zeros(9)
    d = np.zeros(9)
    lb.orc_ciaread(C.byref(c), C.c_double(vv), C.c_double(temp),
                   a.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p))
    return a, d
