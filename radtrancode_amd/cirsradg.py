"""Python mirror of the CIRSRADG_WAVE interface on the gfx950 library.

Reference correspondence:

* :func:`cirsradg_wave` -- ``CALL cirsradg_wave(...)``
  (``radtran/cirsradg/cirsradg_wave.f:1-5``): the same 37 arguments, as
  Fortran-ordered arrays with the fixed leading dimensions of
  ``radtran/includes/arrdef.f``; outputs are written in place.
* :func:`spectrum` -- the wavenumber loop of ``CIRSRTFG_WAVE``
  (``radtran/cirsradg/cirsrtfg_wave.f:361-448``) for one request.
* :func:`register_ktables` / :func:`read_klist` -- the ``/INTERPK/`` registry
  that ``read_klist.f`` fills; :func:`register_lbltables` /
  :func:`read_klbllist` -- ``/INTERPKLBL/`` (``read_klbllist.f``) and
  :func:`set_ilbl` -- ``/LBLTABLE/`` ILBL; :func:`register_cia` -- ``/CIATABLE/``;
  :func:`register_dust` -- ``/DUST/`` for :func:`cirsradg_wave` (batched
  requests carry their own dust arrays); :func:`register_parah2` --
  ``/FPTABLE/``; :func:`register_e3` -- ``/E3TABLE/`` for imod 18;
  :func:`register_n2nir` -- N2's near-infrared CIA tables;
  :func:`register_gascon` -- ``/GASCONDEF/`` with ngascon.f's tables.
* :class:`Batch` -- many requests resident on the device, launched together.
* :func:`cirsconv` -- ``cirsconv.f``, the convolution of the caller's tail
  (``cirsrtfg_wave.f:468-516``), for many spectra at once on the device.

Errors raise :class:`RadgError` where the reference would STOP.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import MODE_BF16, MODE_F16, MODE_FP64, MODE_MIXED, RadgError, Request, check, lib  # noqa: F401

MAXLAY, MAXPAT, MAXGAS, MAXCON, MAXSEC = 400, 200, 30, 10, 16000
NPARMAX = MAXGAS + 2 + MAXCON


def _vp(a):
    return C.c_void_p(a.ctypes.data)


def set_mode(mode: int) -> None:
    check(lib().radg_set_mode(C.c_int(mode)))


def get_mode() -> int:
    return lib().radg_get_mode()


def register_ktables(tables) -> None:
    """Register ``synth.KTable`` objects sharing one P/T/g grid (with
    ``temp2`` set, one temperature grid per pressure; with ``delv <= 0``, the
    irregular wavenumber grid ``vcen``)."""
    t0 = tables[0]
    n = len(tables)
    ids = np.array([t.idgas for t in tables], dtype=np.int32)
    isos = np.array([t.isogas for t in tables], dtype=np.int32)
    npts = np.array([t.npoint for t in tables], dtype=np.int32)
    vmin = np.array([t.vmin for t in tables], dtype=np.float64)
    delv = np.array([t.delv for t in tables], dtype=np.float64)
    ks = [np.ascontiguousarray(t.k, dtype=np.float64) for t in tables]
    for t in tables[1:]:
        if (t.np_, t.nt, t.ng) != (t0.np_, t0.nt, t0.ng):
            raise ValueError("all tables must share the P/T/g grid")
    ptrs = (C.c_void_p * n)(*[k.ctypes.data for k in ks])
    vcs = [None if getattr(t, "vcen", None) is None else
           np.ascontiguousarray(t.vcen, dtype=np.float64) for t in tables]
    vptrs = (C.c_void_p * n)(*[None if v is None else v.ctypes.data for v in vcs])
    press = np.ascontiguousarray(t0.press, dtype=np.float64)
    t2 = getattr(t0, "temp2", None)
    temp = np.ascontiguousarray(t0.temp if t2 is None else t2, dtype=np.float64)
    gord = np.ascontiguousarray(t0.g_ord, dtype=np.float64)
    delg = np.ascontiguousarray(t0.delg, dtype=np.float64)
    check(lib().radg_register_ktables(
        C.c_int(n), _vp(ids), _vp(isos), _vp(npts), _vp(vmin), _vp(delv),
        C.c_int(t0.np_), C.c_int(t0.nt_signed), C.c_int(t0.ng), _vp(press), _vp(temp),
        _vp(gord), _vp(delg), ptrs, vptrs))


def ktable_codes(mode: int, itab: int, shape):
    """16-bit codes (``MODE_BF16`` or ``MODE_F16``) of registered table
    ``itab`` as made on the device, and its offsets [npoint, ng]:
    log10(k) = code + offset[point, g]."""
    codes = np.empty(shape, dtype=np.uint16)
    offs = np.zeros((shape[0], shape[-1]), dtype=np.float32)
    check(lib().radg_ktable_codes(C.c_int(mode), C.c_int(itab), _vp(codes), _vp(offs)))
    return codes, offs


def read_klist(path: str, recl: int = 4) -> None:
    """Register the ``.kta`` files named in a ``.kls`` list."""
    check(lib().radg_read_klist(path.encode(), int(recl)))


def register_lbltables(tables) -> None:
    """Register line-by-line tables (``/INTERPKLBL/``, ILBL = 2): one-ordinate
    ``KTable`` objects (``kta.lbl_table``) on one wavenumber grid; the P/T
    grid is the last table's, as ``read_klbllist.f`` leaves it."""
    t0, tl = tables[0], tables[-1]
    n = len(tables)
    for t in tables:
        if t.ng != 1:
            raise ValueError("line-by-line tables have one g-ordinate")
        if (t.np_, t.nt_signed, t.npoint, t.vmin, t.delv) != (t0.np_, t0.nt_signed, t0.npoint,
                                                               t0.vmin, t0.delv):
            raise ValueError("line-by-line tables must share their wavenumber and P/T grid sizes")
    ids = np.array([t.idgas for t in tables], dtype=np.int32)
    isos = np.array([t.isogas for t in tables], dtype=np.int32)
    ks = [np.ascontiguousarray(t.k, dtype=np.float64) for t in tables]
    ptrs = (C.c_void_p * n)(*[k.ctypes.data for k in ks])
    press = np.ascontiguousarray(tl.press, dtype=np.float64)
    t2 = getattr(tl, "temp2", None)
    temp = np.ascontiguousarray(tl.temp if t2 is None else t2, dtype=np.float64)
    check(lib().radg_register_lbltables(
        C.c_int(n), _vp(ids), _vp(isos), C.c_int(t0.npoint), C.c_double(t0.vmin),
        C.c_double(t0.delv), C.c_int(t0.np_), C.c_int(t0.nt_signed), _vp(press), _vp(temp), ptrs))


def read_klbllist(path: str, recl: int = 4) -> None:
    """Register the ``.lta`` files named in an ``.lls`` list."""
    check(lib().radg_read_klbllist(path.encode(), int(recl)))


def set_ilbl(ilbl: int) -> None:
    """``/LBLTABLE/`` ILBL: 0 k-tables, 2 line-by-line tables."""
    check(lib().radg_set_ilbl(C.c_int(ilbl)))


def get_ilbl() -> int:
    return lib().radg_get_ilbl()


def register_cia(cia) -> None:
    """``/CIATABLE/`` and, if ``cia.para`` is set, ``/FPTABLE/`` (else the
    para-H2 table is cleared); likewise ``/E3TABLE/`` from ``cia.e3`` and
    N2's near-infrared CIA from ``cia.n2nir``."""
    temps = np.ascontiguousarray(cia.temps, dtype=np.float64)
    kcia = np.ascontiguousarray(cia.kcia, dtype=np.float64)
    if kcia.shape != (1501, 25, 9):
        raise ValueError("kcia must be [1501][25][9]")
    check(lib().radg_register_cia(C.c_double(cia.dnu), _vp(temps), _vp(kcia)))
    register_parah2(getattr(cia, "para", None))
    register_e3(getattr(cia, "e3", None))
    register_n2nir(getattr(cia, "n2nir", None))
    register_solar(getattr(cia, "solar", None))
    register_gascon(getattr(cia, "gascon", None))


class _GasCon(C.Structure):
    _fields_ = [("ih2o", C.c_int), ("ich4", C.c_int), ("io3", C.c_int), ("inh3", C.c_int),
                ("iuvscat", C.c_int), ("nh2o", C.c_int), ("h2o_v0", C.c_double),
                ("h2o_dv", C.c_double), ("h2o_c2", C.c_double), ("h2o_ts", C.c_double),
                ("h2o_ps", C.c_double), ("h2o_s296", C.c_void_p), ("h2o_s260", C.c_void_p),
                ("h2o_f296", C.c_void_p), ("nch4", C.c_int), ("nnh3", C.c_int),
                ("ch4_wl", C.c_void_p), ("ch4_k", C.c_void_p), ("nh3_wl", C.c_void_p),
                ("nh3_k", C.c_void_p)]


def register_gascon(g) -> None:
    """``/GASCONDEF/`` and ngascon.f's tables from a ``synth.GasCon`` (None
    clears it: every flag 0, the H- continua still active)."""
    if g is None:
        check(lib().radg_register_gascon(None))
        return
    keep = []

    def arr(x):
        if x is None:
            return None, 0
        a = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data, a.shape[0]

    s96, n96 = arr(g.h2o_s296)
    s60, n60 = arr(g.h2o_s260)
    f96, nf = arr(g.h2o_f296)
    if not (n96 == n60 == nf):
        raise ValueError("CKD H2O tables must have one length")
    cw, nc = arr(g.ch4_wl)
    ck, nck = arr(g.ch4_k)
    nw, nn = arr(g.nh3_wl)
    nk, nnk = arr(g.nh3_k)
    if nc != nck or nn != nnk:
        raise ValueError("CH4/NH3 tables: wavelengths and coefficients must match")
    st = _GasCon(int(g.ih2o), int(g.ich4), int(g.io3), int(g.inh3), int(g.iuvscat), n96,
                 float(g.h2o_v0), float(g.h2o_dv), float(g.h2o_c2), float(g.h2o_ts),
                 float(g.h2o_ps), s96, s60, f96, nc, nn, cw, ck, nw, nk)
    check(lib().radg_register_gascon(C.byref(st)))


class _Conv(C.Structure):
    _fields_ = [("fwhm", C.c_double), ("nfwhm", C.c_int), ("vfwhm", C.c_void_p),
                ("xfwhm", C.c_void_p), ("nfil", C.c_int), ("fil_n", C.c_void_p),
                ("fil_vcen", C.c_void_p), ("fil_v", C.c_void_p), ("fil_f", C.c_void_p)]


def cirsconv(vwave, y, vconv, fwhm, vfwhm=None, xfwhm=None, filters=None, stream=None):
    """``cirsconv.f`` (``radtran/cirsrad/cirsconv.f:12-246``) on the device for
    many spectra at once: ``y`` [nwave] or [nwave, ncol] at ``vwave`` ->
    [nconv] or [nconv, ncol] at ``vconv``.  ``y`` may be a numpy array (the
    result is one too) or a CUDA ``torch`` tensor of float64 (the result
    stays on the device).  ``vfwhm``/``xfwhm``: the FWHM table
    (FWHMEXIST); ``filters`` (fwhm < 0): a list of (vcentral, v[], f[]), the
    ``runname.fil`` content."""
    import torch
    vw = np.ascontiguousarray(vwave, dtype=np.float64)
    vc = np.ascontiguousarray(vconv, dtype=np.float64)
    on_dev = isinstance(y, torch.Tensor)
    yd = y if on_dev else torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).cuda()
    if yd.dtype != torch.float64 or not yd.is_cuda or not yd.is_contiguous():
        raise ValueError("y must be a contiguous float64 CUDA tensor or a numpy array")
    one = yd.dim() == 1
    ncol = 1 if one else int(yd.shape[1])
    if yd.shape[0] != vw.shape[0]:
        raise ValueError("y's first axis must match vwave")
    out = torch.empty((vc.shape[0],) if one else (vc.shape[0], ncol), dtype=torch.float64,
                      device=yd.device)
    keep = []

    def arr(x, dt=np.float64):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    cv = _Conv(float(fwhm), 0, None, None, 0, None, None, None, None)
    if vfwhm is not None:
        cv.nfwhm, cv.vfwhm, cv.xfwhm = len(vfwhm), arr(vfwhm), arr(xfwhm)
    if filters:
        cv.nfil = len(filters)
        cv.fil_n = arr([len(f[1]) for f in filters], np.int32)
        cv.fil_vcen = arr([f[0] for f in filters])
        cv.fil_v = arr(np.concatenate([f[1] for f in filters]))
        cv.fil_f = arr(np.concatenate([f[2] for f in filters]))
    st = torch.cuda.current_stream().cuda_stream if stream is None else (
        stream if isinstance(stream, int) else stream.cuda_stream)
    check(lib().radg_cirsconv(int(vw.shape[0]), _vp(vw), int(vc.shape[0]), _vp(vc), C.byref(cv),
                              int(ncol), C.c_void_p(yd.data_ptr()), C.c_void_p(out.data_ptr()),
                              C.c_void_p(st)))
    return out if on_dev else out.cpu().numpy()


def read_cia(path: str, dnu: float) -> None:
    """``/CIATABLE/`` from a CIA table file (ciaread.f's layout), as the
    run's ``.cia`` file names it with its ``dnu``; leaves the other tables
    registered with it as they are."""
    check(lib().radg_read_cia(path.encode(), C.c_double(dnu)))


def register_solar(s) -> None:
    """``/SOLARDAT/``'s stellar spectrum (iread = 999) from a
    ``synth.SolarSpectrum`` (None clears it)."""
    if s is None:
        check(lib().radg_register_solar(0, None, None))
        return
    w = np.ascontiguousarray(s.wave, dtype=np.float64)
    r = np.ascontiguousarray(s.rad, dtype=np.float64)
    if w.shape != r.shape or w.ndim != 1:
        raise ValueError("stellar spectrum: wave and rad must be 1-D of one length")
    check(lib().radg_register_solar(int(w.shape[0]), _vp(w), _vp(r)))


def register_n2nir(t) -> None:
    """N2's near-infrared CIA from a ``synth.N2NirTable`` (None clears it):
    ``n2n2cia.f`` and ``n2h2cia.f`` values at their knots."""
    if t is None:
        check(lib().radg_register_n2nir(0, None, None, 0, None, None))
        return
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (t.wn_n2n2, t.k_n2n2, t.wn_n2h2, t.k_n2h2)]
    if a[0].shape != a[1].shape or a[2].shape != a[3].shape:
        raise ValueError("N2 near-infrared CIA: knots and values must match")
    check(lib().radg_register_n2nir(int(a[0].shape[0]), _vp(a[0]), _vp(a[1]),
                                     int(a[2].shape[0]), _vp(a[2]), _vp(a[3])))


def register_e3(e3) -> None:
    """``/E3TABLE/`` (e3interp.f) from a ``synth.E3Table`` of 1001 points
    (None clears it); needed by imod 18 paths."""
    if e3 is None:
        check(lib().radg_register_e3(0, None, None))
        return
    t = np.ascontiguousarray(e3.trans, dtype=np.float64)
    v = np.ascontiguousarray(e3.e3, dtype=np.float64)
    if t.shape != v.shape or t.ndim != 1:
        raise ValueError("E3 table: trans and e3 must be 1-D of one length")
    check(lib().radg_register_e3(int(t.shape[0]), _vp(t), _vp(v)))


def register_parah2(para) -> None:
    """``/FPTABLE/`` from a ``synth.ParaCIATable`` (None clears it)."""
    if para is None:
        check(lib().radg_register_parah2(C.c_double(0.0), None, 0, None, None, None, None))
        return
    a = [np.ascontiguousarray(x, dtype=np.float64)
         for x in (para.temps, para.fp, para.kh2h2, para.kh2he, para.kcia5)]
    npara = a[1].shape[0]
    if a[2].shape != (1501, 25, npara) or a[3].shape != a[2].shape or a[4].shape != (1501, 25, 5):
        raise ValueError("para-H2 tables must be [1501][25][npara] and [1501][25][5]")
    check(lib().radg_register_parah2(C.c_double(para.dnu), _vp(a[0]), int(npara), _vp(a[1]),
                                      _vp(a[2]), _vp(a[3]), _vp(a[4])))


def register_dust(vsec=None, xsec=None) -> None:
    """``/DUST/``: extinction cross-sections ``xsec`` [ncont, nsec] (cm2) on
    ``vsec`` [nsec] for the fixed-argument entry point; None clears it."""
    if vsec is None:
        check(lib().radg_register_dust(0, 0, None, None))
        return
    vs = np.ascontiguousarray(vsec, dtype=np.float64)
    xs = np.ascontiguousarray(xsec, dtype=np.float64)
    if xs.ndim != 2 or xs.shape[1] != vs.shape[0]:
        raise ValueError("xsec must be [ncont, nsec]")
    check(lib().radg_register_dust(int(vs.shape[0]), int(xs.shape[0]), _vp(vs), _vp(xs)))


def set_iform(iform=0, iread=0) -> None:
    """``/SOLARDAT/`` iform and iread for :func:`cirsradg_wave`; iread = 999
    must match a registered stellar spectrum (:func:`register_solar`)."""
    check(lib().radg_set_iform(int(iform), int(iread)))


def set_nlte(flag=0, k1=0.0, a=0.0) -> None:
    """``/NLTE_FLAGS/`` of calc_nlte_t.f for :func:`cirsradg_wave`."""
    check(lib().radg_set_nlte(int(flag), float(k1), float(a)))


class _ReqArrays:
    """Contiguous copies of one request, kept alive while the batch exists."""

    def __init__(self, req):
        d = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.press, self.temp, self.delh = d(req.press), d(req.temp), d(req.delh)
        self.pp, self.amount = d(req.pp), d(req.amount)
        self.idgas, self.isogas = i(req.idgas), i(req.isogas)
        self.nlayin, self.imod, self.layinc = i(req.nlayin), i(req.imod), i(req.layinc)
        self.scale, self.emtemp = d(req.scale), d(req.emtemp)
        self.isw, self.vwave = i(req.isw), d(req.vwave)
        self.vem, self.emis = d(req.vem), d(req.emissivity)
        self.pw = d(req.path_weight)
        nc = req.ncont
        dust = (d(req.cont), d(req.vsec), d(req.xsec)) if nc else ()
        self.dust = dust
        self.hfp = d(req.hfp) if req.flagh2p == 1 else None
        self.s = Request(
            req.nlayer, req.npath, req.ngas, len(self.isw), len(self.vwave), len(self.vem),
            req.inormal, req.iray, req.ispace, req.itype, req.flagh2p, float(req.tsurf),
            *[a.ctypes.data for a in (self.press, self.temp, self.delh, self.pp, self.amount,
                                      self.idgas, self.isogas, self.nlayin, self.imod,
                                      self.layinc, self.scale, self.emtemp, self.isw,
                                      self.vwave, self.vem, self.emis, self.pw)],
            nc, len(dust[1]) if nc else 0, *([a.ctypes.data for a in dust] if nc else [None] * 3),
            int(req.inlte), float(req.nlte_k1), float(req.nlte_a),
            int(req.iform), float(req.radius1), float(req.radextra),
            None if self.hfp is None else self.hfp.ctypes.data, float(req.dist))


class StateMap(C.Structure):
    """``radg_statemap`` (include/radg.h)."""
    _fields_ = [("nv", C.c_int), ("npro", C.c_int), ("dam", C.c_void_p), ("dte", C.c_void_p),
                ("xmap", C.c_void_p), ("dco", C.c_void_p), ("dfp", C.c_void_p)]


def _requests(reqs):
    keep = [_ReqArrays(r) for r in reqs]
    arr = (Request * len(keep))(*[k.s for k in keep])
    return keep, arr


def batch_bytes(reqs, mode=MODE_MIXED) -> int:
    """Device bytes a batch of these requests would allocate."""
    keep, arr = _requests(reqs)
    out = C.c_size_t(0)
    check(lib().radg_batch_bytes(len(reqs), C.cast(arr, C.c_void_p), int(mode), C.byref(out)))
    return int(out.value)


class Batch:
    """A set of requests packed into device memory and launched together."""

    def __init__(self, reqs, mode=MODE_MIXED, budget_bytes=0):
        self.reqs = list(reqs)
        self.mode = mode
        self._keep, arr = _requests(self.reqs)
        h = C.c_void_p()
        check(lib().radg_batch_create(len(self.reqs), C.cast(arr, C.c_void_p), int(mode),
                                      C.c_size_t(budget_bytes), C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        """Launch on a HIP stream: ``None``, an int handle or a torch stream."""
        sh = None
        if stream is not None:
            sh = stream if isinstance(stream, int) else stream.cuda_stream
        check(lib().radg_batch_launch(self.h, C.c_void_p(sh)))

    def sync(self):
        check(lib().radg_batch_sync(self.h))

    def read(self, r: int, weighted=True):
        q = self.reqs[r]
        nw, npath, L, npar = q.nwave, q.npath, q.nlayer, q.nparam
        out = np.zeros((nw, npath))
        gts = np.zeros((nw, npath))
        jac = np.zeros((nw, npath, L, npar))
        res = dict(output=out, gtsurf=gts, doutputdq=jac)
        wo = wg = wj = None
        if weighted:
            wo, wg, wj = np.zeros(nw), np.zeros(nw), np.zeros((nw, L, npar))
            res.update(wout=wo, wgts=wg, wjac=wj)
        p = lambda a: None if a is None else _vp(a)
        check(lib().radg_batch_read(self.h, int(r), p(out), p(gts), p(jac), p(wo), p(wg), p(wj)))
        return res

    def set_statemaps(self, maps, per_path=False):
        """Attach layer -> profile -> state maps (map2pro.f + map2xvec.f) to
        the requests: ``maps[r]`` is None or a dict with ``dam``, ``dte``
        [nlayer, npro], ``xmap`` [nv, nparam, npro] and, for a request with
        dust, ``dco``, with FLAGH2P = 1, ``dfp`` [nlayer, npro] (FP64)."""
        n = len(self.reqs)
        arr = (StateMap * n)()
        keep = []
        for r, m in enumerate(maps):
            if m is None:
                continue
            dam, dte, xm = (np.ascontiguousarray(m[k], dtype=np.float64)
                            for k in ("dam", "dte", "xmap"))
            q = self.reqs[r]
            if dam.shape != dte.shape or dam.shape[0] != q.nlayer or \
                    xm.shape[1:] != (q.nparam, dam.shape[1]):
                raise ValueError(f"request {r}: statemap shapes do not match the request")
            extra = []
            for key in ("dco", "dfp"):
                e = None
                if m.get(key) is not None:
                    e = np.ascontiguousarray(m[key], dtype=np.float64)
                    if e.shape != dam.shape:
                        raise ValueError(f"request {r}: {key} must be [nlayer, npro]")
                extra.append(e)
            keep += [dam, dte, xm] + extra
            arr[r] = StateMap(xm.shape[0], dam.shape[1], dam.ctypes.data, dte.ctypes.data,
                              xm.ctypes.data, *[None if e is None else e.ctypes.data for e in extra])
        self._maps = [None if m is None else m["xmap"].shape[0] for m in maps]
        check(lib().radg_batch_set_statemaps(self.h, C.cast(arr, C.c_void_p), int(per_path)))

    def read_mapped(self, r: int, per_path=False):
        """Request r's mapped Jacobian: ``kk`` [nwave, nv] (path-weighted) and,
        if asked for, ``doutdx`` [nwave, npath, nv]."""
        q, nv = self.reqs[r], self._maps[r]
        kk = np.zeros((q.nwave, nv))
        dx = np.zeros((q.nwave, q.npath, nv)) if per_path else None
        check(lib().radg_batch_read_mapped(self.h, int(r), None if dx is None else _vp(dx), _vp(kk)))
        return dict(kk=kk, doutdx=dx)

    def device_outputs(self, r: int):
        """Device addresses (ints) of request r's wout[nwave] and kk[nwave, nv]."""
        w, k = C.c_void_p(), C.c_void_p()
        check(lib().radg_batch_device_outputs(self.h, int(r), C.byref(w), C.byref(k)))
        return w.value, k.value

    def timings(self):
        ms = np.zeros(3, dtype=np.float32)
        check(lib().radg_batch_timings(self.h, _vp(ms)))
        return ms

    def timing_stats(self):
        """(launches, summed ms per kernel) over the last <=128 launches."""
        n = C.c_int()
        ms = np.zeros(3, dtype=np.float32)
        check(lib().radg_batch_timing_stats(self.h, C.byref(n), _vp(ms)))
        return n.value, ms

    def info(self):
        ab = np.zeros(3)
        ns, na = C.c_int(), C.c_int()
        check(lib().radg_batch_info(self.h, _vp(ab), C.byref(ns), C.byref(na)))
        return dict(table_bytes=ab[0], output_bytes=ab[1], input_bytes=ab[2],
                    alg_bytes=float(ab.sum()), nspectra=ns.value, items_a=na.value)

    def close(self):
        if getattr(self, "h", None):
            lib().radg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spectrum(req, mode=MODE_FP64):
    """Run one request; returns ``(output, gtsurf, doutputdq)`` as the oracle."""
    b = Batch([req], mode=mode)
    try:
        b.launch()
        b.sync()
        r = b.read(0, weighted=False)
    finally:
        b.close()
    return r["output"], r["gtsurf"], r["doutputdq"]


def fortran_arrays(req, real=np.float32):
    """The request laid out as the Fortran caller holds it (arrdef.f limits)."""
    L, ng, npath = req.nlayer, req.ngas, req.npath
    F = dict(order="F")
    a = dict(
        delh=np.asarray(req.delh, dtype=real),
        press=np.zeros(MAXLAY, real), temp=np.zeros(MAXLAY, real), hfp=np.zeros(MAXLAY, real),
        pp=np.zeros((MAXLAY, MAXGAS), real, **F), amount=np.zeros((MAXLAY, MAXGAS), real, **F),
        nlayin=np.zeros(MAXPAT, np.int32), layinc=np.zeros((MAXLAY, MAXPAT), np.int32, **F),
        cont=np.zeros((MAXCON, MAXLAY), real, **F), scale=np.zeros((MAXLAY, MAXPAT), real, **F),
        emtemp=np.zeros((MAXLAY, MAXPAT), real, **F), imod=np.zeros(MAXPAT, np.int32),
        idgas=np.zeros(MAXGAS, np.int32), isogas=np.zeros(MAXGAS, np.int32),
        vem=np.zeros(MAXSEC, real), emissivity=np.zeros(MAXSEC, real),
        isw=np.zeros(NPARMAX, np.int32))
    a["press"][:L] = req.press
    a["temp"][:L] = req.temp
    a["pp"][:L, :ng] = req.pp
    a["amount"][:L, :ng] = req.amount
    a["nlayin"][:npath] = req.nlayin
    a["layinc"][:L, :npath] = req.layinc.T
    a["scale"][:L, :npath] = req.scale.T
    a["emtemp"][:L, :npath] = req.emtemp.T
    a["imod"][:npath] = req.imod
    a["idgas"][:ng] = req.idgas
    a["isogas"][:ng] = req.isogas
    a["vem"][:len(req.vem)] = req.vem
    a["emissivity"][:len(req.emissivity)] = req.emissivity
    a["isw"][:len(req.isw)] = req.isw
    if req.hfp is not None:
        a["hfp"][:L] = req.hfp
    if req.ncont:
        a["cont"][:req.ncont, :L] = np.asarray(req.cont).T
    return a


def cirsradg_wave(dist, inormal, iray, delh, nlayer, npath, ngas, press, temp, pp, amount,
                  iwave, ispace, amform, vwave, nlayin, layinc, cont, scale, imod, idgas,
                  isogas, emtemp, itype, nem, vem, emissivity, tsurf, gtsurf, radius1,
                  flagh2p, hfp, radextra, nsw, isw, output, doutputdq):
    """``CALL cirsradg_wave(...)`` with REAL(4) arrays (arrdef.f leading dims).

    ``gtsurf``, ``output`` (``[MAXPAT]`` float32) and ``doutputdq``
    (``[MAXPAT, MAXLAY, NPARMAX]`` float32, Fortran order) are written in place.
    """
    f4 = lambda x: C.byref(C.c_float(float(x)))
    i4 = lambda x: C.byref(C.c_int(int(x)))
    arr = lambda a, dt: _vp(_checked(a, dt))
    lib().cirsradg_wave_(
        f4(dist), i4(inormal), i4(iray), arr(delh, np.float32), i4(nlayer), i4(npath), i4(ngas),
        arr(press, np.float32), arr(temp, np.float32), arr(pp, np.float32),
        arr(amount, np.float32), i4(iwave), i4(ispace), i4(amform), f4(vwave),
        arr(nlayin, np.int32), arr(layinc, np.int32), arr(cont, np.float32),
        arr(scale, np.float32), arr(imod, np.int32), arr(idgas, np.int32),
        arr(isogas, np.int32), arr(emtemp, np.float32), i4(itype), i4(nem),
        arr(vem, np.float32), arr(emissivity, np.float32), f4(tsurf), arr(gtsurf, np.float32),
        f4(radius1), i4(flagh2p), arr(hfp, np.float32), f4(radextra), i4(nsw),
        arr(isw, np.int32), arr(output, np.float32), arr(doutputdq, np.float32))
    check(lib().radg_cirsradg_status())


def _checked(a, dt):
    if not isinstance(a, np.ndarray) or a.dtype != dt:
        raise TypeError(f"expected a numpy array of {np.dtype(dt).name}")
    if not (a.flags.f_contiguous or a.ndim == 1):
        raise TypeError("arrays must be Fortran-ordered")
    return a
