"""Seeded synthetic inputs for the CIRSRADG_WAVE workloads.

Every array here is built the way NEMESIS hands them to ``cirsradg_wave``
(reference ``radtran/cirsradg/cirsrtfg_wave.f:397-401``):

* layers are stored bottom (deepest, index 0) to top, and a nadir path visits
  them top to bottom, ``layinc(j) = nlayer - j + 1`` (``radtran/path/atmg.f:348-349``);
* pressures are in atm (``cirsradg_wave.f:23-24``), amounts in molecule cm-2;
* k-tables hold ``k * 1e20`` so that ``amount * 1e-20`` keeps the optical depth
  in range (``cirsradg_wave.f:386-391,704``);
* the CIA table has the reference's fixed ``kcia(9,25,1501)`` grid
  (``radtran/includes/ciacom.f:3-9``, ``radtran/radtran/ciaread.f:77-98``).

The generator choices (seeds, column formula, ids) are stated in the
docstrings; nothing here reads a file.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# Emission-angle quadrature used by the reference's nadir example
# (Example_calculations/Jupiter_CIRS_nadir_thermal_emission/cirstest.set:3-7).
MU5 = np.array([0.165278957666387, 0.477924949810444, 0.738773865105505,
                0.919533908166459, 1.0])
WT5 = np.array([0.327539761183898, 0.292042683679684, 0.224889342063117,
                0.133305990851069, 0.0222222222222222])

BAR_PER_ATM = 1.01325
AMU_KG = 1.66053907e-27
KB_SI = 1.380649e-23

# CIA grid fixed by the reference (ciacom.f:3-5).
CIA_NUMWN = 1501
CIA_NUMT = 25
CIA_NPAIRS = 9

# Absorbers with k-tables. Ids 2, 6 and 22 would switch on extra CIA pairs
# (nciacon.f:112-137), so the k-table gases avoid them.
KGAS_IDS = [11, 27, 26, 23, 28, 30, 5, 24]
H2_ID, HE_ID = 39, 40


@dataclass
class KTable:
    """One gas's correlated-k table, the content of one ``.kta`` file.

    Header fields follow ``radtran/cirsrad/read_khead.f:38-109``; ``k`` is
    indexed ``[wavenumber, P, T, g]``, i.e. the record order of
    ``calc_fnktablec.f:593-609`` (g fastest, then T, then P).
    """
    idgas: int
    isogas: int
    vmin: float
    delv: float
    fwhm: float
    press: np.ndarray      # [np] atm (read_khead takes the log)
    temp: np.ndarray       # [nt] K
    g_ord: np.ndarray      # [ng]
    delg: np.ndarray       # [ng]
    k: np.ndarray          # [npoint, np, nt, ng] float64, k_phys * 1e20
    # per-pressure temperature grids [np, nt] (the file's NT < 0; temp then
    # unused) and the wavenumbers of an irregular grid [npoint] (DELV <= 0)
    temp2: Optional[np.ndarray] = None
    vcen: Optional[np.ndarray] = None

    @property
    def nt_signed(self) -> int:
        """NT as the .kta header holds it: negative with per-pressure grids."""
        return -self.nt if self.temp2 is not None else self.nt

    @property
    def npoint(self) -> int:
        return int(self.k.shape[0])

    @property
    def np_(self) -> int:
        return int(self.k.shape[1])

    @property
    def nt(self) -> int:
        return int(self.k.shape[2])

    @property
    def ng(self) -> int:
        return int(self.k.shape[3])


@dataclass
class CIATable:
    """The ``/CIATABLE/`` content: ``kcia(9,25,1501)`` as ``[wn, T, pair]``;
    ``para``, when set, is the ``/FPTABLE/`` used with FLAGH2P = 1."""
    dnu: float
    temps: np.ndarray      # [25]
    kcia: np.ndarray       # [1501, 25, 9] float64
    para: Optional["ParaCIATable"] = None
    e3: Optional["E3Table"] = None
    n2nir: Optional["N2NirTable"] = None
    solar: Optional["SolarSpectrum"] = None
    gascon: Optional["GasCon"] = None
    file: Optional[str] = None  # the raddata table file it came from, if any


@dataclass
class GasCon:
    """``/GASCONDEF/`` (gascom.f) and the tables ngascon.f reads: the CKD
    H2O continuum (``h2o_s296``/``h2o_s260`` self at 296/260 K,
    ``h2o_f296`` foreign, on ``h2o_v0 + i*h2o_dv`` cm-1, with /COMCON/'s
    ``c2``, ``ts``, ``ps``), Karkoschka's CH4 and Lutz & Owen's NH3 tables
    (wavelength micron, coefficient as the data files hold it).  The H-
    continua need no flag."""
    ih2o: int = 0
    ich4: int = 0
    io3: int = 0
    inh3: int = 0
    iuvscat: int = 0
    h2o_v0: float = 0.0
    h2o_dv: float = 10.0
    h2o_c2: float = 1.4387863
    h2o_ts: float = 296.0
    h2o_ps: float = 1.0
    h2o_s296: Optional[np.ndarray] = None
    h2o_s260: Optional[np.ndarray] = None
    h2o_f296: Optional[np.ndarray] = None
    ch4_wl: Optional[np.ndarray] = None
    ch4_k: Optional[np.ndarray] = None
    nh3_wl: Optional[np.ndarray] = None
    nh3_k: Optional[np.ndarray] = None


@dataclass
class SolarSpectrum:
    """``/SOLARDAT/``'s stellar spectrum (get_solar_wave.f): total power
    ``rad`` at ``wave`` in the calculation's units (iread = 999)."""
    wave: np.ndarray
    rad: np.ndarray


def make_solar(v0, v1, n=120, seed=17) -> SolarSpectrum:
    """A smooth positive stellar spectrum on [v0, v1]: a 5800 K Planck shape
    in wavenumber scaled to the Sun's total power, with a 5% ripple."""
    rng = np.random.default_rng(seed)
    v = np.linspace(v0, v1, n)
    b = 1.1911e-12 * v ** 3 / np.expm1(1.439 * v / 5800.0)
    rad = b * 4 * np.pi ** 2 * (6.96e10) ** 2 * (1 + 0.05 * np.sin(v / 37.0 + rng.uniform(0, 6)))
    return SolarSpectrum(wave=v, rad=rad)


@dataclass
class N2NirTable:
    """N2's near-infrared CIA, ``n2n2cia.f`` and ``n2h2cia.f``: each
    function's value (cm-1 amagat-2) at its knots, linearly interpolated
    between them and zero outside."""
    wn_n2n2: np.ndarray
    k_n2n2: np.ndarray
    wn_n2h2: np.ndarray
    k_n2h2: np.ndarray


@dataclass
class E3Table:
    """``/E3TABLE/`` of e3interp.f: the exponential integral E3 against
    transmission on 1001 points (imod 18)."""
    trans: np.ndarray      # [1001]
    e3: np.ndarray         # [1001]


def make_e3_table() -> E3Table:
    """E3(tau) at tau = -ln(tr), tr = i/1000 (the grid e3interp.f assumes),
    from scipy's exponential integral; E3 = 0 at tr = 0 and 1/2 at tr = 1.
    The reference's raddata/e3tab.dat agrees with it to 1e-5 from tr = 0.02
    up and departs below (28% at tr = 0.001); register that file's columns
    for results that match the reference's."""
    from scipy.special import expn
    tr = np.arange(1001) / 1000.0
    e3 = np.zeros(1001)
    e3[1:] = expn(3, -np.log(tr[1:]))
    e3[-1] = 0.5
    return E3Table(trans=tr, e3=e3)


@dataclass
class ParaCIATable:
    """The ``/FPTABLE/`` content that fpread_all.f reads (ciacom.f): H2-H2
    and H2-He coefficients on a grid of para-H2 fractions ``fp`` (the first
    at 0.25, uniform spacing, as nparacon_all.f's indexing assumes) and the
    five other pairs (H2-N2, N2-CH4, N2-N2, CH4-CH4, H2-CH4)."""
    dnu: float
    temps: np.ndarray      # [25]
    fp: np.ndarray         # [npara]
    kh2h2: np.ndarray      # [1501, 25, npara]  (kh2h2(npara,25,1501))
    kh2he: np.ndarray      # [1501, 25, npara]
    kcia5: np.ndarray      # [1501, 25, 5]      (kcia5(5,25,1501))


@dataclass
class Request:
    """All arguments of one ``cirsradg_wave`` spectrum (one call per wavenumber).

    Array axes are C order and 0-based; ``layinc`` keeps the reference's
    1-based layer numbers.
    """
    press: np.ndarray      # [L] atm
    temp: np.ndarray       # [L] K
    pp: np.ndarray         # [L, ngas] atm
    amount: np.ndarray     # [L, ngas] molecule cm-2
    delh: np.ndarray       # [L] km
    idgas: np.ndarray      # [ngas] int
    isogas: np.ndarray     # [ngas] int
    vwave: np.ndarray      # [nwave] cm-1
    nlayin: np.ndarray     # [npath] int
    layinc: np.ndarray     # [npath, L] int, 1-based
    scale: np.ndarray      # [npath, L]
    emtemp: np.ndarray     # [npath, L]
    imod: np.ndarray       # [npath] int
    isw: np.ndarray        # [nsw] int, 1-based parameter ids
    path_weight: np.ndarray  # [npath] caller-side angle weights
    hfp: Optional[np.ndarray] = None
    dist: float = 5.2
    inormal: int = 0
    iray: int = 0
    ispace: int = 0
    amform: int = 0
    itype: int = 0
    tsurf: float = 0.0
    vem: np.ndarray = field(default_factory=lambda: np.array([0.0, 1.0e5]))
    emissivity: np.ndarray = field(default_factory=lambda: np.array([1.0, 1.0]))
    radius1: float = 71492.0
    flagh2p: int = 0
    radextra: float = 0.0
    # dust (cirsradg_wave.f:550-577): cont [L, ncont] particles cm-2; the
    # cross-section grid vsec [nsec] (cm-1) and extinction xsec [ncont, nsec]
    # (cm2 per particle) that the reference reads from /DUST/
    cont: Optional[np.ndarray] = None
    vsec: Optional[np.ndarray] = None
    xsec: Optional[np.ndarray] = None
    # non-LTE source function (calc_nlte_t.f, /NLTE_FLAGS/): 0 LTE, 1 the
    # k1/A collisional form, 2 the pressure form
    inlte: int = 0
    nlte_k1: float = 0.0
    nlte_a: float = 0.0
    # /SOLARDAT/ iform (cirsradg_wave.f:912-925): 1 and 3 turn imod 3
    # radiances into the planet's total flux, 4 pi^2 (radius1 + radextra)^2
    # (km -> cm), times 1e-18 for 3
    iform: int = 0

    @property
    def ncont(self) -> int:
        return 0 if self.cont is None else int(self.cont.shape[1])

    @property
    def nparam(self) -> int:
        # cirsradg_wave.f:375-376
        return self.ngas + 1 + self.ncont + (1 if self.flagh2p == 1 else 0)

    @property
    def nlayer(self) -> int:
        return int(self.press.shape[0])

    @property
    def ngas(self) -> int:
        return int(self.idgas.shape[0])

    @property
    def npath(self) -> int:
        return int(self.imod.shape[0])

    @property
    def nwave(self) -> int:
        return int(self.vwave.shape[0])



def gauss_legendre_g(ng: int):
    """Gauss-Legendre ordinates and weights mapped onto g in [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(ng)
    return 0.5 * (x + 1.0), 0.5 * w


def make_ktables(ids, vmin, delv, npoint, ng=20, np_=20, nt=10,
                 pmin_bar=5e-5, pmax_bar=20.0, tmin=40.0, tmax=220.0,
                 seed=2, mu=-22.0, sigma=2.0, distinct=None) -> List[KTable]:
    """k-tables with log10(k_phys) ~ N(mu, sigma), sorted ascending in g.

    One independent draw per (gas, wavenumber, P, T); the table stores
    ``k_phys * 1e20``.  With ``distinct`` set, only that many wavenumber
    slices are drawn per gas and repeated along the grid (each wavenumber
    still owns its own copy in memory): large tables without the cost of
    drawing every value.
    """
    rng = np.random.default_rng(seed)
    g_ord, delg = gauss_legendre_g(ng)
    press = np.geomspace(pmin_bar, pmax_bar, np_) / BAR_PER_ATM
    temp = np.linspace(tmin, tmax, nt)
    out = []
    for gid in ids:
        nd = npoint if distinct is None else min(int(distinct), npoint)
        lk = rng.normal(mu, sigma, size=(nd, np_, nt, ng))
        lk.sort(axis=-1)
        k = np.power(10.0, lk + 20.0)
        if nd < npoint:
            k = k[np.arange(npoint) % nd]
        out.append(KTable(idgas=int(gid), isogas=0, vmin=float(vmin),
                          delv=float(delv), fwhm=0.0, press=press,
                          temp=temp, g_ord=g_ord, delg=delg, k=k))
    return out


def make_cia(seed=3, dnu=1.0, pairs=(0, 1)) -> CIATable:
    """Smooth positive CIA coefficients (cm-1 amagat-2) on the fixed grid.

    Only the listed pairs are non-zero (0: H2-H2, 1: H2-He for INORMAL=0,
    ``nciacon.f:209-215``).  Each is a low-order random log-polynomial in
    (wavenumber, T), so it is smooth and positive.
    """
    rng = np.random.default_rng(seed)
    temps = np.linspace(40.0, 400.0, CIA_NUMT)
    v = np.arange(CIA_NUMWN) * dnu
    kcia = np.zeros((CIA_NUMWN, CIA_NUMT, CIA_NPAIRS))
    vs = (v - 1000.0) / 1000.0
    ts = (temps - 150.0) / 150.0
    for p in pairs:
        c = rng.normal(0.0, 0.4, size=(3, 3))
        base = -5.5 + rng.normal(0.0, 0.3)
        lg = base + sum(c[a, b] * vs[:, None] ** (a + 1) * ts[None, :] ** b
                        for a in range(3) for b in range(3))
        kcia[:, :, p] = np.power(10.0, lg)
    return CIATable(dnu=float(dnu), temps=temps, kcia=kcia)


def make_para_cia(seed=6, dnu=1.0, npara=24, other_pairs=(0, 3, 4)) -> ParaCIATable:
    """Smooth positive para-H2-dependent CIA: each fraction's H2-H2 and H2-He
    coefficients are a random log-polynomial in (wavenumber, T) plus a term
    linear in the para fraction; ``other_pairs`` of the five extra pairs are
    non-zero."""
    rng = np.random.default_rng(seed)
    temps = np.linspace(40.0, 400.0, CIA_NUMT)
    fp = np.linspace(0.25, 1.0, npara)
    v = np.arange(CIA_NUMWN) * dnu
    vs = (v - 1000.0) / 1000.0
    ts = (temps - 150.0) / 150.0

    def surface(base):
        c = rng.normal(0.0, 0.4, size=(3, 3))
        return base + sum(c[a, b] * vs[:, None] ** (a + 1) * ts[None, :] ** b
                          for a in range(3) for b in range(3))

    kh2h2 = np.empty((CIA_NUMWN, CIA_NUMT, npara))
    kh2he = np.empty((CIA_NUMWN, CIA_NUMT, npara))
    s_h2, s_he = surface(-5.4), surface(-5.8)
    g_h2, g_he = rng.normal(0.0, 0.3, 2)
    for i, f in enumerate(fp):
        kh2h2[:, :, i] = np.power(10.0, s_h2 + g_h2 * (f - 0.25))
        kh2he[:, :, i] = np.power(10.0, s_he + g_he * (f - 0.25))
    kcia5 = np.zeros((CIA_NUMWN, CIA_NUMT, 5))
    for p in other_pairs:
        kcia5[:, :, p] = np.power(10.0, surface(-6.0))
    return ParaCIATable(dnu=float(dnu), temps=temps, fp=fp, kh2h2=kh2h2, kh2he=kh2he, kcia5=kcia5)


def make_lbltables(ids, vmin, delv, npoint, np_=20, nt=10, pmin_bar=5e-5, pmax_bar=20.0,
                   tmin=40.0, tmax=220.0, seed=12, mu=-23.0, sigma=2.5) -> List[KTable]:
    """Line-by-line tables (the ``.lta`` content, one g-ordinate of weight
    1), all on one wavenumber and P/T grid as read_klbllist.f requires:
    log10(k_phys) ~ N(mu, sigma) per (gas, wavenumber, P, T), stored as
    ``k_phys * 1e20``."""
    from .kta import lbl_table

    rng = np.random.default_rng(seed)
    press = np.geomspace(pmin_bar, pmax_bar, np_) / BAR_PER_ATM
    temp = np.linspace(tmin, tmax, nt)
    out = []
    for gid in ids:
        lk = rng.normal(mu, sigma, size=(npoint, np_, nt))
        out.append(lbl_table(int(gid), 0, vmin, delv, press, temp, np.power(10.0, lk + 20.0)))
    return out


def make_atmosphere(nlayer=120, nkgas=3, seed=1, p_bot_bar=10.0,
                    p_top_bar=1e-4, vmr_lo=1e-9, vmr_hi=1e-3, q_h2=0.86,
                    q_he=0.14, mu_bar=2.3, grav=24.79, tnoise=2.0):
    """Layer arrays for a giant-planet nadir atmosphere.

    * level pressures log-spaced from ``p_bot_bar`` to ``p_top_bar``; the
      layer pressure is the geometric mean of its levels, converted to atm;
    * ``T = 110 + 60 tanh(log10(p / 0.1 bar)) + N(0, tnoise)``;
    * column ``N = dp / (mu m_u g)`` (cm-2) and thickness
      ``delh = (k T / (mu m_u g)) ln(p_bot / p_top)`` (km);
    * ``nkgas`` absorbers with log-uniform VMR per layer, then H2 and He.
    """
    rng = np.random.default_rng(seed)
    plev = np.geomspace(p_bot_bar, p_top_bar, nlayer + 1)
    pbar = np.sqrt(plev[:-1] * plev[1:])
    press = pbar / BAR_PER_ATM
    temp = 110.0 + 60.0 * np.tanh(np.log10(pbar / 0.1)) + rng.normal(0.0, tnoise, nlayer)
    dp_pa = (plev[:-1] - plev[1:]) * 1e5
    mass = mu_bar * AMU_KG
    ncol = dp_pa / (mass * grav) * 1e-4
    delh = KB_SI * temp / (mass * grav) * np.log(plev[:-1] / plev[1:]) / 1e3
    vmr = np.exp(rng.uniform(np.log(vmr_lo), np.log(vmr_hi), size=(nlayer, nkgas)))
    vmr = np.concatenate([vmr, np.full((nlayer, 1), q_h2), np.full((nlayer, 1), q_he)], axis=1)
    amount = vmr * ncol[:, None]
    pp = vmr * press[:, None]
    idgas = np.array(KGAS_IDS[:nkgas] + [H2_ID, HE_ID], dtype=np.int32)
    isogas = np.zeros_like(idgas)
    return press, temp, pp, amount, delh, idgas, isogas


def nadir_paths(nlayer, mu=MU5, wt=WT5, temp=None):
    """One ``imod=3`` path per emission angle, layers top to bottom."""
    npath = len(mu)
    layinc = np.tile(np.arange(nlayer, 0, -1, dtype=np.int32), (npath, 1))
    scale = np.repeat((1.0 / np.asarray(mu))[:, None], nlayer, axis=1)
    emtemp = np.zeros((npath, nlayer)) if temp is None else temp[layinc - 1]
    nlayin = np.full(npath, nlayer, dtype=np.int32)
    imod = np.full(npath, 3, dtype=np.int32)
    return nlayin, layinc, scale, emtemp, imod, np.asarray(wt, dtype=np.float64)


def make_request(nlayer=120, nkgas=3, nwave=1000, v0=600.0, v1=1400.0,
                 seed=1, wave_offset=0, table_vmin=None, table_delv=None,
                 mu=MU5, wt=WT5, isw=None) -> Request:
    """One spectrum request; the wavenumbers are grid points of the table."""
    press, temp, pp, amount, delh, idgas, isogas = make_atmosphere(
        nlayer=nlayer, nkgas=nkgas, seed=seed)
    delv = (v1 - v0) / (nwave - 1) if table_delv is None else table_delv
    vmin = v0 if table_vmin is None else table_vmin
    vwave = vmin + delv * (wave_offset + np.arange(nwave))
    nlayin, layinc, scale, emtemp, imod, pw = nadir_paths(nlayer, mu, wt, temp)
    ngas = idgas.shape[0]
    if isw is None:
        isw = np.arange(1, ngas + 2, dtype=np.int32)
    return Request(press=press, temp=temp, pp=pp, amount=amount, delh=delh,
                   idgas=idgas, isogas=isogas, vwave=vwave, nlayin=nlayin,
                   layinc=layinc, scale=scale, emtemp=emtemp, imod=imod,
                   isw=np.asarray(isw, dtype=np.int32), path_weight=pw,
                   hfp=np.zeros(nlayer))


def config5_tables(nwave=10000, nkgas=8, ng=20, np_=30, nt=15, distinct=64):
    """Tables of the large-spectrum stress: 8 k-gases on 10000 wavenumbers at
    0.1 cm-1 (600-1599.9 cm-1), a 30 P x 15 T grid and 20 g-ordinates
    (8 x 10000 x 9000 values, 5.76 GB in FP64, 1.44 GB as BF16)."""
    return make_ktables(KGAS_IDS[:nkgas], 600.0, 0.1, nwave, ng=ng, np_=np_, nt=nt,
                        distinct=distinct)


def config5_request(seed, nwave=10000, nlayer=200, nkgas=8):
    """One request of the large-spectrum stress: 200 layers, the 8 k-gases +
    H2/He, every wavenumber of the 0.1 cm-1 table grid."""
    return make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed,
                        table_vmin=600.0, table_delv=0.1)


def config1(nwave=1000, nlayer=120, nkgas=3, ng=20, np_=20, nt=10):
    """The 120 layer x 1000 wavenumber x 20 g workload (3 k-gases + H2, He).

    Seeds: atmosphere 1, k-tables 2, CIA 3.  The table grid equals the
    calculation grid, 600-1400 cm-1, so every lookup is coincident.
    """
    v0, v1 = 600.0, 1400.0
    delv = (v1 - v0) / (nwave - 1) if nwave > 1 else 1.0
    tables = make_ktables(KGAS_IDS[:nkgas], v0, delv, nwave, ng=ng, np_=np_, nt=nt)
    cia = make_cia()
    req = make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, v0=v0, v1=v1)
    return tables, cia, req
