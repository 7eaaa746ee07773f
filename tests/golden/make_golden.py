"""Generate the golden cases from the reference's own CIRSRADG_WAVE.

Run in a checkout where ``/root/reference`` is present::

    make -C oracle ref
    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py rayleigh_j # only the named ones

Each case stores its inputs (request arrays and k-tables) and the outputs of
the reference compiled with REAL promoted to 8 bytes (``ref64_*``: the FP64
oracle) and as written (``ref32_*``).  The inputs come from
``radtrancode_amd.synth`` with fixed seeds; edits for the edge cases are made
here and saved with the case, so the fixtures do not depend on the
generator staying unchanged.
"""
from __future__ import annotations

import copy
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from radtrancode_amd import synth  # noqa: E402
from oracle import ref  # noqa: E402
from tests.golden.golden_io import load_e3, load_gascon, load_n2nir, save_case, save_cia  # noqa: E402

REF_E3TAB = "/root/reference/raddata/e3tab.dat"


# the knots of n2n2cia.f (66 points from 4500 cm-1) and n2h2cia.f (201 from
# 3995 cm-1), 5 cm-1 apart
N2N2_KNOTS = 4500.0 + 5.0 * np.arange(66)
N2H2_KNOTS = 3995.0 + 5.0 * np.arange(201)


def save_n2nir(path=os.path.join(HERE, "n2nir.npz")):
    """N2's near-infrared CIA from the reference's own N2N2CIA and N2H2CIA
    (REAL(8) build): the values at the knots (the table a caller registers)
    and, to pin the knots and the interpolation, at every mid-point and at
    points just outside each range."""
    d = {}
    for which, knots in (("n2n2", N2N2_KNOTS), ("n2h2", N2H2_KNOTS)):
        probe = np.concatenate([0.5 * (knots[1:] + knots[:-1]), knots[:1] - 0.5, knots[-1:] + 0.5,
                                knots[:1] + 1.25, knots[-1:] - 3.75])
        d[f"{which}_wn"] = knots
        d[f"{which}_k"] = np.array([ref.n2nir(64, which, v) for v in knots])
        d[f"{which}_probe_wn"] = probe
        d[f"{which}_probe_k"] = np.array([ref.n2nir(64, which, v) for v in probe])
    np.savez_compressed(path, **d)


def save_e3tab(path=os.path.join(HERE, "e3tab.npz")):
    """The reference's own E3 table (raddata/e3tab.dat, the file e3interp.f:33-44
    reads: a count, a header line, then tau, transmission, E3 per row) as a
    fixture, the values as the file holds them."""
    with open(REF_E3TAB) as f:
        n = int(f.readline().split()[0])
        f.readline()
        rows = np.array([[float(x) for x in f.readline().split()] for _ in range(n)])
    np.savez_compressed(path, tau=rows[:, 0], trans=rows[:, 1], e3=rows[:, 2])


def slice_tables(tables, i0, n):
    out = []
    for kt in tables:
        k2 = copy.copy(kt)
        k2.k = kt.k[i0:i0 + n].copy()
        k2.vmin = kt.vmin + i0 * kt.delv
        out.append(k2)
    return out


def run_ref(tables, cia, req, ilbl=0):
    outs = {}
    for prec in (64, 32):
        rr = ref.RefRunner(tables, cia, req, prec=prec, ilbl=ilbl)
        o, g, d = rr.spectrum()
        rr.close()
        outs[f"ref{prec}_output"] = o
        outs[f"ref{prec}_gtsurf"] = g
        outs[f"ref{prec}_doutputdq"] = d
    return outs


def cases():
    # tiny: 4 layers, 4 g, 2 k-gases + H2/He, 3 wavenumbers
    t, c, r = synth.config1(nwave=3, nlayer=4, nkgas=2, ng=4)
    yield "tiny", t, r, "L=4 ng=4 2 k-gases + H2/He, 3 wavenumbers"

    t, c, r = synth.config1(nwave=16, nlayer=20, nkgas=3, ng=8)
    yield "medium", t, r, "L=20 ng=8 3 k-gases + H2/He, 16 wavenumbers"

    t, c, r = synth.config1(nwave=1000, nlayer=120, nkgas=3, ng=20)
    i0, n = 496, 8
    ts = slice_tables(t, i0, n)
    r.vwave = r.vwave[i0:i0 + n].copy()
    yield "cfg1_slice", ts, r, "120 layers x 20 g x 3+2 gases, 5 paths, wavenumbers 496..503 of 1000"

    t, c, r = synth.config1(nwave=12, nlayer=30, nkgas=2, ng=10)
    r.tsurf = 150.0
    r.vem = np.array([500.0, 2000.0])
    r.emissivity = np.array([0.9, 0.95])
    r.amount = r.amount * 1e-4
    r.pp = r.pp * 1e-4
    r.imod = np.array([0, 1, 3, 3, 0], dtype=np.int32)
    yield "surface_imod", t, r, "tsurf>0 with emissivity, optically thin, imod 0/1/3 paths"

    t, c, r = synth.config1(nwave=12, nlayer=30, nkgas=3, ng=10)
    r.vwave = r.vwave[:-1] + 0.37 * t[0].delv
    yield "noncoincident", t, r, "wavenumbers between table points (rankk branch)"

    t, c, r = synth.config1(nwave=6, nlayer=25, nkgas=6, ng=12)
    yield "six_gases", t, r, "6 k-gases + H2/He"

    t, c, r = synth.config1(nwave=6, nlayer=25, nkgas=3, ng=8)
    r.temp[:5] = 300.0
    r.temp[-3:] = 20.0
    r.emtemp = r.temp[r.layinc - 1]
    r.press[:2] = 40.0
    yield "clamps", t, r, "layer T and P outside the table grid"

    t, c, r = synth.config1(nwave=6, nlayer=25, nkgas=3, ng=8)
    t[1].k[2] = 0.0
    t[0].k[3, :, :, 0] = 0.0
    t[2].k[4, :, :, :3] = 0.0
    yield "zeros", t, r, "all-zero slice, zero g=1 corners (linear interpolation), zero low-g"

    t, c, r = synth.config1(nwave=4, nlayer=200, nkgas=4, ng=16)
    yield "deep200", t, r, "200 layers, ng=16, 4 k-gases"

    t, c, r = synth.config1(nwave=5, nlayer=40, nkgas=1, ng=20)
    yield "one_kgas", t, r, "one k-gas + H2/He"

    t, c, r = synth.config1(nwave=5, nlayer=36, nkgas=3, ng=20)
    # a path that stops half way down (no surface term) and one at a
    # different scale with every layer counted twice
    r.nlayin = np.array([36, 18, 36, 36, 36], dtype=np.int32)
    r.scale[3] = 1.7
    yield "partial_paths", t, r, "a half-depth path (surface test fails) and an odd scale"

    t, c, r = synth.config1(nwave=8, nlayer=30, nkgas=2, ng=10)
    r = add_dust(r, 2, 40, seed=21)
    yield "dust", t, r, "two dust types (cloud deck, haze) on a 40-point cross-section grid"

    t, c, r = synth.config1(nwave=5, nlayer=24, nkgas=3, ng=8)
    r = add_dust(r, 1, 1, seed=22, peak_tau=(3.0,))
    r.isw = np.array([1, 3, 4, 6, 7], dtype=np.int32)
    yield "dust_const", t, r, ("one dust type with one cross-section (nsec = 1), optically "
                               "thick; isw skips gases 2 and 5")

    # wavelength space (ispace = 1): tables, vwave and emissivity grid in
    # micron; CIA and Rayleigh at 1e4/vwave cm-1, Planck in its wavelength
    # form, dust cross-sections on a micron grid
    t, r = wavelength_inputs(41, 7.0, 0.01)
    yield "wavelength", t, r, "ispace = 1 at 7.00-7.05 micron (CIA active)"

    t, r = wavelength_inputs(42, 2.0, 0.01)
    r.iray, r.itype = 1, 11
    r = add_dust(r, 1, 12, seed=43, peak_tau=(0.8,))
    yield "wavelength_ray_dust", t, r, "ispace = 1 at 2.00-2.05 micron with Rayleigh and one dust type"

    # disc flux (iform = 3, as the direct-imaging example runs nemesisdisc):
    # imod 3 outputs times 4 pi^2 ((radius1 + radextra) 1e5)^2 1e-18
    t, r = wavelength_inputs(44, 3.0, 0.02)
    r.iform, r.radius1, r.radextra = 3, 80000.0, 150.0
    yield "iform3", t, r, "iform = 3 total flux, ispace = 1, radius1 + radextra = 80150 km"

    t, c, r = synth.config1(nwave=5, nlayer=24, nkgas=2, ng=8)
    r.iform, r.radius1 = 1, 60268.0
    r.imod = np.array([3, 0, 3, 1, 3], dtype=np.int32)
    yield "iform1", t, r, "iform = 1 total flux on imod 3 paths; imod 0 and 1 paths unscaled"

    # para-H2 (FLAGH2P = 1, nparacon_all.f): a para fraction per layer from
    # 0.25 (deep, equilibrium at high T) rising aloft, with layers at the
    # fraction grid's ends, one inside its last interval and one with
    # fpara < 0 (the table's last fraction)
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.flagh2p = 1
    r.hfp = np.linspace(0.25, 0.95, r.nlayer)
    r.hfp[0], r.hfp[7], r.hfp[12], r.hfp[20] = 0.2, 1.0, 0.99, -1.0
    r.isw = np.arange(1, r.nparam + 1, dtype=np.int32)
    yield "parah2", t, r, "para-H2 fraction per layer (FLAGH2P = 1), 24 fractions"

    t, c, r = synth.config1(nwave=5, nlayer=24, nkgas=2, ng=8)
    r.flagh2p = 1
    r.hfp = np.linspace(0.3, 0.8, r.nlayer)
    r = add_dust(r, 1, 10, seed=45)
    r.isw = np.arange(1, r.nparam + 1, dtype=np.int32)
    yield "parah2_dust", t, r, "para-H2 with a 12-fraction table and one dust type (para column last)"

    # non-LTE source function (calc_nlte_t.f).  k1 = 1e-11 cm3/s and
    # A = 1e5 /s put f = k1 N / (A + k1 N) near 0.3 at the top and 1 at depth
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.inlte, r.nlte_k1, r.nlte_a = 1, 1e-11, 1e5
    yield "nlte_k1", t, r, "non-LTE, collisional form (inlte = 1, k1 = 1e-11, A = 1e5)"

    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.inlte = 2
    r.tsurf = 160.0
    yield "nlte_p", t, r, "non-LTE, pressure form (inlte = 2), surface at 160 K"

    # Titan-like: N2 and CH4 with H2, the N2 and CH4 CIA pairs from a CIA
    # table of the case's own (CASE_CIA).  Amounts are scaled by 0.05 and the
    # k-tables centred on 1e-25 so the column stays semi-transparent: the
    # N2 pairs then set 64% of the radiance and every output is within FP32
    # range
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=2, ng=10)
    ids = [6, 27]
    t = synth.make_ktables(ids, t[0].vmin, t[0].delv, t[0].npoint, ng=10, seed=31, mu=-25.0)
    r.idgas = np.array(ids + [synth.H2_ID, 22], dtype=np.int32)
    q = np.array([0.015, 3e-6, 0.001, 0.984])
    r.pp = r.press[:, None] * q[None, :] * 0.05
    r.amount = r.amount.sum(axis=1)[:, None] * q[None, :] * 0.05
    yield "titan_n2", t, r, "N2 + CH4 + H2 with the N2 and CH4 CIA pairs (own CIA table)"

    # Rayleigh (iray > 0, itype = 11 as forwardavfovX sets it).  Above the
    # CIA table's 1500 cm-1 the H2 and He Jacobian columns are Rayleigh alone.
    # At 1450 cm-1 (rayleigh_cia) the CIA terms dominate and Rayleigh adds
    # to their derivatives.
    for iray, tag, seed, v0 in ((1, "j", 11, 5000.0), (2, "v", 12, 5200.0),
                                (3, "a", 13, 4800.0), (4, "ls", 14, 5400.0),
                                (4, "cia", 15, 1450.0)):
        t, r = rayleigh_inputs(iray, seed, v0)
        yield f"rayleigh_{tag}", t, r, (f"iray={iray} itype=11 at {v0:.0f} cm-1"
                                        + (", CH4 and NH3 present" if iray == 4 else "")
                                        + (", CIA overlapping" if v0 < 1500 else ", no CIA"))

    # SCR radiometer paths (cirsradg_wave.f:1037-1102): cells are one-layer
    # paths (imod 13 sideband 1 - e^-tau, imod 14 wideband (1 + e^-tau)/2);
    # an imod 8 path multiplies the cell named by its layinc(1) with the
    # atmosphere named by layinc(2), per g-ordinate
    # (cell layers and scales chosen for transmissions of 0.4-0.9)
    for tag, atm, cells, cl in (("scr", 3, (13, 14), ((2, 1.0), (4, 1.0))),
                                ("scr_trans", 0, (14, 13), ((3, 1.0), (10, 2.0)))):
        t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
        r.imod = np.array([atm, cells[0], 8, cells[1], 8], dtype=np.int32)
        r.nlayin = np.array([30, 1, 2, 1, 2], dtype=np.int32)
        lay = r.layinc.copy()
        sc = r.scale.copy()
        (lay[1, 0], sc[1, 0]), (lay[3, 0], sc[3, 0]) = cl  # the cells' layers
        lay[2, :2] = [2, 1]                # cell 2 x atmosphere 1
        lay[4, :2] = [4, 1]                # cell 4 x atmosphere 1
        r.layinc, r.scale = lay, sc
        if atm == 0:
            r.amount, r.pp = r.amount * 0.02, r.pp * 0.02
        else:  # semi-transparent, with a surface gradient for the products to carry
            r.amount, r.pp = r.amount * 0.01, r.pp * 0.01
            r.tsurf = 150.0
            r.vem = np.array([500.0, 2000.0])
            r.emissivity = np.array([0.9, 0.95])
        yield tag, t, r, (f"SCR: imod {atm} atmosphere, imod {cells[0]} and {cells[1]} cells, "
                          "two imod 8 products")

    # with a stellar spectrum (iread = 999): iform 1 gives imod 3 paths as a
    # fraction of the star's total power (cirsradg_wave.f:912-925), iform 4
    # multiplies imod 0 transmissions by its flux at dist AU (:862-880)
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.iform, r.radius1 = 1, 71492.0
    r.imod = np.array([3, 3, 0, 3, 1], dtype=np.int32)
    yield "solar_ratio", t, r, "iform = 1 with a stellar spectrum: planet / star flux"

    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.amount, r.pp = r.amount * 0.02, r.pp * 0.02
    r.iform, r.dist = 4, 9.54
    r.imod = np.array([0, 3, 0, 1, 0], dtype=np.int32)
    yield "solar_trans", t, r, "iform = 4 with a stellar spectrum: transmissions times the flux at 9.54 AU"

    # the imod 8 rows are nlayin(1) (cirsradg_wave.f:1251-1252): path 1 a
    # 15-layer imod 1 path, the product's atmosphere path 2 of 30 layers
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.amount, r.pp = r.amount * 0.01, r.pp * 0.01
    r.imod = np.array([1, 3, 14, 8, 13], dtype=np.int32)
    r.nlayin = np.array([15, 30, 1, 2, 1], dtype=np.int32)
    lay, sc = r.layinc.copy(), r.scale.copy()
    lay[2, 0], sc[2, 0] = 4, 1.0
    lay[3, :2] = [3, 2]
    lay[4, 0], sc[4, 0] = 2, 1.0
    r.layinc, r.scale = lay, sc
    yield "scr_rows", t, r, "imod 8 of a 30-layer atmosphere summed over path 1's 15 rows"

    # N2's near-infrared CIA (n2n2cia.f on 4500-4825 cm-1, n2h2cia.f on
    # 3995-4995 cm-1): the Titan-like N2 + CH4 + H2 column at wavenumbers
    # inside, at the edges of and outside each table, in wavenumber and in
    # wavelength space
    for tag, ispace, grid, vw in (
            ("titan_nir", 0, (4440.0, 20.0, 30), [4460.0, 4500.0, 4580.0, 4660.0, 4820.0, 4840.0,
                                                  4980.0, 5000.0]),
            ("titan_nir_um", 1, (1.99, 0.01, 56), [2.0, 2.02, 2.08, 2.13, 2.17, 2.21, 2.25, 2.5])):
        t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=2, ng=10)
        ids = [6, 27]
        t = synth.make_ktables(ids, grid[0], grid[1], grid[2], ng=10, np_=10, nt=6, seed=33 + ispace,
                               mu=-25.0)
        r.idgas = np.array(ids + [synth.H2_ID, 22], dtype=np.int32)
        r.isogas = np.zeros(4, dtype=np.int32)
        q = np.array([0.015, 3e-6, 0.001, 0.984])
        r.pp = r.press[:, None] * q[None, :] * 0.05
        r.amount = r.amount.sum(axis=1)[:, None] * q[None, :] * 0.05
        r.ispace = ispace
        r.vwave = np.array(vw)
        r.vem = np.array([1.5, 3.0]) if ispace else np.array([3900.0, 5100.0])
        r.emissivity = np.array([1.0, 1.0])
        r.isw = np.arange(1, r.nparam + 1, dtype=np.int32)
        yield tag, t, r, (f"N2 + CH4 + H2 with N2's near-infrared CIA, ispace = {ispace}, "
                          "wavenumbers inside and outside both tables")

    # disc-integrated emission (imod 18, cirsradg_wave.f:1116-1211) with the
    # reference's E3 table (e3tab.npz): the layer weights are differences of
    # 2 pi E3 of the transmission, the surface term 2 pi E3(tr) B(tsurf) is
    # always on.  Amounts are scaled so
    # the column is semi-transparent and the surface shows; iform 1 or 3
    # multiply by the disc area 4 pi ((radius1 + radextra) 1e5)^2 (times
    # 1e-18 for 3)
    t, c, r = synth.config1(nwave=6, nlayer=30, nkgas=3, ng=10)
    r.amount, r.pp = r.amount * 0.05, r.pp * 0.05
    r.tsurf = 170.0
    r.imod = np.array([18, 0, 18, 3, 18], dtype=np.int32)
    r.nlayin[2] = 20
    yield "disc", t, r, "imod 18 paths (one stopping at layer 20) beside imod 0 and 3, tsurf 170 K"

    t, r = wavelength_inputs(46, 4.0, 0.02)
    r.amount, r.pp = r.amount * 0.05, r.pp * 0.05
    r.tsurf = 180.0
    r.imod = np.full(r.npath, 18, dtype=np.int32)
    r.iform, r.radius1, r.radextra = 3, 71492.0, 90.0
    yield "disc_iform3", t, r, "imod 18 with iform = 3 in wavelength space (total flux)"

    t, c, r = synth.config1(nwave=5, nlayer=24, nkgas=2, ng=8)
    r.amount, r.pp = r.amount * 0.05, r.pp * 0.05
    r.tsurf = 150.0
    r.vem = np.array([500.0, 2000.0])
    r.emissivity = np.array([0.85, 0.9])
    r.imod = np.array([18, 18, 3, 18, 1], dtype=np.int32)
    r.inlte, r.nlte_k1, r.nlte_a = 1, 1e-11, 1e5
    r = add_dust(r, 1, 10, seed=47)
    r.iform, r.radius1 = 1, 6052.0
    yield "disc_nlte_dust", t, r, "imod 18 with non-LTE, one dust type, emissivity < 1, iform = 1"

    # .kta forms: per-pressure temperature grids (NT < 0) and irregular
    # wavenumber grids (DELV < 0: nearest point; DELV = 0: merge of the
    # points either side)
    t, r = kta_forms_inputs(35, t2=True)
    yield "kta_t2", t, r, "k-tables with a temperature grid per pressure (NT < 0)"
    t, r = kta_forms_inputs(36, delv=-1.0)
    yield "kta_irregular_nearest", t, r, "irregular wavenumber grid, DELV < 0 (nearest point)"
    t, r = kta_forms_inputs(37, delv=0.0)
    yield "kta_irregular_between", t, r, "irregular wavenumber grid, DELV = 0 (two-point merge)"
    t, r = kta_forms_inputs(38, t2=True, delv=0.0)
    yield "kta_t2_irregular", t, r, "NT < 0 and DELV = 0 together"

    # the reference's own CIA table file (raddata/isotest.tab, dnu = 1 as
    # the Jupiter CIRS example's .cia has it), read by ciaread.f itself in
    # the REAL(4) build
    t, c, r = synth.config1(nwave=8, nlayer=60, nkgas=2, ng=10)
    yield "cia_isotest", t, r, "CIA from raddata/isotest.tab (dnu 1), 60 layers, 600-1400 cm-1"

    # gas continua (ngascon.f): H- free-free and bound-free, the CKD H2O
    # continuum (IH2O = 1) and the CH4/NH3 tables (ICH4, INH3)
    t, r = hminus_inputs(31, 2000.0, 2.0)
    yield "hminus_ir", t, r, "H- free-free (91, 80, 48) in a 1270-2730 K column at 2000 cm-1"
    t, r = hminus_inputs(32, 8000.0, 3.0)
    yield "hminus_nir", t, r, "H- bound-free and free-free at 8000 cm-1 (1.25 micron)"
    t, r = gascon_inputs(33, (27, 26), 1010.0, 37.0, 150.0, 0.8, 1.0)
    add_continuum_gas(r, 1, 2e-2, seed=33)
    yield "h2o_ckd", t, r, ("CKD H2O continuum (IH2O = 1, H2O with no k-table) at 1010-1195 cm-1, "
                            "190-290 K")
    t, r = gascon_inputs(34, (6, 11), 15300.0, 20.0, 900.0, 1.5, 1.0)
    yield "ch4_nh3_vis", t, r, "Karkoschka CH4 and Lutz-Owen NH3 continua (ICH4, INH3) at 0.65 micron"


def hminus_inputs(seed, v0, delv, nwave=6, nlayer=30, ng=10, vmr=(1e-9, 3e-7, 1e-2)):
    """A hot (1270-2730 K) 30-layer column with two k-gases, H2/He and then
    H- (91), e- (80) and H (48), all isotope 0, whose VMRs scatter about
    ``vmr``: H- bound-free (hmin_bf.f) acts above 6090 cm-1, free-free
    (hmin_ff.f, with the e- pressure) everywhere in the infrared."""
    ids = list(synth.KGAS_IDS[:2])
    t = synth.make_ktables(ids, v0, delv, nwave, ng=ng, np_=10, nt=6, seed=seed, tmin=800.0,
                           tmax=3000.0)
    r = synth.make_request(nlayer=nlayer, nkgas=2, nwave=nwave, seed=seed, table_vmin=v0,
                           table_delv=delv)
    r.temp = 1300.0 + 12.0 * (r.temp - 50.0)
    r.emtemp = r.temp[r.layinc - 1]
    ncol = r.amount[:, -1] / 0.14  # the He column over its mole fraction
    x = np.asarray(vmr)[None, :] * np.exp(np.random.default_rng(seed).normal(0, 0.3, (nlayer, 3)))
    r.idgas = np.concatenate([r.idgas, [91, 80, 48]]).astype(np.int32)
    r.isogas = np.zeros_like(r.idgas)
    r.amount = np.concatenate([r.amount, x * ncol[:, None]], axis=1)
    r.pp = np.concatenate([r.pp, x * r.press[:, None]], axis=1)
    r.isw = np.arange(1, r.ngas + 2, dtype=np.int32)
    return t, r


def gascon_inputs(seed, ids, v0, delv, tshift, tscale, vmr_scale, nwave=6, nlayer=30, ng=10):
    """A 30-layer column whose k-gases are ``ids`` (with H2/He), the
    temperatures ``tshift + tscale * T``, the k-gas amounts times
    ``vmr_scale`` (a gas-continuum case: H2O with IH2O = 1, CH4 with ICH4,
    NH3 with INH3)."""
    t = synth.make_ktables(list(ids), v0, delv, nwave, ng=ng, np_=10, nt=6, seed=seed,
                           tmin=min(40.0, tshift), tmax=tshift + tscale * 230.0)
    r = synth.make_request(nlayer=nlayer, nkgas=len(ids), nwave=nwave, seed=seed,
                           table_vmin=v0, table_delv=delv)
    r.idgas = np.array(list(ids) + [synth.H2_ID, synth.HE_ID], dtype=np.int32)
    r.temp = tshift + tscale * r.temp
    r.emtemp = r.temp[r.layinc - 1]
    n = len(ids)
    r.amount[:, :n] *= vmr_scale
    r.pp[:, :n] *= vmr_scale
    return t, r


def add_continuum_gas(r, gid, vmr, seed):
    """Append gas ``gid`` (isotope 0, no k-table) with a VMR scattered
    about ``vmr`` to every layer."""
    ncol = r.amount[:, -1] / 0.14  # the He column over its mole fraction
    x = vmr * np.exp(np.random.default_rng(seed).normal(0, 0.3, r.nlayer))
    r.idgas = np.concatenate([r.idgas, [gid]]).astype(np.int32)
    r.isogas = np.zeros_like(r.idgas)
    r.amount = np.concatenate([r.amount, (x * ncol)[:, None]], axis=1)
    r.pp = np.concatenate([r.pp, (x * r.press)[:, None]], axis=1)
    r.isw = np.arange(1, r.ngas + 2, dtype=np.int32)
    return r


def kta_forms_inputs(seed, t2=False, delv=None, nwave=7, nlayer=30, ng=10, nkgas=3):
    """A 30-layer request on k-tables with per-pressure temperature grids
    (``t2``: the .kta's NT < 0, each pressure's grid shifted and stretched)
    and/or an irregular wavenumber grid (``delv`` 0 or -1: DELV <= 0, the
    calculation wavenumbers between the grid points, or near them)."""
    rng = np.random.default_rng(seed)
    npt = 12
    ids = list(synth.KGAS_IDS[:nkgas])
    t = synth.make_ktables(ids, 800.0, 3.0, npt, ng=ng, np_=10, nt=6, seed=seed)
    vc = 800.0 + np.cumsum(rng.uniform(1.0, 5.0, npt))
    if t2:
        base = np.linspace(40.0, 220.0, 6)
        grid = np.array([base * (0.9 + 0.02 * i) + 4.0 * i for i in range(10)])
    for kt in t:
        if t2:
            kt.temp2 = grid.copy()
            kt.temp = grid[0].copy()
        if delv is not None:
            kt.delv, kt.vmin, kt.vcen = float(delv), float(vc[0]), vc.copy()
    r = synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed, table_vmin=800.0,
                           table_delv=3.0)
    if delv == 0:  # between grid points (read_klist.f:299-312 takes the one below)
        r.vwave = vc[2:2 + nwave] + rng.uniform(0.1, 0.9, nwave) * np.diff(vc)[2:2 + nwave]
    elif delv is not None:  # near grid points: the nearest is taken (:313-323)
        r.vwave = vc[1:1 + nwave] + rng.uniform(-0.3, 0.3, nwave)
    return t, r


def lbl_inputs(seed, nwave=6, nlayer=30, nkgas=3, t2=False, npoint=40, delv=0.5, v0=800.0):
    """A request on line-by-line tables (ILBL = 2): nkgas LBL gases then H2
    and He, which have no table; wavenumbers on grid points."""
    ids = list(synth.KGAS_IDS[:nkgas])
    t = synth.make_lbltables(ids, v0, delv, npoint, np_=10, nt=6, seed=seed)
    if t2:  # NT < 0: each pressure's temperature grid shifted and stretched
        base = np.linspace(40.0, 220.0, 6)
        grid = np.array([base * (0.9 + 0.02 * i) + 4.0 * i for i in range(10)])
        for kt in t:
            kt.temp2 = grid.copy()
            kt.temp = grid[0].copy()
    r = synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed, table_vmin=v0,
                           table_delv=delv * 5, wave_offset=1)
    return t, r


def lbl_cases():
    """The ILBL = 2 cases (get_klblg.f + noverlapg1.f), saved as lbl_*.npz.
    Their ref64 T column is not the routine's intended one (noverlapg1.f:108
    stores the first pair's derivative into grad; later gases read the
    first pair's dk/dT); the tests leave it out."""
    t, r = lbl_inputs(31)
    yield "basic", t, r, "L=30, 3 LBL gases + H2/He without tables, on grid points"

    t, r = lbl_inputs(32, t2=True)
    yield "t2", t, r, "LBL tables with per-pressure temperature grids (NT < 0)"

    # between grid points (the slice at or below v + 0.02 delv), just below
    # the grid (INT truncates toward zero: slice 1) and past its last point
    t, r = lbl_inputs(33, npoint=40, delv=0.5)
    rng = np.random.default_rng(33)
    vb = 800.0 + 0.5 * np.array([3, 7, 11, 20]) + rng.uniform(0.05, 0.45, 4)
    r.vwave = np.concatenate([vb, [800.0 - 0.3, 800.0 + 0.5 * 39 + 0.4]])
    yield "between", t, r, "LBL between grid points, below the first and past the last"

    # zero terms: an all-zero slice of gas 1 (k1 = 0 branch), a zero slice of
    # gas 2 with gas 1 of zero amount in some layers (a zero running sum
    # before gas 3 zeroes the columns before it, noverlapg1.f:123-128),
    # corners <= 0 (KLOG false, linear interpolation)
    t, r = lbl_inputs(34, nwave=6)
    iw = ((r.vwave - t[0].vmin) / t[0].delv + 0.02).astype(int)
    t[0].k[iw[1]] = 0.0
    t[1].k[iw[2]] = 0.0
    t[1].k[iw[3]] = 0.0
    t[0].k[iw[3], :, :3] = 0.0
    t[2].k[iw[4], 2:5, :] = 0.0
    r.amount = r.amount.copy()
    r.amount[5:15, 0] = 0.0
    yield "zeros", t, r, "LBL zero slices, zero corners and zero amounts"

    # two tables for gas 2 (read_klbllist.f:190-198: the last one listed
    # wins) whose interior pressures differ from the others', so the shared
    # P/T grid is the last file's (read_klblhead.f fills one grid per file)
    t, r = lbl_inputs(35)
    extra = copy.deepcopy(t[1])
    extra.k = extra.k * 3.0
    last = copy.deepcopy(t[2])
    last.press = last.press.copy()
    last.press[3:7] *= 1.3
    t = [t[0], t[1], extra, last]
    yield "last_table", t, r, "LBL: two tables for one gas, the last file's P/T grid"


def cia_fixture(name="isotest", dnu=1.0):
    """raddata/<name>.tab as a fixture: its temperatures and kcia as the
    file holds them (REAL*4), read as ciaread.f:77-89 reads it."""
    from radtrancode_amd.cia import read_cia_table
    fn = f"{name}.tab"
    temps, k = read_cia_table(os.path.join(ref.REF_TREE, "raddata", fn))
    np.savez_compressed(os.path.join(HERE, f"cia_{name}.npz"), dnu=np.array(dnu), temps=temps,
                        kcia=k.astype(np.float32), table=np.array(fn))


def gascon_tables(path=os.path.join(HERE, "gascon_tables.npz")):
    """The gas-continuum data as the reference holds it: the CKD H2O tables
    of ngascon.f's BLOCK DATA with /COMCON/ (copied out of the REAL(8)
    build's COMMON blocks by the harness), raddata/karkoschka.dat and
    raddata/lutzowen.dat (read as ch4cont.f and nh3cont.f read them)."""
    from radtrancode_amd.gascon import read_karkoschka, read_lutzowen
    grid, npt, s96, s60, f96, con = ref.h2o_tables()
    cw, ck = read_karkoschka(os.path.join(ref.REF_TREE, "raddata", "karkoschka.dat"))
    nw, nk = read_lutzowen(os.path.join(ref.REF_TREE, "raddata", "lutzowen.dat"))
    assert npt == 2001
    np.savez_compressed(path, h2o_grid=grid, h2o_s296=s96, h2o_s260=s60, h2o_f296=f96,
                        comcon=con, ch4_wl=cw, ch4_k=ck, nh3_wl=nw, nh3_k=nk)


def gascon_vectors(n=64, seed=12):
    """The reference's own HMIN_BF and HMIN_FF (each from a zero sum),
    CH4CONT and NH3CONT at log-uniform wavenumbers (and temperatures for
    HMIN_FF), with both ranges' ends, REAL(8) build."""
    rng = np.random.default_rng(seed)
    v = np.exp(rng.uniform(np.log(500.0), np.log(60000.0), n))
    v = np.concatenate([v, [1e4 / 1.6419 * (1 + 1e-6), 1e4 / 0.3645 * (1 - 1e-6),
                            1e4 / 0.3645 * (1 + 1e-6), 1e4 / 0.1823 * (1 - 1e-6)]])
    t = rng.uniform(800.0, 6000.0, v.shape[0])
    vt = np.concatenate([np.exp(rng.uniform(np.log(9000.0), np.log(34000.0), n)),
                         1e4 / np.array([0.3004, 1.06, 0.547615, 0.660327, 0.55, 0.65])])
    return dict(v=v, t=t, vt=vt,
                ref64_bf=np.array([ref.hmin(64, "bf", x) for x in v]),
                ref64_ff=np.array([ref.hmin(64, "ff", x, y) for x, y in zip(v, t)]),
                ref64_ch4=np.array([ref.tabcont(64, "ch4", x) for x in vt]),
                ref64_nh3=np.array([ref.tabcont(64, "nh3", x) for x in vt]))


def conv_vectors(seed=21):
    """The reference's own CIRSCONV on seeded spectra: boxcars that pass
    either end of the grid (extrapolated end points), a wide one, an FWHM
    table (FWHMEXIST), filter functions (fwhm < 0, a runname.fil file) and
    fwhm = 0, on an irregular 120-point grid; REAL(8) build."""
    rng = np.random.default_rng(seed)
    v = np.unique(np.sort(600.0 + 100.0 * rng.random(120)))
    y = 1e-7 * (1 + 0.3 * np.sin(v / 3.0)) * np.exp(rng.normal(0, 0.05, v.size))
    out = dict(v=v, y=y)
    cases = {"box_inside": (5.0, np.linspace(606.0, 694.0, 23), None),
             "box_ends": (5.0, np.linspace(600.0, 700.0, 21), None),
             "box_wide": (12.5, np.linspace(598.0, 702.0, 9), None),
             "box_table": (5.0, np.linspace(605.0, 695.0, 19), (np.array([550., 650., 750.]),
                                                               np.array([3.0, 6.0, 4.0]))),
             "identity": (0.0, v[:40].copy(), None)}
    for name, (fw, vc, tab) in cases.items():
        vf, xf = (None, None) if tab is None else tab
        out[f"{name}_fwhm"] = np.array(fw)
        out[f"{name}_vconv"] = vc
        if tab is not None:
            out[f"{name}_vfwhm"], out[f"{name}_xfwhm"] = vf, xf
        out[f"{name}_ref64"] = ref.cirsconv(64, fw, v, y, vc, vf, xf)
    vc = np.linspace(610.0, 690.0, 9)
    fil = [(c, np.linspace(c - 4.0, c + 4.0, 17), np.exp(-0.5 * (np.linspace(-4, 4, 17) / 1.7) ** 2))
           for c in vc]
    out["filter_vconv"] = vc
    out["filter_n"] = np.array([len(f[1]) for f in fil], dtype=np.int32)
    out["filter_v"] = np.concatenate([f[1] for f in fil])
    out["filter_f"] = np.concatenate([f[2] for f in fil])
    out["filter_ref64"] = ref.cirsconv(64, -1.0, v, y, vc, filters=fil)
    return out


def add_dust(r, ncont, nsec, seed, peak_tau=(1.5, 0.4)):
    """Dust on a request (cirsradg_wave.f:550-577): particle type 0 is a
    cloud deck near 0.7 bar, type 1 a haze spread over the upper layers;
    extinction cross-sections ~3e-8 cm2 varying smoothly over a ``nsec``-point
    grid that spans the wavenumbers (``nsec = 1``: one constant value)."""
    rng = np.random.default_rng(seed)
    L = r.nlayer
    lp = np.log(r.press * 1.01325)  # ln(bar)
    prof = [np.exp(-0.5 * ((lp - np.log(0.7)) / 0.35) ** 2),
            np.exp(-0.5 * ((lp - np.log(0.01)) / 1.5) ** 2)]
    if nsec == 1:
        vsec = np.array([0.5 * (r.vwave[0] + r.vwave[-1])])
    else:
        vsec = np.linspace(r.vwave[0] - 7.3, r.vwave[-1] + 11.9, nsec)
    xsec = np.empty((ncont, nsec))
    cont = np.empty((L, ncont))
    for k in range(ncont):
        a, b = rng.uniform(0.2, 0.5, 2)
        ph = rng.uniform(0, 2 * np.pi)
        x = (vsec - vsec[0]) / max(vsec[-1] - vsec[0], 1.0)
        xsec[k] = 3e-8 * (1 + a * np.sin(2 * np.pi * x + ph) + b * x)
        c = prof[k % 2] * rng.uniform(0.8, 1.2, L)
        # scale so that the column optical depth is peak_tau at the grid centre
        cont[:, k] = c * peak_tau[k % 2] / (c.sum() * np.median(xsec[k]))
    r.cont, r.vsec, r.xsec = cont, vsec, xsec
    r.isw = np.arange(1, r.ngas + 2 + ncont, dtype=np.int32)
    return r


def wavelength_inputs(seed, w0, dw, nwave=6, nkgas=3, nlayer=30, ng=10):
    """A request in wavelength space: the table grid and vwave in micron."""
    ids = list(synth.KGAS_IDS[:nkgas])
    t = synth.make_ktables(ids, w0, dw, nwave, ng=ng, np_=10, nt=6, seed=seed)
    r = synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed,
                           table_vmin=w0, table_delv=dw)
    r.ispace = 1
    return t, r


def rayleigh_inputs(iray, seed, v0, nkgas=3, nlayer=30, ng=10, nwave=6, delv=2.0,
                    np_=10, nt=6):
    """A 30-layer request at ``v0`` with Rayleigh on.  For iray=4 the second
    k-gas is CH4 (id 6), so H2, He, CH4 and NH3 (id 11) all set the mixture
    of rayleighls.f."""
    ids = list(synth.KGAS_IDS[:nkgas])
    if iray == 4:
        ids[1] = 6
    t = synth.make_ktables(ids, v0, delv, nwave, ng=ng, np_=np_, nt=nt, seed=seed)
    r = synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed,
                           table_vmin=v0, table_delv=delv)
    r.idgas = np.array(ids + [synth.H2_ID, synth.HE_ID], dtype=np.int32)
    r.iray = iray
    r.itype = 11
    return t, r


RAYLEIGH_FN = {1: "rayleighj", 2: "rayleighv", 3: "rayleigha", 4: "rayleighls"}


def rayleigh_vectors(n=64, seed=7):
    """The reference's own RAYLEIGHJ/V/A/LS on log-spaced wavenumbers
    (100..30000 cm-1), pressures (1e-4..100 atm), temperatures (50..800 K)
    and mixtures (He/H2, CH4/H2, NH3 fraction), REAL(8) and REAL(4) builds."""
    rng = np.random.default_rng(seed)
    v = np.exp(rng.uniform(np.log(100.0), np.log(30000.0), n))
    p = np.exp(rng.uniform(np.log(1e-4), np.log(100.0), n))
    t = rng.uniform(50.0, 800.0, n)
    mix = rng.uniform(0.0, 0.3, (n, 3))
    out = dict(v=v, p=p, t=t, mix=mix)
    for iray in (1, 2, 3, 4):
        for prec in (64, 32):
            out[f"ref{prec}_iray{iray}"] = np.array(
                [ref.rayleigh(prec, iray, v[i], p[i], t[i], *mix[i]) for i in range(n)])
    return out


def _band(vwave, dnu, n, half):
    keep = np.zeros(n, dtype=bool)
    for v in vwave:
        i = int(v / dnu)
        keep[max(0, i - half):i + half + 2] = True
    return keep


def banded_cia(cia, vwave, half=3):
    """The table with every wavenumber row farther than ``half`` rows from
    the case's wavenumbers set to zero: ciaread.f reads only the two rows
    bracketing a wavenumber, so results are unchanged and the stored case
    stays small."""
    keep = _band(vwave, cia.dnu, cia.kcia.shape[0], half)
    k = cia.kcia.copy()
    k[~keep] = 0.0
    para = None
    if cia.para is not None:
        pa = cia.para
        kp = _band(vwave, pa.dnu, pa.kh2h2.shape[0], half)
        arrs = [a.copy() for a in (pa.kh2h2, pa.kh2he, pa.kcia5)]
        for a in arrs:
            a[~kp] = 0.0
        para = synth.ParaCIATable(dnu=pa.dnu, temps=pa.temps, fp=pa.fp, kh2h2=arrs[0],
                                  kh2he=arrs[1], kcia5=arrs[2])
    return synth.CIATable(dnu=cia.dnu, temps=cia.temps, kcia=k, para=para)


# cases that run with a CIA table of their own: every pair the Titan case
# touches (H2-H2, H2-He, H2-N2, N2-CH4, N2-N2, CH4-CH4, H2-CH4)
CASE_CIA = {"titan_n2": dict(seed=5, pairs=(0, 1, 4, 5, 6, 7, 8)),
            "parah2": dict(seed=7, pairs=(0, 1)), "parah2_dust": dict(seed=8, pairs=(0, 1))}
# cases whose CIA table also carries a para-H2 table (FLAGH2P = 1)
CASE_PARA = {"parah2": dict(seed=9), "parah2_dust": dict(seed=10, npara=12)}
# cases with imod 18 paths, which run with the E3 table
CASE_E3 = ("disc", "disc_iform3", "disc_nlte_dust")
# cases with N2 at 3995-4995 cm-1, which need N2's near-infrared CIA
CASE_N2NIR = ("titan_nir", "titan_nir_um")
# cases run with a stellar spectrum (/SOLARDAT/, iread = 999)
CASE_SOLAR = ("solar_ratio", "solar_trans")
# cases run with a reference CIA table file (its fixture name)
CASE_CIA_FILE = {"cia_isotest": "isotest"}
# cases run with /GASCONDEF/ flags (ih2o, ich4, io3, inh3, iuvscat)
CASE_GASCON = {"hminus_ir": (0, 0, 0, 0, 0), "hminus_nir": (0, 0, 0, 0, 0),
               "h2o_ckd": (1, 0, 0, 0, 0), "ch4_nh3_vis": (0, 1, 0, 1, 0)}


def main(names=()):
    """Regenerate every fixture, or only the named cases (``rayleigh`` is the
    function-level Rayleigh vectors)."""
    cia = synth.make_cia()
    if not names:
        save_cia(os.path.join(HERE, "cia.npz"), cia)
    if not names or "e3tab" in names:
        save_e3tab()
    if not names or "n2nir" in names:
        save_n2nir()
    if not names or "rayleigh" in names:
        np.savez_compressed(os.path.join(HERE, "rayleigh.npz"), **rayleigh_vectors())
    if not names or "cia_fixture" in names:
        cia_fixture()
    if not names or "conv" in names:
        np.savez_compressed(os.path.join(HERE, "conv.npz"), **conv_vectors())
    if not names or "gascon" in names:
        gascon_tables()
        np.savez_compressed(os.path.join(HERE, "gascon_funcs.npz"), **gascon_vectors())
    for name, t, r, note in lbl_cases():
        if names and "lbl_" + name not in names:
            continue
        outs = run_ref(t, cia, r, ilbl=2)
        path = os.path.join(HERE, f"lbl_{name}.npz")
        save_case(path, t, r, outs, note)
        print(f"lbl_{name:12s} {os.path.getsize(path) / 1024:8.1f} KiB  "
              f"|R|max {np.abs(outs['ref64_output']).max():.3e}")
    for name, t, r, note in cases():
        if names and name not in names:
            continue
        own = None
        if name in CASE_CIA:
            own = synth.make_cia(**CASE_CIA[name])
            if name in CASE_PARA:
                own.para = synth.make_para_cia(**CASE_PARA[name])
            own = banded_cia(own, r.vwave, half=1 if name in CASE_PARA else 3)
        use = own or cia
        e3 = nir = None
        if name in CASE_E3:
            e3 = load_e3()
            use = copy.copy(use)
            use.e3 = e3
        if name in CASE_N2NIR:  # for the oracle and the GPU; the reference has its own
            nir = load_n2nir()
        sol = None
        if name in CASE_SOLAR:
            sol = synth.make_solar(500.0, 1500.0)
            use = copy.copy(use)
            use.solar = sol
        cfx = CASE_CIA_FILE.get(name)
        if cfx is not None:
            from tests.golden.golden_io import load_cia_fixture
            use = load_cia_fixture(cfx)
        gflags = CASE_GASCON.get(name)
        if gflags is not None:
            use = copy.copy(use)
            use.gascon = load_gascon(gflags)
        outs = run_ref(t, use, r)
        path = os.path.join(HERE, f"case_{name}.npz")
        save_case(path, t, r, outs, note, cia=own, e3=e3, n2nir=nir, solar=sol,
                  gascon_flags=gflags, hmin_static=bool(np.isin(r.idgas, (48, 91)).any()),
                  cia_fixture=cfx)
        print(f"{name:16s} {os.path.getsize(path) / 1024:8.1f} KiB  "
              f"|R|max {np.abs(outs['ref64_output']).max():.3e}")


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
