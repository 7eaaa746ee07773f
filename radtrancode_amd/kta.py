"""Reader and writer for the reference's ``.kta`` k-table files.

A ``.kta`` file is a Fortran direct-access file of fixed-width records
(``ISYS()`` bytes each: 4 for REAL(4) builds, ``radtran/rtm_util/isys.f:28``).
Layout, from ``radtran/cirsrad/read_khead.f:38-109`` and the writer
``radtran/radtran/calc_fnktablec.f:308-367,593-609``:

====================  =====================================================
record                content
====================  =====================================================
1                     IREC0, first data record (integer)
2                     NPOINT (integer)
3, 4, 5               VMIN, DELV, FWHM (real)
6, 7, 8               NP, NT, NG (integer)
9, 10                 IDGAS, ISOGAS (integer)
11 .. 10+NG           G_ORD
11+NG .. 10+2NG       DEL_G
11+2NG, 12+2NG        unused
13+2NG ..             PRESS(NP) in atm, then TEMP(NT) (or TEMP2(NP,|NT|))
then (DELV <= 0)      VCEN(NPOINT) (read_khead.f:93-100)
IREC0 ..              k * 1e20 for each wavenumber: NP*|NT|*NG values,
                      g fastest, then T, then P
====================  =====================================================

Integers occupy the first 4 bytes of a record.  ``recl=8`` gives the
layout the REAL(8) build of the reference reads.
"""
from __future__ import annotations

import numpy as np

from .synth import KTable


def _header_records(kt: KTable, vcen=None):
    ng, np_, nt = kt.ng, kt.np_, kt.nt
    ntemp = np_ * nt if kt.temp2 is not None else nt
    if vcen is None:
        vcen = kt.vcen
    nrec_head = 10 + 2 * ng + 2 + np_ + ntemp + (0 if kt.delv > 0 else kt.npoint)
    irec0 = nrec_head + 1
    ints = {1: irec0, 2: kt.npoint, 6: np_, 7: kt.nt_signed, 8: ng, 9: kt.idgas, 10: kt.isogas}
    reals = {3: kt.vmin, 4: kt.delv, 5: kt.fwhm}
    for j in range(ng):
        reals[11 + j] = kt.g_ord[j]
        reals[11 + ng + j] = kt.delg[j]
    r = 11 + 2 * ng + 2
    for j in range(np_):
        reals[r + j] = kt.press[j]
    r += np_
    temps = np.asarray(kt.temp2).reshape(-1) if kt.temp2 is not None else kt.temp
    for j in range(ntemp):  # TEMP2(I,J) with J fastest (read_khead.f:86-92)
        reals[r + j] = temps[j]
    r += ntemp
    if kt.delv <= 0:
        for j in range(kt.npoint):
            reals[r + j] = vcen[j]
    return irec0, nrec_head, ints, reals


def write_kta(path: str, kt: KTable, recl: int = 4, vcen=None) -> None:
    """Write ``kt`` as a ``.kta`` file with ``recl``-byte records (4 or 8)."""
    if recl not in (4, 8):
        raise ValueError("recl must be 4 or 8")
    rdt = np.float32 if recl == 4 else np.float64
    irec0, nhead, ints, reals = _header_records(kt, vcen)
    head = np.zeros((nhead, recl), dtype=np.uint8)
    for r, v in ints.items():
        head[r - 1, :4] = np.frombuffer(np.int32(v).tobytes(), dtype=np.uint8)
    for r, v in reals.items():
        head[r - 1, :] = np.frombuffer(rdt(v).tobytes(), dtype=np.uint8)
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(np.ascontiguousarray(kt.k, dtype=rdt).tobytes())


def read_kta(path: str, recl: int = 4) -> KTable:
    """Read a ``.kta`` file into a :class:`KTable` (``temp2`` set for NT < 0,
    ``vcen`` for DELV <= 0)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size % recl:
        raise ValueError(f"{path}: size is not a multiple of {recl}")
    rec = raw.reshape(-1, recl)
    rdt = np.float32 if recl == 4 else np.float64

    def gi(r):
        return int(rec[r - 1, :4].view(np.int32)[0])

    def gr(r):
        return float(rec[r - 1].view(rdt)[0])

    irec0, npoint, np_, nts, ng = gi(1), gi(2), gi(6), gi(7), gi(8)
    nt = abs(nts)
    vmin, delv, fwhm = gr(3), gr(4), gr(5)
    g_ord = np.array([gr(11 + j) for j in range(ng)])
    delg = np.array([gr(11 + ng + j) for j in range(ng)])
    r = 11 + 2 * ng + 2
    press = np.array([gr(r + j) for j in range(np_)])
    r += np_
    temp2 = None
    if nts < 0:
        temp2 = np.array([gr(r + j) for j in range(np_ * nt)]).reshape(np_, nt)
        temp = temp2[0].copy()
        r += np_ * nt
    else:
        temp = np.array([gr(r + j) for j in range(nt)])
        r += nt
    vcen = np.array([gr(r + j) for j in range(npoint)]) if delv <= 0 else None
    ntab = np_ * nt * ng
    data = rec[irec0 - 1: irec0 - 1 + npoint * ntab].copy().view(rdt).reshape(
        npoint, np_, nt, ng).astype(np.float64)
    return KTable(idgas=gi(9), isogas=gi(10), vmin=vmin, delv=delv, fwhm=fwhm,
                  press=press, temp=temp, g_ord=g_ord, delg=delg, k=data, temp2=temp2,
                  vcen=vcen)


# ------------------------------------------------------------------ .lta
# Line-by-line tables (/LBLTABLE/ ILBL = 2), read by
# radtran/cirsrad/read_klblhead.f:1-75 and written by
# radtran/radtran/calc_lbltable.f:228-368:
#
#   record 1 IREC0, 2 NPOINT (integer); 3 VMIN, 4 DELV (real);
#   5 NP, 6 NT, 7 IDGAS, 8 ISOGAS (integer); 9 .. PRESS(NP) in atm, then
#   TEMP(NT) or TEMP2(NP,|NT|) (J fastest); IREC0 = 12 + NP + NP*|NT| (NT < 0)
#   or 12 + NP + NT; from IREC0, k * 1e20 per wavenumber: NP*|NT| values,
#   T fastest, then P.
#
# An LBL table is held as a KTable with one g-ordinate of weight 1 (the
# ng = 1, delg = 1 that cirsradg_wave.f:336-339 sets).


def lbl_table(idgas, isogas, vmin, delv, press, temp, k, temp2=None) -> KTable:
    """An LBL table as a one-ordinate :class:`KTable`; ``k`` is
    ``[npoint, np, nt]`` (k * 1e20)."""
    k = np.asarray(k, dtype=np.float64)
    return KTable(idgas=int(idgas), isogas=int(isogas), vmin=float(vmin), delv=float(delv),
                  fwhm=0.0, press=np.asarray(press, dtype=np.float64),
                  temp=np.asarray(temp, dtype=np.float64), g_ord=np.array([1.0]),
                  delg=np.array([1.0]), k=k.reshape(k.shape + (1,)),
                  temp2=None if temp2 is None else np.asarray(temp2, dtype=np.float64))


def write_lta(path: str, kt: KTable, recl: int = 4) -> None:
    """Write a one-ordinate table as a ``.lta`` file with ``recl``-byte records."""
    if recl not in (4, 8):
        raise ValueError("recl must be 4 or 8")
    if kt.ng != 1:
        raise ValueError("an .lta table has one g-ordinate")
    rdt = np.float32 if recl == 4 else np.float64
    np_, nt = kt.np_, kt.nt
    ntemp = np_ * nt if kt.temp2 is not None else nt
    irec0 = 10 + np_ + ntemp + 2
    head = np.zeros((irec0 - 1, recl), dtype=np.uint8)
    for r, v in {1: irec0, 2: kt.npoint, 5: np_, 6: kt.nt_signed, 7: kt.idgas, 8: kt.isogas}.items():
        head[r - 1, :4] = np.frombuffer(np.int32(v).tobytes(), dtype=np.uint8)
    temps = np.asarray(kt.temp2).reshape(-1) if kt.temp2 is not None else kt.temp
    reals = {3: kt.vmin, 4: kt.delv}
    for j in range(np_):
        reals[9 + j] = kt.press[j]
    for j in range(ntemp):
        reals[9 + np_ + j] = temps[j]
    for r, v in reals.items():
        head[r - 1, :] = np.frombuffer(rdt(v).tobytes(), dtype=np.uint8)
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(np.ascontiguousarray(kt.k[..., 0], dtype=rdt).tobytes())


def read_lta(path: str, recl: int = 4) -> KTable:
    """Read a ``.lta`` file into a one-ordinate :class:`KTable`."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size % recl:
        raise ValueError(f"{path}: size is not a multiple of {recl}")
    rec = raw.reshape(-1, recl)
    rdt = np.float32 if recl == 4 else np.float64

    def gi(r):
        return int(rec[r - 1, :4].view(np.int32)[0])

    def gr(r):
        return float(rec[r - 1].view(rdt)[0])

    irec0, npoint, np_, nts = gi(1), gi(2), gi(5), gi(6)
    nt = abs(nts)
    press = np.array([gr(9 + j) for j in range(np_)])
    r = 9 + np_
    temp2 = None
    if nts < 0:
        temp2 = np.array([gr(r + j) for j in range(np_ * nt)]).reshape(np_, nt)
        temp = temp2[0].copy()
    else:
        temp = np.array([gr(r + j) for j in range(nt)])
    data = rec[irec0 - 1: irec0 - 1 + npoint * np_ * nt].copy().view(rdt).reshape(
        npoint, np_, nt).astype(np.float64)
    return lbl_table(gi(7), gi(8), gr(3), gr(4), press, temp, data, temp2)
