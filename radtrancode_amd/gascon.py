"""Readers for the gas-continuum data files ngascon.f's helpers read.

* :func:`read_karkoschka` -- ``karkoschka.dat``, as ``ch4cont.f:35-43``
  reads it: 1810 rows of wavelength (nm) and CH4 coefficient; the wavelength
  is returned in micron (the reference's ``0.001*x``).
* :func:`read_lutzowen` -- ``lutzowen.dat``, as ``nh3cont.f:35-44`` reads it:
  three header lines, then 130 rows of wavelength (micron) and NH3
  coefficient (m-1 amagat-1).

The tables go into ``synth.GasCon`` (``ch4_wl``/``ch4_k``,
``nh3_wl``/``nh3_k``) for :func:`radtrancode_amd.cirsradg.register_gascon`.
"""
from __future__ import annotations

import numpy as np


def read_karkoschka(path: str, n: int = 1810):
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2:
                rows.append((float(parts[0]), float(parts[1])))
            if len(rows) == n:
                break
    if len(rows) != n:
        raise ValueError(f"{path}: {len(rows)} rows, expected {n}")
    a = np.array(rows)
    return 0.001 * a[:, 0], a[:, 1]


def read_lutzowen(path: str, n: int = 130):
    with open(path) as f:
        lines = f.readlines()[3:]
    rows = [tuple(float(x) for x in l.split()[:2]) for l in lines if len(l.split()) >= 2][:n]
    if len(rows) != n:
        raise ValueError(f"{path}: {len(rows)} rows, expected {n}")
    a = np.array(rows)
    return a[:, 0], a[:, 1]
