"""Write the golden vectors of the optimal-estimation step and the layer ->
state maps from the reference's own routines.

Run in a container where ``/root/reference`` is present and
``make -C oracle ref`` has built ``oracle/_ref/liboeref64.so``:

    python tests/golden/make_golden_oe.py

Cases:
* ``oe_small``: nx = 12, ny = 30 (calc_gain_matrix's nx < ny branch);
* ``oe_wide``:  nx = 40, ny = 25 (its nx >= ny branch);
* ``oe_medium``: nx = 121, ny = 300;
* ``oe_cfg4``: nx = 363, ny = 1000, the batched-retrieval shape.  Its inputs
  are regenerated from the seed by :func:`oe_inputs`; only the outputs, the
  row sums of the gain matrix and the diagonal of the averaging kernels are
  stored;
* ``map_small``: map2pro + map2xvec for 3 paths x 20 layers x 2 gases,
  21 levels, 30 state elements;
* ``map_dust``: the same with 2 dust types mapped through DCO;
* ``map_para``: one dust type (DCO) and the para-H2 fraction (DFP);
* ``map_scr``: path 3 an imod 8 product, mapped over path 1's layers.

``python tests/golden/make_golden_oe.py maps`` writes only the two map cases.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def oe_inputs(nx, ny, seed):
    """Seeded OE inputs: K ~ N(0, 1), Sa = A A^T / nx + 0.5 I, diagonal
    Se in [0.5, 2], xa ~ N(0, 1), xn = xa + 0.1 N(0, 1), y ~ N(0, 1),
    yn = y + 0.2 N(0, 1)."""
    rng = np.random.default_rng(seed)
    kk = rng.normal(size=(ny, nx))
    a = rng.normal(size=(nx, nx))
    sa = a @ a.T / nx + 0.5 * np.eye(nx)
    se = rng.uniform(0.5, 2.0, ny)
    xa = rng.normal(size=nx)
    xn = xa + 0.1 * rng.normal(size=nx)
    y = rng.normal(size=ny)
    yn = y + 0.2 * rng.normal(size=ny)
    return dict(kk=kk, sa=sa, se_var=se, xa=xa, xn=xn, y=y, yn=yn)


def map_inputs(seed=11, npath=3, L=20, ngas=2, npro=21, nv=30, ncont=0, para=False, scr=False):
    """Layer Jacobians and maps; with ``ncont`` dust types the Jacobian and
    xmap gain their columns and a DCO matrix is added; with ``scr`` path 3
    is an imod 8 product (mapped over path 1's 20 layers)."""
    rng = np.random.default_rng(seed)
    npar = ngas + 1 + ncont + (1 if para else 0)
    m = dict(dq=rng.normal(size=(npath, L, npar)), nlayin=np.array([20, 18, 20]),
             layinc=np.array([rng.permutation(L) + 1 for _ in range(npath)]),
             dam=rng.normal(size=(L, npro)), dte=rng.normal(size=(L, npro)),
             xmap=rng.normal(size=(nv, npar, npro)))
    if ncont:
        m.update(dco=rng.normal(size=(L, npro)), ncont=ncont)
    if para:
        m.update(dfp=rng.normal(size=(L, npro)))
    if scr:
        m.update(imod=np.array([3, 3, 8]), nlayin=np.array([20, 18, 2]))
    return m


def main(only_maps=False):
    from oracle import oe_ref
    assert oe_ref.available(), "build oracle/_ref/liboeref64.so first (make -C oracle ref)"
    for name, nx, ny, seed, full in [] if only_maps else [("oe_small", 12, 30, 1, True), ("oe_wide", 40, 25, 2, True),
                                     ("oe_medium", 121, 300, 3, True),
                                     ("oe_cfg4", 363, 1000, 4, False)]:
        inp = oe_inputs(nx, ny, seed)
        r = oe_ref.oe_step(**inp)
        out = dict(nx=nx, ny=ny, seed=seed, x_out=r["x_out"], phi=r["phi"], chisq=r["chisq"],
                   dd_rowsum=r["dd"].sum(axis=1), aa_diag=np.diag(r["aa"]).copy())
        if full:
            out.update({k: v for k, v in inp.items()})
            out.update(dd=r["dd"], aa=r["aa"], sai=r["sai"])
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, nx, ny, "phi", r["phi"])
    for name, kw in (("map_small", {}), ("map_dust", dict(seed=12, ncont=2)),
                     ("map_para", dict(seed=13, ncont=1, para=True)),
                     ("map_scr", dict(seed=14, scr=True))):
        m = map_inputs(**kw)
        dx = oe_ref.map_state(**m)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), doutdx=dx, **m)
        print(name, dx.shape)


if __name__ == "__main__":
    main(only_maps=sys.argv[1:] == ["maps"])
