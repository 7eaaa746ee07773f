"""GPU parity of the line-by-line path (/LBLTABLE/ ILBL = 2: get_klblg.f and
noverlapg1.f on one-ordinate tables).

* every lbl_* golden case in each form: radiance, gtsurf and the gas
  columns of the Jacobian against the reference built with REAL promoted to
  8 bytes; the T column (undefined in the reference, noverlapg1.f:108)
  against the FP64 oracle; the 16-bit forms against the oracle run on the
  tables they decode;
* a 120-layer request on 5 line-by-line gases against the oracle;
* .lta files through radg_read_klbllist (4- and 8-byte records) against the
  same tables registered from arrays; a batch of requests of different
  sizes against each alone; the entry point that allocates nothing against
  the batch entry point; ILBL switched back to 0 leaves the k-table path as
  it was; the refusals (ILBL = 1, a wavenumber the tables cannot reach).
"""
import ctypes as C
import os

import numpy as np
import pytest

from radtrancode_amd import kta, synth
from radtrancode_amd._native import RadgError
from tests.bf16_codec import quantize_tables
from tests.golden.golden_io import case_cia, lbl_case_paths, load_case, load_cia
from tests.tolerance import (TOL_JAC_FP64, TOL_JAC_MIXED, TOL_RAD_FP64, TOL_RAD_MIXED,
                             fp32_range, fp32_range_jac, jac_err, rad_err)

pytestmark = pytest.mark.gpu
LBL = lbl_case_paths()
FORMS = {0: "fp64", 1: "mixed", 2: "bf16", 3: "f16"}


@pytest.fixture
def lbl(gpu_lib):
    yield gpu_lib
    gpu_lib.set_ilbl(0)


def _gas(d, ngas):
    keep = np.ones(d.shape[-1], bool)
    keep[ngas] = False
    return d[..., keep]


def _tol(mode):
    return (TOL_RAD_FP64, TOL_JAC_FP64) if mode == 0 else (TOL_RAD_MIXED, TOL_JAC_MIXED)


@pytest.mark.parametrize("mode", [0, 1, 2, 3], ids=lambda m: FORMS[m])
@pytest.mark.parametrize("path", LBL, ids=[os.path.basename(p)[4:-4] for p in LBL])
def test_lbl_golden_case(lbl, path, mode):
    from oracle import oracle
    tables, req, outs, note = load_case(path)
    print(f"lbl case {os.path.basename(path)} mode {FORMS[mode]}", flush=True)
    lbl.register_lbltables(tables)
    lbl.set_ilbl(2)
    o, g, d = lbl.spectrum(req, mode=mode)
    tr, tj = _tol(mode)
    ng = req.ngas
    if mode >= 2:  # against the oracle on the tables the form decodes
        fmt = "bf16" if mode == 2 else "f16"
        ro, rg, rd = oracle.spectrum(quantize_tables(tables, fmt), case_cia(path), req, ilbl=2)
        rt = rd[..., ng]
    else:
        ro, rg, rd = outs["ref64_output"], outs["ref64_gtsurf"], outs["ref64_doutputdq"]
        rt = oracle.spectrum(tables, case_cia(path), req, ilbl=2)[2][..., ng]
    dt = d[..., ng]
    if mode >= 1:  # FP32 stores hold no value below FP32's normal range
        ro, rg = fp32_range(ro), fp32_range(rg)
        rd = fp32_range_jac(rd, ng, gas_scale=1e-20)
        d = fp32_range_jac(d, ng, gas_scale=1e-20)
        rt, dt = fp32_range(rt), fp32_range(dt)
    er, eg = rad_err(o, ro), rad_err(g, rg)
    ej, et = jac_err(_gas(d, ng), _gas(rd, ng)), jac_err(dt[..., None], rt[..., None])
    print(f"  radiance {er:.2e} gtsurf {eg:.2e} gas jacobian {ej:.2e} T {et:.2e}", flush=True)
    assert er < tr and eg < tr, note
    assert ej < tj and et < tj, note


def _seeded(seed=61, nlayer=120, nkgas=5, nwave=64):
    tables = synth.make_lbltables(synth.KGAS_IDS[:nkgas], 600.0, 0.01, 5000, np_=20, nt=10, seed=seed)
    req = synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=seed, table_vmin=600.0,
                             table_delv=0.01)
    rng = np.random.default_rng(seed)
    req.vwave = np.sort(600.0 + rng.uniform(0.0, 49.98, nwave))
    return tables, req


def test_lbl_seeded_against_oracle(lbl):
    from oracle import oracle
    tables, req = _seeded()
    print("lbl seeded: 120 layers, 5 LBL gases + H2/He, 64 wavenumbers", flush=True)
    ro, rg, rd = oracle.spectrum(tables, load_cia(), req, ilbl=2, nthreads=8)
    lbl.register_lbltables(tables)
    lbl.set_ilbl(2)
    for mode in (0, 1):
        o, g, d = lbl.spectrum(req, mode=mode)
        tr, tj = _tol(mode)
        er, ej = rad_err(o, ro), jac_err(d, rd)
        print(f"  mode {mode}: radiance {er:.2e} jacobian {ej:.2e}", flush=True)
        assert er < tr and ej < tj


@pytest.mark.parametrize("recl", [4, 8])
def test_read_klbllist_equals_register(lbl, tmp_path, recl):
    path = [p for p in LBL if p.endswith("lbl_t2.npz")][0]
    tables, req, _, _ = load_case(path)
    names = []
    for i, t in enumerate(tables):
        n = str(tmp_path / f"g{i}")  # no extension: file.f appends .lta
        kta.write_lta(n + ".lta", t, recl)
        names.append(n)
    lst = tmp_path / "tables.lls"
    lst.write_text("\n".join(names) + "\n\n")
    print(f"read_klbllist recl {recl}", flush=True)
    # the tables as the file holds them (REAL(4) records round k)
    back = [kta.read_lta(n + ".lta", recl) for n in names]
    lbl.set_ilbl(2)
    lbl.register_lbltables(back)
    a = lbl.spectrum(req, mode=0)
    lbl.read_klbllist(str(lst), recl)
    b = lbl.spectrum(req, mode=0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lbl_batch_equals_alone(lbl):
    tables, big = _seeded(nlayer=90, nwave=20)
    small = synth.make_request(nlayer=33, nkgas=3, nwave=7, seed=5, table_vmin=600.0,
                               table_delv=0.37, wave_offset=3)
    lbl.register_lbltables(tables)
    lbl.set_ilbl(2)
    reqs = [big, small, big]
    print("lbl batch of 3 requests", flush=True)
    b = lbl.Batch(reqs, mode=1)
    b.launch()
    b.sync()
    res = [b.read(i, weighted=False) for i in range(len(reqs))]
    b.close()
    for q, r in zip(reqs, res):
        o, g, d = lbl.spectrum(q, mode=1)
        assert np.array_equal(o, r["output"]) and np.array_equal(d, r["doutputdq"])


def test_lbl_noalloc_equals_batch(lbl):
    import torch
    from tests.test_zz_no_alloc import _batch, _noalloc
    tables, req = _seeded(nlayer=60, nwave=24)
    lbl.register_lbltables(tables)
    lbl.set_ilbl(2)
    lbl.set_mode(1)
    try:
        print("lbl noalloc vs batch", flush=True)
        torch.cuda.set_device(0)
        for x, y in zip(_noalloc(req, 1), _batch(req, 1)):
            assert np.array_equal(x, y)
    finally:
        lbl.set_mode(0)


def test_ilbl_back_to_zero_keeps_ktables(lbl):
    tables, cia, req = synth.config1(nwave=6, nlayer=40, nkgas=3, ng=10)
    lbl.register_ktables(tables)
    a = lbl.spectrum(req, mode=0)
    lt, lreq = _seeded(nlayer=40, nwave=6)
    lbl.register_lbltables(lt)
    lbl.set_ilbl(2)
    assert lbl.get_ilbl() == 2
    lbl.spectrum(lreq, mode=0)
    lbl.set_ilbl(0)
    b = lbl.spectrum(req, mode=0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lbl_refusals(lbl):
    with pytest.raises(RadgError) as e:
        lbl.set_ilbl(1)
    assert e.value.code == 2
    tables, req = _seeded(nlayer=10, nwave=3)
    lbl.register_lbltables(tables)
    lbl.set_ilbl(2)
    req.vwave = np.array([600.0 - 0.0105])  # below slice 1's reach (x <= -1)
    with pytest.raises(RadgError) as e:
        lbl.spectrum(req, mode=0)
    assert e.value.code == 4
    req.vwave = np.array([600.0 - 0.0095])  # INT truncates toward zero: slice 1
    lbl.spectrum(req, mode=0)
