"""GPU parity of the 16-bit log10(k) table forms (BF16 and IEEE-half codes).

A 16-bit form reads k-tables the library encodes on the device: log10 k
minus one FP32 offset per (wavenumber point, g-ordinate) -- the mean of
log10 k over that point's (P, T) values -- rounded to 16 bits.  Handing the
FP64 oracle tables that hold exactly the decoded values (tests/bf16_codec.py)
isolates the kernel: the form must then match it to the mixed-form
tolerance.  The storage error itself -- the form against the oracle on the
original FP64 tables -- is bounded by the rounding of log10 k: the half
form's 11-bit significand keeps radiance within 1e-4; BF16's 8 bits keep it
within 2e-3 on these tables, whose log10 k are independent normal draws
spanning some 14 decades at each point.
"""
import numpy as np
import pytest

from radtrancode_amd import synth
from tests.bf16_codec import encode, quantize_tables, slice_offsets
from tests.golden.golden_io import case_paths, load_case, load_cia
from tests.tolerance import TOL_JAC_MIXED, TOL_RAD_MIXED, jac_err, rad_err

pytestmark = pytest.mark.gpu
TOL_RAD_STORAGE = {"bf16": 2e-3, "f16": 1e-4}
FORMS = [("bf16", 2), ("f16", 3)]


@pytest.mark.parametrize("fmt,mode", FORMS, ids=[f for f, _ in FORMS])
@pytest.mark.parametrize("shape", [(120, 3, 20, 24, 20, 10), (200, 8, 20, 8, 30, 15),
                                   (40, 1, 20, 12, 20, 10)],
                         ids=lambda s: "L%d_k%d_g%d_w%d_P%dxT%d" % s)
def test_code16_against_oracle(gpu_lib, shape, fmt, mode):
    from oracle import oracle
    L, nk, ng, nw, np_, nt = shape
    print(f"{fmt} shape {shape}", flush=True)
    tables, _, req = synth.config1(nwave=nw, nlayer=L, nkgas=nk, ng=ng, np_=np_, nt=nt)
    cia = load_cia()
    gpu_lib.register_ktables(tables)
    o, g, d = gpu_lib.spectrum(req, mode=mode)
    oq, gq, dq = oracle.spectrum(quantize_tables(tables, fmt), cia, req, nthreads=8)
    er, eg, ej = rad_err(o, oq), rad_err(g, gq), jac_err(d, dq)
    print(f"  vs oracle on decoded tables: radiance {er:.2e} gtsurf {eg:.2e} jacobian {ej:.2e}",
          flush=True)
    assert er < TOL_RAD_MIXED and eg < TOL_RAD_MIXED and ej < TOL_JAC_MIXED
    o0, g0, d0 = oracle.spectrum(tables, cia, req, nthreads=8)
    sr, sj = rad_err(o, o0), jac_err(d, d0)
    print(f"  {fmt} storage error vs FP64 oracle: radiance {sr:.2e} jacobian {sj:.2e}", flush=True)
    assert sr < TOL_RAD_STORAGE[fmt]


@pytest.mark.parametrize("fmt,mode", FORMS, ids=[f for f, _ in FORMS])
@pytest.mark.parametrize("name", ["zeros", "clamps", "six_gases", "noncoincident"])
def test_code16_golden_edge_cases(gpu_lib, name, fmt, mode):
    """Zero slices, zero corners (linear interpolation), P/T clamps and
    wavenumbers between table points (two slices, two offset sets)."""
    from oracle import oracle
    path = [p for p in case_paths() if f"case_{name}." in p][0]
    tables, req, _, _ = load_case(path)
    print(f"{fmt} golden {name}", flush=True)
    gpu_lib.register_ktables(tables)
    o, g, d = gpu_lib.spectrum(req, mode=mode)
    oq, gq, dq = oracle.spectrum(quantize_tables(tables, fmt), load_cia(), req, nthreads=8)
    er, eg, ej = rad_err(o, oq), rad_err(g, gq), jac_err(d, dq)
    print(f"  radiance {er:.2e} gtsurf {eg:.2e} jacobian {ej:.2e}", flush=True)
    assert er < TOL_RAD_MIXED and eg < TOL_RAD_MIXED and ej < TOL_JAC_MIXED


@pytest.mark.parametrize("fmt,mode", FORMS, ids=[f for f, _ in FORMS])
@pytest.mark.parametrize("name", ["six_gases", "zeros", "medium"])
def test_code16_matches_host_codec(gpu_lib, name, fmt, mode):
    """The device offsets equal the host restatement's, and so do the codes,
    but for isolated codes one unit apart: log10 k on the device and in numpy
    may differ in their last bit, which moves a value lying on a rounding tie
    of the FP32 -> 16-bit step to the neighbouring code."""
    path = [p for p in case_paths() if f"case_{name}." in p][0]
    tables, _, _, _ = load_case(path)
    print(f"{fmt} codes {name}", flush=True)
    gpu_lib.register_ktables(tables)
    for it, t in enumerate(tables):
        codes, offs = gpu_lib.ktable_codes(mode, it, t.k.shape)
        off_h = slice_offsets(t.k)
        assert np.array_equal(offs, off_h), np.max(np.abs(offs - off_h))
        ref = encode(t.k, off_h, fmt)
        diff = codes != ref
        nbad = int(np.count_nonzero(diff))
        if nbad:
            i = np.argwhere(diff)[0]
            print(f"  table {it}: {nbad} codes differ, first at {tuple(i)}: device "
                  f"{codes[tuple(i)]:#06x} host {ref[tuple(i)]:#06x} k {t.k[tuple(i)]!r}", flush=True)
        step = np.abs(codes[diff].astype(np.int64) - ref[diff].astype(np.int64))
        assert np.all(step == 1) and nbad <= max(2, 1e-4 * codes.size)
