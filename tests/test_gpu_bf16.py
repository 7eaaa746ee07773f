"""GPU parity of the BF16 log10(k) table form.

The BF16 form reads k-tables the library encodes on the device (log10 k,
one offset and scale per table, rounded to BF16).  Handing the FP64 oracle
tables that hold exactly the decoded values (tests/bf16_codec.py) isolates
the kernel: the BF16 form must then match it to the mixed-form tolerance.
The storage error itself -- the BF16 form against the oracle on the
original FP64 tables -- is reported; its radiance part is bounded here by
2e-3, about 3x the largest value measured on these shapes (6e-4).
Jacobian columns carry larger storage error (up to ~0.2 of a column's
largest entry for weak absorbers, since dk/dT comes from differences of
rounded table nodes) and are reported only.
"""
import numpy as np
import pytest

from radtrancode_amd import synth
from tests.bf16_codec import encode, quantize_tables, table_scaling
from tests.golden.golden_io import case_paths, load_case, load_cia
from tests.tolerance import TOL_JAC_MIXED, TOL_RAD_MIXED, jac_err, rad_err

pytestmark = pytest.mark.gpu
TOL_RAD_BF16_STORAGE = 2e-3


@pytest.mark.parametrize("shape", [(120, 3, 20, 24), (200, 8, 20, 8), (40, 1, 20, 12)],
                         ids=lambda s: "L%d_k%d_g%d_w%d" % s)
def test_bf16_against_oracle(gpu_lib, shape):
    from oracle import oracle
    L, nk, ng, nw = shape
    print(f"bf16 shape {shape}", flush=True)
    tables, _, req = synth.config1(nwave=nw, nlayer=L, nkgas=nk, ng=ng)
    cia = load_cia()
    gpu_lib.register_ktables(tables)
    o, g, d = gpu_lib.spectrum(req, mode=gpu_lib.MODE_BF16)
    oq, gq, dq = oracle.spectrum(quantize_tables(tables), cia, req, nthreads=8)
    er, eg, ej = rad_err(o, oq), rad_err(g, gq), jac_err(d, dq)
    print(f"  vs oracle on decoded tables: radiance {er:.2e} gtsurf {eg:.2e} jacobian {ej:.2e}",
          flush=True)
    assert er < TOL_RAD_MIXED and eg < TOL_RAD_MIXED and ej < TOL_JAC_MIXED
    o0, g0, d0 = oracle.spectrum(tables, cia, req, nthreads=8)
    sr, sj = rad_err(o, o0), jac_err(d, d0)
    print(f"  BF16 storage error vs FP64 oracle: radiance {sr:.2e} jacobian {sj:.2e}", flush=True)
    assert sr < TOL_RAD_BF16_STORAGE


@pytest.mark.parametrize("name", ["zeros", "clamps", "six_gases"])
def test_bf16_golden_edge_cases(gpu_lib, name):
    """Zero slices, zero corners (linear interpolation) and P/T clamps."""
    from oracle import oracle
    path = [p for p in case_paths() if f"case_{name}." in p][0]
    tables, req, _, _ = load_case(path)
    print(f"bf16 golden {name}", flush=True)
    gpu_lib.register_ktables(tables)
    o, g, d = gpu_lib.spectrum(req, mode=gpu_lib.MODE_BF16)
    oq, gq, dq = oracle.spectrum(quantize_tables(tables), load_cia(), req, nthreads=8)
    er, eg, ej = rad_err(o, oq), rad_err(g, gq), jac_err(d, dq)
    print(f"  radiance {er:.2e} gtsurf {eg:.2e} jacobian {ej:.2e}", flush=True)
    assert er < TOL_RAD_MIXED and eg < TOL_RAD_MIXED and ej < TOL_JAC_MIXED


@pytest.mark.parametrize("name", ["six_gases", "zeros", "medium"])
def test_bf16_codes_match_host_codec(gpu_lib, name):
    """The device encoding equals the host restatement code for code."""
    path = [p for p in case_paths() if f"case_{name}." in p][0]
    tables, _, _, _ = load_case(path)
    print(f"bf16 codes {name}", flush=True)
    gpu_lib.register_ktables(tables)
    for it, t in enumerate(tables):
        codes, off, sc = gpu_lib.ktable_bf16(it, t.k.shape)
        off_h, sc_h = table_scaling(t.k)
        assert abs(off - off_h) <= 1e-12 * max(1.0, abs(off_h)) and abs(sc - sc_h) <= 1e-12 * sc_h
        ref = encode(t.k, off_h, sc_h)
        nbad = int(np.count_nonzero(codes != ref))
        if nbad:
            i = np.argwhere(codes != ref)[0]
            print(f"  table {it}: {nbad} codes differ, first at {tuple(i)}: device "
                  f"{codes[tuple(i)]:#06x} host {ref[tuple(i)]:#06x} k {t.k[tuple(i)]!r}", flush=True)
        assert nbad == 0
