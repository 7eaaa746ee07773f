"""The host restatement of the BF16 log10(k) codec, pinned by brute force.

bf16_rne must pick the nearest of all 65536 BF16 values (ties to the even
code), and a round trip must keep log10(k) within |log10(k) - off| * 2^-8
(unit roundoff of an 8-bit significand), plus FP32 rounding.
"""
import numpy as np
import pytest

from tests import bf16_codec as bc


def _all_finite_bf16():
    codes = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    with np.errstate(invalid="ignore"):
        vals = bc.bf16_value(codes).astype(np.float64)
    ok = np.isfinite(vals)
    return codes[ok], vals[ok]


def test_rne_matches_brute_force():
    codes, vals = _all_finite_bf16()
    order = np.argsort(vals, kind="stable")
    sv, sc = vals[order], codes[order]
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-1, 1, 4000), rng.uniform(-1e-3, 1e-3, 500),
                        [0.0, 1.0, -1.0, 0.5, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8]]).astype(np.float32)
    got = bc.bf16_rne(x)
    for xi, gi in zip(x.astype(np.float64), got):
        d = np.abs(sv - xi)
        best = d.min()
        cand = sc[d == best]
        # +0/-0 both represent 0: accept either; exact ties go to the even code
        if best == 0:
            assert bc.bf16_value(gi) == xi
            continue
        if len(cand) > 1:
            cand = cand[(cand & 1) == 0]
        assert gi in set(cand.tolist()), (xi, gi, cand)


def test_ties_to_even():
    # 1 + 2^-8 lies halfway between 1 (code 0x3F80) and 1 + 2^-7 (0x3F81): even wins
    assert bc.bf16_rne(np.float32(1 + 2.0 ** -8)) == 0x3F80
    # 1 + 3*2^-8 lies halfway between 0x3F81 and 0x3F82: even wins
    assert bc.bf16_rne(np.float32(1 + 3 * 2.0 ** -8)) == 0x3F82


@pytest.mark.parametrize("lo,hi", [(-11.0, 7.8), (-3.0, -2.5), (2.0, 2.0)])
def test_round_trip_bound(lo, hi):
    rng = np.random.default_rng(7)
    k = 10.0 ** rng.uniform(lo, hi, 5000)
    k[:3] = [0.0, -1.0, 10.0 ** lo]
    off, sc = bc.table_scaling(k)
    code = bc.encode(k, off, sc)
    assert code[0] == bc.NONPOS and code[1] == bc.NONPOS
    kq = bc.decode(code, off, sc)
    assert kq[0] == 0.0 and kq[1] == 0.0
    pos = k > 0
    err = np.abs(np.log10(kq[pos]) - np.log10(k[pos]))
    bound = np.abs(np.log10(k[pos]) - off) * 2.0 ** -8 + 4 * abs(off) * 2.0 ** -24 + 4 * sc * 2.0 ** -24
    assert np.all(err <= bound), float(np.max(err - bound))
    # the codes span [-1, 1] over the positive range
    x = bc.bf16_value(code[pos]).astype(np.float64)
    assert x.min() >= -1.0 - 2 ** -8 and x.max() <= 1.0 + 2 ** -8


def test_quantize_tables_keeps_zero_slices():
    from radtrancode_amd import synth

    tables, _, _ = synth.config1(nwave=4, nlayer=10, nkgas=1, ng=8)
    t = tables[0]
    t.k[1] = 0.0
    q = bc.quantize_tables([t])[0]
    assert np.all(q.k[1] == 0.0)
    assert np.all((q.k > 0) == (t.k > 0))
