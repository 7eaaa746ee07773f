"""The host restatement of the 16-bit log10(k) codecs, pinned by brute force.

bf16_rne and f16_rne must pick the nearest of all 65536 values (ties to the
even code); the offsets are the plane means a plain loop gives; a round trip
must keep log10(k) within |log10(k) - off| * 2^-8 (BF16: unit roundoff of an
8-bit significand) or 2^-11 (half), plus FP32 rounding.
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


def test_f16_rne_matches_brute_force():
    codes = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    with np.errstate(invalid="ignore"):
        vals = bc.f16_value(codes).astype(np.float64)
    ok = np.isfinite(vals)
    sv, sc = vals[ok], codes[ok]
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-8, 8, 3000), rng.uniform(-1e-3, 1e-3, 300),
                        [1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11]]).astype(np.float32)
    got = bc.f16_rne(x)
    for xi, gi in zip(x.astype(np.float64), got):
        d = np.abs(sv - xi)
        best = d.min()
        cand = sc[d == best]
        if best == 0:
            assert bc.f16_value(gi) == xi
            continue
        if len(cand) > 1:
            cand = cand[(cand & 1) == 0]
        assert gi in set(cand.tolist()), (xi, gi, cand)


def test_offsets_are_plane_means():
    """Offsets by a plain loop: FP64 sum of log10 k over the positive values of
    each (point, ordinate) plane in pool order, divided, rounded to FP32."""
    rng = np.random.default_rng(8)
    k = 10.0 ** rng.normal(-22, 2, size=(3, 4, 5, 6))
    k[1, :, :, 2] = 0.0
    k[2, 0, 0, 3] = -1.0
    off = bc.slice_offsets(k)
    for p in range(3):
        for g in range(6):
            s, n = 0.0, 0
            for ip in range(4):
                for it in range(5):
                    if k[p, ip, it, g] > 0:
                        s += float(np.log10(k[p, ip, it, g]))
                        n += 1
            assert off[p, g] == np.float32(s / n if n else 0.0)


@pytest.mark.parametrize("fmt,bits", [("bf16", 8), ("f16", 11)])
def test_round_trip_bound(fmt, bits):
    rng = np.random.default_rng(7)
    k = 10.0 ** rng.normal(-22, 2, size=(4, 5, 6, 8))
    k[0, 0, 0, :2] = [0.0, -1.0]
    off = bc.slice_offsets(k)
    code = bc.encode(k, off, fmt)
    assert code[0, 0, 0, 0] == bc.NONPOS[fmt] and code[0, 0, 0, 1] == bc.NONPOS[fmt]
    kq = bc.decode(code, off, fmt)
    assert kq[0, 0, 0, 0] == 0.0 and kq[0, 0, 0, 1] == 0.0
    pos = k > 0
    o = off.astype(np.float64)[:, None, None, :] * np.ones_like(k)
    err = np.abs(np.log10(kq[pos]) - np.log10(k[pos]))
    dev = np.abs(np.log10(k[pos]) - o[pos])
    # one 16-bit rounding of log10 k - off, an FP32 rounding of it and of the
    # decoded sum
    bound = dev * 2.0 ** -bits + 4 * (dev + np.abs(o[pos])) * 2.0 ** -24
    assert np.all(err <= bound), float(np.max(err - bound))


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_quantize_tables_keeps_zero_slices(fmt):
    from radtrancode_amd import synth

    tables, _, _ = synth.config1(nwave=4, nlayer=10, nkgas=1, ng=8)
    t = tables[0]
    t.k[1] = 0.0
    q = bc.quantize_tables([t], fmt)[0]
    assert np.all(q.k[1] == 0.0)
    assert np.all((q.k > 0) == (t.k > 0))
