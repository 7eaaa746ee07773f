"""Host restatement of the 16-bit log10(k) table codecs, for the tests.

The library makes the 16-bit table forms on the device (radg_kernels.hip,
slice_offsets_kernel and to_code16_kernel) from the FP64 tables it was
given; this module computes the same offsets and codes in numpy so a test
can hand the FP64 oracle the tables a 16-bit form actually reads:

  per (wavenumber point, g-ordinate) of a table [npoint, np, nt, ng]:
      off = float32(sum of log10 k over the positive k of the (P, T) plane,
                    taken in pool order, / their number)      (0 if none)
  per value:  code = rne16(float32(log10 k - off)),  k <= 0 -> -inf
              rne16: BF16 ("bf16") or IEEE half ("f16"), round to nearest even
  decode:     k = 10 ** (float32(code) + off)   (FP32 arithmetic)
"""
import numpy as np

NONPOS = {"bf16": 0xFF80, "f16": 0xFC00}


def slice_offsets(k):
    """[npoint, ng] float32 offsets of a table k[npoint, np, nt, ng]."""
    k = np.asarray(k, dtype=np.float64)
    npoint, ng = k.shape[0], k.shape[-1]
    flat = k.reshape(npoint, -1, ng)
    pos = flat > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        lg = np.where(pos, np.log10(np.where(pos, flat, 1.0)), 0.0)
    s = np.cumsum(lg, axis=1)[:, -1, :]   # sequential sum, as the device loop
    n = pos.sum(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        off = np.where(n > 0, s / np.where(n > 0, n, 1), 0.0)
    return off.astype(np.float32)


def bf16_rne(x):
    """float32 array -> uint16 bf16 codes, round to nearest, ties to even."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = b + 0x7FFF + ((b >> 16) & 1)
    return (b >> 16).astype(np.uint16)


def bf16_value(code):
    """uint16 bf16 codes -> float32 values."""
    return (np.asarray(code, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def f16_rne(x):
    return np.asarray(x, dtype=np.float32).astype(np.float16).view(np.uint16)


def f16_value(code):
    return np.asarray(code, dtype=np.uint16).view(np.float16).astype(np.float32)


def encode(k, off, fmt="bf16"):
    """Codes of table k[npoint, ..., ng] with offsets off[npoint, ng]."""
    k = np.asarray(k, dtype=np.float64)
    o = np.asarray(off, dtype=np.float32).astype(np.float64)
    o = o.reshape((k.shape[0],) + (1,) * (k.ndim - 2) + (k.shape[-1],))
    pos = k > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        x = (np.log10(np.where(pos, k, 1.0)) - o).astype(np.float32)
    code = bf16_rne(x) if fmt == "bf16" else f16_rne(x)
    code[~pos] = NONPOS[fmt]
    return code


def decode(code, off, fmt="bf16"):
    code = np.asarray(code, dtype=np.uint16)
    o = np.asarray(off, dtype=np.float32)
    o = o.reshape((code.shape[0],) + (1,) * (code.ndim - 2) + (code.shape[-1],))
    v = bf16_value(code) if fmt == "bf16" else f16_value(code)
    with np.errstate(invalid="ignore"):
        l10 = v + o
    out = np.power(10.0, l10.astype(np.float64))
    out[code == NONPOS[fmt]] = 0.0
    return out


def quantize_tables(tables, fmt="bf16"):
    """Copies of the k-tables holding the values a 16-bit form decodes."""
    import copy

    out = []
    for t in tables:
        off = slice_offsets(t.k)
        t2 = copy.copy(t)
        t2.k = decode(encode(t.k, off, fmt), off, fmt)
        out.append(t2)
    return out
