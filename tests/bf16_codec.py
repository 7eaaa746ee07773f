"""Host restatement of the BF16 log10(k) table codec, for the tests.

The library makes the BF16 table form on the device (radg_kernels.hip,
to_bf16_kernel) from the FP64 tables it was given; this module computes the
same codes in numpy so a test can hand the FP64 oracle the tables the BF16
form actually reads:

  per table:  off = (log10 kmin + log10 kmax) / 2,  sc = max((log10 kmax - log10 kmin) / 2, 1e-3)
              over the table's positive values
  per value:  code = bf16_rne(float32((log10 k - off) / sc)),  k <= 0 -> 0xFF80 (-inf)
  decode:     k = 10 ** (float32(code) * float32(sc) + float32(off))   (FP32 arithmetic)
"""
import numpy as np

NONPOS = 0xFF80


def table_scaling(k):
    k = np.asarray(k, dtype=np.float64)
    pos = k[k > 0]
    if pos.size == 0:
        return 0.0, 1.0
    a, z = np.log10(pos.min()), np.log10(pos.max())
    return 0.5 * (a + z), max(0.5 * (z - a), 1e-3)


def bf16_rne(x):
    """float32 array -> uint16 bf16 codes, round to nearest, ties to even."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = b + 0x7FFF + ((b >> 16) & 1)
    return (b >> 16).astype(np.uint16)


def bf16_value(code):
    """uint16 bf16 codes -> float32 values."""
    return (np.asarray(code, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def encode(k, off, sc):
    k = np.asarray(k, dtype=np.float64)
    pos = k > 0
    x = np.zeros(k.shape, dtype=np.float32)
    x[pos] = ((np.log10(k[pos]) - off) / sc).astype(np.float32)
    code = bf16_rne(x)
    code[~pos] = NONPOS
    return code


def decode(code, off, sc):
    code = np.asarray(code, dtype=np.uint16)
    l10 = bf16_value(code) * np.float32(sc) + np.float32(off)
    out = np.power(10.0, l10.astype(np.float64))
    out[code == NONPOS] = 0.0
    return out


def quantize_tables(tables):
    """Copies of the k-tables holding the values the BF16 form decodes."""
    import copy

    out = []
    for t in tables:
        off, sc = table_scaling(t.k)
        t2 = copy.copy(t)
        t2.k = decode(encode(t.k, off, sc), off, sc)
        out.append(t2)
    return out
