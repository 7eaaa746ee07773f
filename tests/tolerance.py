"""Error measures shared by the parity tests.

* radiance (and gtsurf): relative error per value, ``|a - b| / |b|``, where
  ``b`` is the oracle; values the oracle gives as exactly 0 must come out
  within ``1e-30`` absolute.
* Jacobians: for each (wavenumber, path, parameter) column the error is
  ``max_layer |a - b| / max_layer |b|``.  Per-element relative error is
  meaningless here: single entries can cancel to near zero (an isothermal
  stretch has dR/dtau = 0 exactly), so each column is scaled by its own
  largest entry.
"""
import numpy as np

# FP64 form against the FP64 oracle
TOL_RAD_FP64 = 1e-5
TOL_JAC_FP64 = 1e-4
# mixed-precision form against the FP64 oracle
TOL_RAD_MIXED = 1e-4
TOL_JAC_MIXED = 1e-3


# FP32's smallest normal number.  An FP32 output (the reference's REAL(4)
# build, the mixed form) cannot hold an FP64 value below it, e.g. an opaque
# path's transmission exp(-500) (dpexp.f): such reference values are
# compared as 0.
FP32_TINY = 1.1754943508222875e-38


def fp32_range(b):
    b = np.asarray(b, dtype=np.float64)
    return np.where(np.abs(b) < FP32_TINY, 0.0, b)


def fp32_range_jac(d, ngas, gas_scale=1.0):
    """The Jacobian ``d[..., nparam]`` with entries an FP32 store cannot hold
    set to 0: below FP32_TINY, or FP32_TINY * ``gas_scale`` in the first
    ``ngas`` (gas) columns when the store holds them scaled (the mixed form
    keeps them per 1e20 molecules: gas_scale = 1e-20)."""
    d = np.asarray(d, dtype=np.float64)
    floor = np.full(d.shape[-1], FP32_TINY)
    floor[:ngas] *= gas_scale
    return np.where(np.abs(d) < floor, 0.0, d)


def rad_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nz = b != 0
    err = 0.0
    if nz.any():
        err = float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])))
    if (~nz).any():
        za = float(np.max(np.abs(a[~nz])))
        if za > 1e-30:
            err = max(err, np.inf)
    return err


def jac_err_cols(a, b, layer_axis=-2):
    """jac_err per column (every index but the layer axis)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b), axis=layer_axis)
    diff = np.max(np.abs(a - b), axis=layer_axis)
    ok0 = scale == 0
    return np.where(ok0, np.where(diff > 1e-300, np.inf, 0.0), diff / np.where(ok0, 1.0, scale))


def jac_err(a, b, layer_axis=-2):
    return float(np.max(jac_err_cols(a, b, layer_axis)))
