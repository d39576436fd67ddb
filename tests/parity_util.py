"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np


def same_bits(a, b) -> bool:
    """Bitwise equality of float64 arrays, NaNs compared by position only
    (the sign/payload of a NaN is not specified by IEEE arithmetic)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.int64), b[~nb].view(np.int64))


def describe_diff(a, b) -> str:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    m = ~(na | nb)
    d = np.abs(a[m] - b[m])
    nbad = int(np.count_nonzero(a[m].view(np.int64) != b[m].view(np.int64)))
    return (f"shape={a.shape} nan_mismatch={int(np.count_nonzero(na != nb))} "
            f"bits_differ={nbad} maxabs={d.max() if d.size else 0:.3e}")


def rel_l1(a, b, var):
    """relative L1 distance of conserved variable var over the grid."""
    a = np.asarray(a)[..., var]
    b = np.asarray(b)[..., var]
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / (den if den > 0 else 1.0))


FVTD_KEYS = ("Ut", "FX", "FY", "S", "Prod", "Dest")


def fvtd_view(out, key):
    """The parts of an FVTD dump that the PATANKAR build defines: Ut(1) and
    S(1) are unused by the reference with PATANKAR."""
    a = out[key]
    if key in ("Ut", "S"):
        return a[..., 1:]
    return a
