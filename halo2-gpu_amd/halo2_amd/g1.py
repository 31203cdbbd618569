"""Host-side G1 helpers for the handful of points the host touches (folding
chunked MSM results, decoding the 96-byte MSM output). Bulk curve work runs on
the GPU.

Affine points are (x, y) integer pairs or None for the identity.
"""
from __future__ import annotations

import numpy as np

from .fields import P_MOD, from_limbs, mont_limbs, unmont


def add(a, b):
    """Affine addition on y^2 = x^3 + 3 (handles doubling and inverses)."""
    if a is None:
        return b
    if b is None:
        return a
    x1, y1 = a
    x2, y2 = b
    if x1 == x2:
        if (y1 + y2) % P_MOD == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P_MOD) % P_MOD
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P_MOD) % P_MOD
    x3 = (lam * lam - x1 - x2) % P_MOD
    return x3, (lam * (x1 - x3) - y1) % P_MOD


def from_out12(out) -> tuple | None:
    """Decode the MSM projective-out layout (x, y, 1) / (0, 1, 0) to an affine pair."""
    out = np.asarray(out, dtype=np.uint64).reshape(12)
    z = from_limbs(out[8:12])
    if z == 0:
        return None
    return unmont(from_limbs(out[0:4]), P_MOD), unmont(from_limbs(out[4:8]), P_MOD)


def to_out12(pt) -> np.ndarray:
    out = np.zeros(12, dtype=np.uint64)
    out[4:8] = mont_limbs(1, P_MOD)
    if pt is not None:
        out[0:4] = mont_limbs(pt[0], P_MOD)
        out[4:8] = mont_limbs(pt[1], P_MOD)
        out[8:12] = mont_limbs(1, P_MOD)
    return out
