"""Host-side BN254 scalars for the few constants the host computes (domain roots,
divisors, coset factors). Bulk field work never runs here; it runs on the GPU.

Values are Python ints; `limbs`/`mont_limbs` produce the 32-byte boundary
format (4 little-endian u64 limbs, Montgomery form a * 2^256 mod m).
"""
from __future__ import annotations

import numpy as np

R_MOD = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001  # Fr
P_MOD = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47  # Fq
TWO_ADICITY = 28
# halo2curves bn256 Fr constants: multiplicative generator 7, ROOT_OF_UNITY = 7^((r-1)/2^28),
# ZETA (cube root of unity used as the coset generator g_coset)
GENERATOR = 7
ROOT_OF_UNITY = pow(GENERATOR, (R_MOD - 1) >> TWO_ADICITY, R_MOD)
ZETA = 0x30644E72E131A029048B6E193FD84104CC37A73FEC2BC5E9B8CA0B2D36636F23
DELTA = pow(GENERATOR, 1 << TWO_ADICITY, R_MOD)
MONT = 1 << 256


def limbs(x: int) -> np.ndarray:
    return np.array([(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def from_limbs(a) -> int:
    a = np.asarray(a, dtype=np.uint64).reshape(-1)
    return sum(int(w) << (64 * i) for i, w in enumerate(a[:4]))


def mont(x: int, m: int = R_MOD) -> int:
    return (x % m) * MONT % m


def unmont(x: int, m: int = R_MOD) -> int:
    return x * pow(MONT, -1, m) % m


def mont_limbs(x: int, m: int = R_MOD) -> np.ndarray:
    return limbs(mont(x, m))


def mont_bytes(x: int, m: int = R_MOD) -> bytes:
    return mont_limbs(x, m).tobytes()


def value_of(limb_row, m: int = R_MOD) -> int:
    """Montgomery limbs -> integer value."""
    return unmont(from_limbs(limb_row), m)


def omega(k: int) -> int:
    """Primitive 2^k-th root of unity (ROOT_OF_UNITY squared 28-k times)."""
    if not 0 <= k <= TWO_ADICITY:
        raise ValueError(f"k={k} outside [0, {TWO_ADICITY}]")
    return pow(ROOT_OF_UNITY, 1 << (TWO_ADICITY - k), R_MOD)


def inv(x: int, m: int = R_MOD) -> int:
    return pow(x, -1, m)
