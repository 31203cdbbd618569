"""Pure-Python big-integer oracle for tiny inputs (test infrastructure only).

Everything here is written from the mathematics, independently of the C
oracle and of the GPU code, so that tests can pin both:

* BN254 parameters from the curve seed u (p, r, the trace), the Montgomery
  constants of ``cuda/include/field/alt_bn128.hpp:19-47`` recomputed as powers of 2;
* the 2^28-th root of unity of halo2curves (generator 7) and the cube root of
  unity zeta hard-coded at ``cuda/include/kernel/zeta.h:19-30``;
* affine G1 arithmetic on y^2 = x^3 + 3 with G = (1, 2);
* the naive O(n^2) DFT and the closed forms of the halo2 NTT family
  (``src/poly/domain.rs:730-746, 650-691``, ``src/cpu/evaluator.rs:26-45``).
"""
from __future__ import annotations

import numpy as np

# BN254 curve seed (Barreto-Naehrig parametrisation)
BN_U = 4965661367192848881
P = 36 * BN_U**4 + 36 * BN_U**3 + 24 * BN_U**2 + 6 * BN_U + 1
R = 36 * BN_U**4 + 36 * BN_U**3 + 18 * BN_U**2 + 6 * BN_U + 1
MONT = 1 << 256
S = 28  # 2-adicity of r - 1
GENERATOR = 7  # multiplicative generator used by halo2curves for Fr

# zeta.h:19-30 (Montgomery limbs of zeta and zeta^2 in Fr)
ZETA_1_MONT_LIMBS = (244305545194690131, 8351807910065594880, 14266533074055306532, 404339206190769364)
ZETA_2_MONT_LIMBS = (10657714497315350963, 9029678389775483239, 10080386412464207114, 2070906320917503013)


def limbs_to_int(limbs) -> int:
    v = 0
    for i, w in enumerate(limbs):
        v |= int(w) << (64 * i)
    return v


def int_to_limbs(x: int) -> list[int]:
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def to_mont(x: int, m: int = R) -> int:
    return (x * MONT) % m


def from_mont(x: int, m: int = R) -> int:
    return (x * pow(MONT, -1, m)) % m


def root_of_unity() -> int:
    """halo2curves Fr::ROOT_OF_UNITY = 7^((r-1)/2^28)."""
    return pow(GENERATOR, (R - 1) >> S, R)


def omega(k: int) -> int:
    return pow(root_of_unity(), 1 << (S - k), R)


def zeta() -> int:
    return from_mont(limbs_to_int(ZETA_1_MONT_LIMBS), R)


# ---------------------------------------------------------------- arrays

def fr_array(values, mont: bool = True) -> np.ndarray:
    """Python ints -> (n, 4) uint64 limbs (Montgomery form by default)."""
    out = np.zeros((len(values), 4), dtype=np.uint64)
    for i, v in enumerate(values):
        x = to_mont(v % R, R) if mont else v
        out[i] = int_to_limbs(x)
    return out


def fr_values(arr: np.ndarray, mont: bool = True, m: int = R) -> list[int]:
    arr = np.asarray(arr, dtype=np.uint64).reshape(-1, 4)
    vals = [limbs_to_int(row) for row in arr]
    return [from_mont(v, m) for v in vals] if mont else vals


def fq_array(values) -> np.ndarray:
    out = np.zeros((len(values), 4), dtype=np.uint64)
    for i, v in enumerate(values):
        out[i] = int_to_limbs(to_mont(v % P, P))
    return out


# -------------------------------------------------------------------- G1
# affine points as (x, y) ints; None = identity

G1_GEN = (1, 2)


def g1_on_curve(pt) -> bool:
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - 3) % P == 0


def g1_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    x1, y1 = a
    x2, y2 = b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    y3 = (lam * (x1 - x3) - y1) % P
    return (x3, y3)


def g1_mul(pt, s: int):
    s %= R
    acc = None
    add = pt
    while s:
        if s & 1:
            acc = g1_add(acc, add)
        add = g1_add(add, add)
        s >>= 1
    return acc


def g1_affine_array(points) -> np.ndarray:
    """affine points -> (n, 8) uint64 Montgomery limbs, identity = zeros."""
    out = np.zeros((len(points), 8), dtype=np.uint64)
    for i, pt in enumerate(points):
        if pt is not None:
            out[i, :4] = int_to_limbs(to_mont(pt[0], P))
            out[i, 4:] = int_to_limbs(to_mont(pt[1], P))
    return out


def g1_out12(pt) -> np.ndarray:
    """The 96-byte projective-out layout of the reference MSM."""
    out = np.zeros(12, dtype=np.uint64)
    one = int_to_limbs(to_mont(1, P))
    if pt is None:
        out[4:8] = one
    else:
        out[:4] = int_to_limbs(to_mont(pt[0], P))
        out[4:8] = int_to_limbs(to_mont(pt[1], P))
        out[8:] = one
    return out


def msm_naive(scalars: list[int], points) -> object:
    acc = None
    for s, pt in zip(scalars, points):
        acc = g1_add(acc, g1_mul(pt, s))
    return acc


# ------------------------------------------------------------------- NTT

def dft(a: list[int], w: int) -> list[int]:
    n = len(a)
    return [sum(a[i] * pow(w, i * j, R) for i in range(n)) % R for j in range(n)]


def idft(a: list[int], w: int) -> list[int]:
    n = len(a)
    winv = pow(w, -1, R)
    ninv = pow(n, -1, R)
    return [(v * ninv) % R for v in dft(a, winv)]


def coset_part(a: list[int], k: int, c: int) -> list[int]:
    """CosetFFT_Part: DFT_n(a_i c^i) with the size-n root."""
    tw = [(v * pow(c, i, R)) % R for i, v in enumerate(a)]
    return dft(tw, omega(k))


def coeff_to_extended(a: list[int], k: int, ext_k: int) -> list[int]:
    """pad to 2^ext_k, multiply index i by zeta^(i mod 3), DFT with the extended root."""
    z = zeta()
    n_ext = 1 << ext_k
    padded = list(a) + [0] * (n_ext - len(a))
    tw = [(v * pow(z, i % 3, R)) % R for i, v in enumerate(padded)]
    return dft(tw, omega(ext_k))


def extended_to_coeff(e: list[int], ext_k: int, target_len: int) -> list[int]:
    """iDFT with the extended root, multiply index i by [1, zeta^-1, zeta^-2][i mod 3], truncate."""
    z = zeta()
    zinv = pow(z, -1, R)
    c = idft(e, omega(ext_k))
    c = [(v * pow(zinv, i % 3, R)) % R for i, v in enumerate(c)]
    return c[:target_len]
