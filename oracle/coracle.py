"""ctypes binding of the C oracle ``libhalo2_oracle.so`` (test infrastructure only).

Arrays are numpy uint64: Fr/Fq as (n, 4) Montgomery limbs, affine G1 as (n, 8),
the MSM projective-out point as (12,). Every function copies into fresh output
arrays; inputs are never modified unless the name says ``_inplace``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhalo2_oracle.so")
_lib = None

U64P = ctypes.POINTER(ctypes.c_uint64)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U64P)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _ptr_array(arrs):
    arr = (U64P * len(arrs))(*[_p(a) for a in arrs])
    return arr


def num_threads() -> int:
    return int(lib().or_num_threads())


def field_mul(fld: int, a, b) -> np.ndarray:
    a, b = _c(a).reshape(-1, 4), _c(b).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().or_field_mul(fld, _p(out), _p(a), _p(b), ctypes.c_size_t(len(a)))
    return out


def field_add(fld: int, a, b) -> np.ndarray:
    a, b = _c(a).reshape(-1, 4), _c(b).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().or_field_add(fld, _p(out), _p(a), _p(b), ctypes.c_size_t(len(a)))
    return out


def field_sub(fld: int, a, b) -> np.ndarray:
    a, b = _c(a).reshape(-1, 4), _c(b).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().or_field_sub(fld, _p(out), _p(a), _p(b), ctypes.c_size_t(len(a)))
    return out


def field_inv(fld: int, a) -> np.ndarray:
    a = _c(a).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().or_field_inv(fld, _p(out), _p(a), ctypes.c_size_t(len(a)))
    return out


def fr_pow(base, e: int) -> np.ndarray:
    base = _c(base).reshape(4)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_fr_pow(_p(out), _p(base), ctypes.c_uint64(e))
    return out


def fr_from_u512(raw: bytes) -> np.ndarray:
    n = len(raw) // 64
    out = np.zeros((n, 4), dtype=np.uint64)
    buf = ctypes.create_string_buffer(raw, len(raw))
    lib().or_fr_from_u512(_p(out), buf, ctypes.c_size_t(n))
    return out


def chacha_block(key_words, counter: int) -> bytes:
    key = (ctypes.c_uint32 * 8)(*key_words)
    out = ctypes.create_string_buffer(64)
    lib().or_chacha_block(key, ctypes.c_uint64(counter), out)
    return out.raw


def chacha_seed_key(seed: int) -> list[int]:
    key = (ctypes.c_uint32 * 8)()
    lib().or_chacha_seed_key(ctypes.c_uint64(seed), key)
    return list(key)


def chacha_fr(seed: int, start: int, n: int) -> np.ndarray:
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().or_chacha_fr(ctypes.c_uint64(seed), ctypes.c_uint64(start), ctypes.c_size_t(n), _p(out))
    return out


def fr_inner_product(a, b) -> np.ndarray:
    a, b = _c(a).reshape(-1, 4), _c(b).reshape(-1, 4)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_fr_inner_product(_p(out), _p(a), _p(b), ctypes.c_size_t(len(a)))
    return out


def fr_powers(base, n: int) -> np.ndarray:
    """out[i] = base^i (Montgomery)."""
    from . import pyoracle

    ones = np.tile(np.array(pyoracle.int_to_limbs(pyoracle.to_mont(1)), dtype=np.uint64), (n, 1))
    return distribute_powers_inplace(ones, base)


def g1_generator() -> np.ndarray:
    out = np.zeros(8, dtype=np.uint64)
    lib().or_g1_generator(_p(out))
    return out


def g1_mul_gen(scalars, nthreads: int | None = None) -> np.ndarray:
    s = _c(scalars).reshape(-1, 4)
    out = np.zeros((len(s), 8), dtype=np.uint64)
    lib().or_g1_mul_gen(_p(out), _p(s), ctypes.c_size_t(len(s)), nthreads or num_threads())
    return out


def g1_mul(p_aff, s) -> np.ndarray:
    out = np.zeros(12, dtype=np.uint64)
    lib().or_g1_mul(_p(out), _p(_c(p_aff).reshape(8)), _p(_c(s).reshape(4)))
    return out


def g1_is_on_curve(p_aff) -> bool:
    return bool(lib().or_g1_is_on_curve(_p(_c(p_aff).reshape(8))))


def msm(scalars, points, nthreads: int | None = None) -> np.ndarray:
    s = _c(scalars).reshape(-1, 4)
    pts = _c(points).reshape(-1, 8)
    assert len(s) == len(pts)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_msm(_p(out), _p(s), _p(pts), ctypes.c_size_t(len(s)), nthreads or num_threads())
    return out


def msm_naive(scalars, points) -> np.ndarray:
    s = _c(scalars).reshape(-1, 4)
    pts = _c(points).reshape(-1, 8)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_msm_naive(_p(out), _p(s), _p(pts), ctypes.c_size_t(len(s)))
    return out


def fft(a, log_n: int, omega, nthreads: int | None = None) -> np.ndarray:
    out = _c(a).reshape(-1, 4).copy()
    assert len(out) == 1 << log_n
    lib().or_fft(_p(out), ctypes.c_uint32(log_n), _p(_c(omega).reshape(4)), nthreads or num_threads())
    return out


def dft_naive(a, log_n: int, omega) -> np.ndarray:
    a = _c(a).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().or_dft_naive(_p(out), _p(a), ctypes.c_uint32(log_n), _p(_c(omega).reshape(4)))
    return out


def scale_inplace(a: np.ndarray, c, nthreads: int | None = None) -> np.ndarray:
    lib().or_scale(_p(a), ctypes.c_size_t(len(a)), _p(_c(c).reshape(4)), nthreads or num_threads())
    return a


def distribute_powers_inplace(a: np.ndarray, c, nthreads: int | None = None) -> np.ndarray:
    lib().or_distribute_powers(_p(a), ctypes.c_size_t(len(a)), _p(_c(c).reshape(4)), nthreads or num_threads())
    return a


def distribute_powers_zeta(a, coset_powers) -> np.ndarray:
    out = _c(a).reshape(-1, 4).copy()
    cp = _c(coset_powers).reshape(-1, 4)
    lib().or_distribute_powers_zeta(_p(out), ctypes.c_size_t(len(out)), _p(cp), ctypes.c_uint32(len(cp)))
    return out


def divide_by_vanishing(a, t_evals) -> np.ndarray:
    out = _c(a).reshape(-1, 4).copy()
    t = _c(t_evals).reshape(-1, 4)
    lib().or_divide_by_vanishing(_p(out), ctypes.c_size_t(len(out)), _p(t), ctypes.c_uint32(len(t)))
    return out


def extended_from_lagrange_vec(parts) -> np.ndarray:
    parts = [_c(p).reshape(-1, 4) for p in parts]
    n = len(parts[0])
    out = np.zeros((n * len(parts), 4), dtype=np.uint64)
    lib().or_extended_from_lagrange_vec(_p(out), _ptr_array(parts), ctypes.c_uint32(len(parts)), ctypes.c_size_t(n))
    return out


def batch_invert(fld: int, a) -> np.ndarray:
    out = _c(a).reshape(-1, 4).copy()
    lib().or_batch_invert(fld, _p(out), ctypes.c_size_t(len(out)))
    return out


def grand_product(inp, prefix) -> np.ndarray:
    inp = _c(inp).reshape(-1, 4)
    out = np.zeros_like(inp)
    lib().or_grand_product(_p(out), _p(inp), _p(_c(prefix).reshape(4)), ctypes.c_size_t(len(inp)))
    return out


def lookup_product(pi, pt, ci, ct, beta, gamma) -> np.ndarray:
    arrs = [_c(x).reshape(-1, 4) for x in (pi, pt, ci, ct)]
    out = np.zeros_like(arrs[0])
    lib().or_lookup_product(_p(out), *[_p(x) for x in arrs], _p(_c(beta).reshape(4)),
                            _p(_c(gamma).reshape(4)), ctypes.c_size_t(len(out)))
    return out


def permutation_product(modified, perms, values, beta, gamma, delta, omega, deltaomega):
    """Returns (modified_out, deltaomega_out)."""
    mod = _c(modified).reshape(-1, 4).copy()
    perms = [_c(x).reshape(-1, 4) for x in perms]
    values = [_c(x).reshape(-1, 4) for x in values]
    dw = _c(deltaomega).reshape(4).copy()
    lib().or_permutation_product(_p(mod), _ptr_array(perms), _ptr_array(values), ctypes.c_size_t(len(perms)),
                                 _p(_c(beta).reshape(4)), _p(_c(gamma).reshape(4)), _p(_c(delta).reshape(4)),
                                 _p(_c(omega).reshape(4)), _p(dw), ctypes.c_size_t(len(mod)))
    return mod, dw


def quotient_lookups(values, table_values, product_coset, permuted_input_coset, permuted_table_coset,
                     l0, l_last, l_active_row, beta, gamma, y) -> np.ndarray:
    out = _c(values).reshape(-1, 4).copy()
    arrs = [_c(x).reshape(-1, 4) for x in (table_values, product_coset, permuted_input_coset,
                                          permuted_table_coset, l0, l_last, l_active_row)]
    lib().or_quotient_lookups(_p(out), *[_p(x) for x in arrs], _p(_c(beta).reshape(4)),
                              _p(_c(gamma).reshape(4)), _p(_c(y).reshape(4)), ctypes.c_size_t(len(out)))
    return out


def permutation_quotient(values, zs, perms, cols, l0, l_last, l_active, beta, gamma, y, delta, delta_start,
                         cur_ext_omega, omega, chunk_len, last_rotation, rot_scale, isize) -> np.ndarray:
    out = _c(values).reshape(-1, 4).copy()
    zs = [_c(x).reshape(-1, 4) for x in zs]
    perms = [_c(x).reshape(-1, 4) for x in perms]
    cols = [_c(x).reshape(-1, 4) for x in cols]
    sc = [_c(x).reshape(4) for x in (beta, gamma, y, delta, delta_start, cur_ext_omega, omega)]
    ls = [_c(x).reshape(-1, 4) for x in (l0, l_last, l_active)]
    lib().or_permutation_quotient(_p(out), _ptr_array(zs), _ptr_array(perms), _ptr_array(cols),
                                  *[_p(x) for x in ls], *[_p(x) for x in sc],
                                  ctypes.c_size_t(len(zs)), ctypes.c_size_t(chunk_len), ctypes.c_size_t(len(cols)),
                                  ctypes.c_int32(last_rotation), ctypes.c_int32(rot_scale), ctypes.c_int32(isize),
                                  ctypes.c_size_t(len(out)))
    return out


def quotient_rules(fixed, instance, advice, challenges, constants, y, rules, vp, nq: int, rot_scale: int,
                   row_start: int, row_end: int, prev=None) -> np.ndarray:
    """Oracle of the custom-gate rule interpreter over rows [row_start, row_end)."""
    fixed = [_c(x).reshape(-1, 4) for x in fixed]
    instance = [_c(x).reshape(-1, 4) for x in instance]
    advice = [_c(x).reshape(-1, 4) for x in advice]
    ch = _c(challenges).reshape(-1, 4) if len(challenges) else np.zeros((1, 4), dtype=np.uint64)
    cs = _c(constants).reshape(-1, 4)
    rules = _c(rules).reshape(-1)
    if len(rules) == 0:
        rules = np.zeros(2, dtype=np.uint64)
        num_rules = 0
    else:
        num_rules = len(rules) // 2
    num_vp = len(vp)
    vp = _c(vp).reshape(-1) if num_vp else np.zeros(1, dtype=np.uint64)
    out = np.zeros((row_end - row_start, 4), dtype=np.uint64)
    pv = _c(prev).reshape(-1, 4) if prev is not None else None
    empty = np.zeros((1, 4), dtype=np.uint64)
    lib().or_quotient_rules(_p(out), _p(pv) if pv is not None else None,
                            _ptr_array(fixed or [empty]), ctypes.c_size_t(len(fixed)),
                            _ptr_array(instance or [empty]), ctypes.c_size_t(len(instance)),
                            _ptr_array(advice or [empty]), ctypes.c_size_t(len(advice)),
                            _p(ch), _p(cs), _p(_c(y).reshape(4)), _p(rules), ctypes.c_size_t(num_rules),
                            _p(vp), ctypes.c_size_t(num_vp), ctypes.c_size_t(nq), ctypes.c_int64(rot_scale),
                            ctypes.c_size_t(row_start), ctypes.c_size_t(row_end))
    return out


def eval_poly(a, u) -> np.ndarray:
    """a(u) by Horner (halo2-axiom arithmetic::eval_polynomial)."""
    a = _c(a).reshape(-1, 4)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_eval_poly(_p(out), _p(a) if len(a) else None, ctypes.c_size_t(len(a)), _p(_c(u).reshape(4)))
    return out


def kate_division(a, u) -> np.ndarray:
    """(a(X) - a(u)) / (X - u) (halo2-axiom arithmetic::kate_division): len(a) - 1 coefficients."""
    a = _c(a).reshape(-1, 4)
    q = np.zeros((max(0, len(a) - 1), 4), dtype=np.uint64)
    if len(a) > 1:
        lib().or_kate_division(_p(q), _p(a), ctypes.c_size_t(len(a)), _p(_c(u).reshape(4)))
    return q


def permute_expression_pair(inp, table, n: int, usable_rows: int):
    """(permuted_input, permuted_table), n rows each (src/plonk/lookup/prover.rs:788-882);
    raises ValueError where the reference panics (an input value missing from the table)."""
    a, t = _c(inp).reshape(-1, 4), _c(table).reshape(-1, 4)
    pin = np.zeros((n, 4), dtype=np.uint64)
    ptab = np.zeros((n, 4), dtype=np.uint64)
    err = lib().or_permute_expression_pair(_p(pin), _p(ptab), _p(a), _p(t), ctypes.c_size_t(n),
                                           ctypes.c_size_t(usable_rows))
    if err:
        raise ValueError("input value missing from the table")
    return pin, ptab
