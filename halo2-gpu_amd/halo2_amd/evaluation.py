"""h(X) on the extended domain, part by part, on the device: the orchestration of
`evaluate_h_device` / `evaluate_h_inner` (src/plonk/evaluation.rs:1056-1417 of the
reference; the CPU form it must agree with is src/cpu/evaluator.rs:150-411).

For each of the 2^(extended_k - k) parts (coset shift zeta * extended_omega^p):
  1. every fixed / selector / advice / instance polynomial is moved to the part
     (`coeff_to_extended_part`),
  2. custom gates: the rule interpreter over the part's columns (`_halo2_quotient*`),
  3. permutation identities (`_halo2_quotient_permutation`) over the cosets of the
     grand products and of the sigma polynomials,
  4. per lookup: the compressed (input + beta)(table + gamma) values through the rule
     interpreter with runtime constants, then `_halo2_quotient_lookups` over the cosets
     of the product and permuted polynomials,
then the parts are interleaved into the extended Lagrange vector
(`_halo2_extended_from_lagrange_vec_device`). One circuit (the reference folds a batch
of circuits through the same loop).
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import DELTA, R_MOD, ZETA, mont_limbs
from .quotient import ColumnPool, GraphEvaluator, compress_expressions_with_runtime_constants_device, \
    quotient_device


@dataclass
class LookupCommitted:
    """lookup::prover::Committed: the product polynomial (coefficients) and the permuted
    input / table expressions (Lagrange values), plus the lookup's evaluator graph.
    The permuted columns' coefficient forms, when the caller has them, save their
    inverse transforms in every part."""

    graph: GraphEvaluator
    product_poly: torch.Tensor
    permuted_input: torch.Tensor
    permuted_table: torch.Tensor
    permuted_input_poly: torch.Tensor | None = None
    permuted_table_poly: torch.Tensor | None = None


@dataclass
class PermutationCommitted:
    """permutation::prover::Committed: one grand-product polynomial (coefficients) per set,
    the columns of the argument as (kind, index) with kind in fixed/advice/instance, and
    the proving key's sigma polynomials (coefficients), one per column."""

    product_polys: list
    columns: list
    sigma_polys: list = field(default_factory=list)


_key_lock = threading.Lock()


def _host_scalar(v: int) -> np.ndarray:
    return np.ascontiguousarray(mont_limbs(v % R_MOD))


def _ptr_table(tensors) -> torch.Tensor:
    return dv.to_device(np.array([dv.ptr(t) for t in tensors] or [0], dtype=np.uint64))


def lagrange_to_extend_part(domain, poly: torch.Tensor, extended_omega_factor: int) -> torch.Tensor:
    """domain.rs:281: Lagrange values -> coefficients -> the coset part."""
    coeff = domain.lagrange_to_coeff_device_input(poly)
    return domain.coeff_to_extended_part_many_device([coeff], extended_omega_factor)[0]


def evaluate_h_device(domain, gates: GraphEvaluator, fixed_polys, l0, l_last, l_active, advice_polys,
                      instance_polys, challenges, y: int, beta: int, gamma: int, theta: int,
                      lookups: list[LookupCommitted] = (), permutation: PermutationCommitted | None = None,
                      blinding_factors: int = 0, cs_degree: int = 3, key_parts: dict | None = None) -> torch.Tensor:
    """Extended Lagrange values of h's numerator (before the vanishing division), one
    circuit; all polynomials are device tensors of n coefficients (permuted lookup
    columns: n Lagrange values).

    key_parts: a dict owned by the proving key that keeps, per part, the coset values of
    the polynomials that do not change between proofs (fixed columns, l0, l_last,
    l_active, sigma) -- the proving key's fixed_cosets / permutation cosets of halo2
    (keygen.rs), 4 (nf + 3 + sigmas) x 32 n bytes of HBM; filled on first use."""
    lib = _lib.load()
    n = domain.n
    num_parts = domain.extended_len >> domain.k
    stream = dv.stream()
    cur = 1
    nf, na = len(fixed_polys), len(advice_polys)
    value_parts = []
    sigma_polys = list(permutation.sigma_polys) if permutation is not None and permutation.product_polys else []
    # the lookup identities' beta, gamma, y: one upload for every part
    d_lsc = dv.to_device(np.stack([mont_limbs(v % R_MOD) for v in (beta, gamma, y)])) if lookups else None
    d_sc = [d_lsc[4 * i:4 * i + 4] for i in range(3)] if lookups else []
    for part in range(num_parts):
        key = key_parts.get(part) if key_parts is not None else None
        if key is None:
            with _key_lock:  # concurrent first proofs on one key: one builder per part
                key = key_parts.get(part) if key_parts is not None else None
                if key is None:
                    key = domain.coeff_to_extended_part_many_device(
                        list(fixed_polys) + [l0, l_last, l_active] + sigma_polys, cur)
                    if key_parts is not None:
                        # readers run on other streams: publish only once the parts are written
                        torch.cuda.current_stream().synchronize()
                        key_parts[part] = key
        fixed = key[:nf]
        l0_e, ll_e, la_e = key[nf:nf + 3]
        sig = key[nf + 3:]
        # every per-proof polynomial of the part in one batched transform
        per_proof = list(advice_polys) + list(instance_polys)
        if permutation is not None and permutation.product_polys:
            per_proof += list(permutation.product_polys)
        for lk in lookups:
            per_proof.append(lk.product_poly)
            if lk.permuted_input_poly is not None and lk.permuted_table_poly is not None:
                per_proof += [lk.permuted_input_poly, lk.permuted_table_poly]
        parts = domain.coeff_to_extended_part_many_device(per_proof, cur)
        advice = parts[:na]
        instance = parts[na:na + len(instance_polys)]
        rest = parts[na + len(instance_polys):]
        pool = ColumnPool(n)
        pool.try_init_device(fixed, [], advice, instance)

        values = quotient_device(gates, fixed, instance, advice, challenges, y, n, 1)

        if permutation is not None and permutation.product_polys:
            chunk_len = cs_degree - 2
            last_rotation = -(blinding_factors + 1)
            zs, rest = rest[:len(permutation.product_polys)], rest[len(permutation.product_polys):]
            groups = {"fixed": fixed, "advice": advice, "instance": instance}
            cols = [groups[kind][idx] for kind, idx in permutation.columns]
            ptrs = _ptr_table(list(zs) + list(sig) + list(cols))  # one upload for the three tables
            tz, ts = ptrs[:len(zs)], ptrs[len(zs):len(zs) + len(sig)]
            tc = ptrs[len(zs) + len(sig):]
            sc = [_host_scalar(v) for v in (beta, gamma, y, DELTA, beta * ZETA, cur, domain.omega)]
            _lib.check(lib._halo2_quotient_permutation(
                dv.ptr(values), dv.ptr(l0_e), dv.ptr(ll_e), dv.ptr(la_e), dv.ptr(tz), dv.ptr(ts), dv.ptr(tc),
                *[ctypes.c_void_p(x.ctypes.data) for x in sc], len(zs), chunk_len, len(cols), last_rotation, 1, n,
                n, stream), "_halo2_quotient_permutation")

        if lookups:
            table = dv.empty_u64(4 * n)
            for lk in lookups:
                compress_expressions_with_runtime_constants_device(lk.graph, theta, beta, gamma, y, n, 1, pool,
                                                                   challenges, table)
                prod, rest = rest[0], rest[1:]
                if lk.permuted_input_poly is not None and lk.permuted_table_poly is not None:
                    (pin, ptab), rest = rest[:2], rest[2:]
                else:
                    pin = lagrange_to_extend_part(domain, lk.permuted_input, cur)
                    ptab = lagrange_to_extend_part(domain, lk.permuted_table, cur)
                _lib.check(lib._halo2_quotient_lookups(
                    dv.ptr(values), dv.ptr(table), dv.ptr(prod), dv.ptr(pin), dv.ptr(ptab), dv.ptr(l0_e),
                    dv.ptr(ll_e), dv.ptr(la_e), *[dv.ptr(x) for x in d_sc], n, stream), "_halo2_quotient_lookups")
        value_parts.append(values)
        cur = cur * domain.extended_omega % R_MOD
    return domain.extended_from_lagrange_vec_device(value_parts)
