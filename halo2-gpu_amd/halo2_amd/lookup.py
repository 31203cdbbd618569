"""Lookup-argument wrappers over the C ABI, mirroring src/cuda/funcs/lookup.rs of the
reference: the lookup product (host chunked and device-resident) and the lookup
permutation pair (permute_expression_pair_device).

Device polynomials are int64 torch tensors of 4*n words (Montgomery Fr limbs);
host polynomials are (n, 4) uint64 numpy arrays; scalars are Python ints.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import R_MOD, mont_limbs
from .quotient import query_device_free_bytes_for_chunking


def _scalar_dev(x: int) -> torch.Tensor:
    return dv.to_device(mont_limbs(x % R_MOD).reshape(1, 4))


def _len(t: torch.Tensor) -> int:
    return t.numel() // 4


def lookup_product_gpu(permuted_input: np.ndarray, permuted_table: np.ndarray, compressed_input: np.ndarray,
                       compressed_table: np.ndarray, beta: int, gamma: int,
                       free_bytes: int | None = None) -> np.ndarray:
    """lookup.rs:22-90: host polynomials, chunked by the free device memory;
    out[i] = (A_i + beta)(S_i + gamma) / ((A'_i + beta)(S'_i + gamma))."""
    lib = _lib.load()
    arrs = [np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
            for a in (permuted_input, permuted_table, compressed_input, compressed_table)]
    n = len(arrs[0])
    assert all(len(a) == n for a in arrs)
    out = np.zeros((n, 4), dtype=np.uint64)
    if n == 0:
        return out
    free = query_device_free_bytes_for_chunking() if free_bytes is None else free_bytes
    chunk = min(n, int(lib._halo2_commit_product_max_len(n, free)))
    if chunk == 0:
        raise _lib.Halo2GpuError(2, "lookup_product_gpu: not enough device memory")
    ws = int(lib._halo2_commit_product_workspace_size(chunk))
    scratch = dv.empty_bytes(ws)
    b, g = _scalar_dev(beta), _scalar_dev(gamma)
    objs = [_lib.obj(a.ctypes.data) for a in [out] + arrs]
    for off in range(0, n, chunk):
        _lib.check(lib._halo2_commit_product(*objs, dv.ptr(b), dv.ptr(g), min(chunk, n - off), off,
                                             dv.ptr(scratch), ws, dv.stream()), "_halo2_commit_product")
    return out


def lookup_product_device(d_permuted_input: torch.Tensor, d_permuted_table: torch.Tensor,
                          d_compressed_input: torch.Tensor, d_compressed_table: torch.Tensor, beta: int,
                          gamma: int) -> torch.Tensor:
    """lookup.rs:96-147: device inputs, a fresh device output."""
    lib = _lib.load()
    n = _len(d_permuted_input)
    assert _len(d_permuted_table) == n and _len(d_compressed_input) == n and _len(d_compressed_table) == n
    out = dv.empty_u64(4 * n)
    b, g = _scalar_dev(beta), _scalar_dev(gamma)
    objs = [_lib.obj(dv.ptr(t)) for t in (out, d_permuted_input, d_permuted_table, d_compressed_input,
                                          d_compressed_table)]
    _lib.check(lib._halo2_commit_product_device_inputs(*objs, dv.ptr(b), dv.ptr(g), n, 0, dv.stream()),
               "_halo2_commit_product_device_inputs")
    return out


def permute_expression_pair_device(d_compressed_input: torch.Tensor, d_compressed_table: torch.Tensor,
                                   d_permuted_input: torch.Tensor, d_permuted_table: torch.Tensor, n: int,
                                   usable_rows: int) -> None:
    """lookup.rs:212-256: the sorted permutation pair into caller-owned device buffers of
    n rows; rows [usable_rows, n) are zero."""
    for t in (d_compressed_input, d_compressed_table, d_permuted_input, d_permuted_table):
        assert _len(t) == n
    assert usable_rows <= n
    lib = _lib.load()
    ws = int(lib._halo2_permute_expression_pair_workspace_size(n, usable_rows))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_permute_expression_pair(dv.ptr(d_compressed_input), dv.ptr(d_compressed_table),
                                                  dv.ptr(d_permuted_input), dv.ptr(d_permuted_table), n,
                                                  usable_rows, dv.ptr(scratch), ws, dv.stream()),
               "_halo2_permute_expression_pair")
