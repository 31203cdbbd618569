"""Opening-phase and key-generation polynomial helpers over the C ABI, mirroring
the reference's host wrappers in src/cuda/funcs/polynomial_ops.rs,
src/cuda/funcs/multiopen.rs and src/cuda/funcs/omega.rs (same names, argument
meaning and chunking).

Device polynomials are int64 torch tensors of 4*n words (Montgomery Fr limbs);
host polynomials are (n, 4) uint64 numpy arrays; scalars are Python ints.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import R_MOD, from_limbs, mont_limbs, unmont
from .quotient import query_device_free_bytes_for_chunking

DENSE_POWER_DEGREE = 10  # src/arithmetic.rs (DENSE_POWER_DEGREE), cuda/include/kernel/omega.h


def _host(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(-1, 4))


def _hobj(a: np.ndarray) -> _lib.FFITraitObject:
    return _lib.obj(a.ctypes.data)


def _scalar_host(x: int) -> np.ndarray:
    return np.ascontiguousarray(mont_limbs(x % R_MOD).reshape(1, 4))


def _scalar_dev(x: int) -> torch.Tensor:
    return dv.to_device(mont_limbs(x % R_MOD).reshape(1, 4))


def _value(row) -> int:
    return unmont(from_limbs(row))


def _len(t: torch.Tensor) -> int:
    return t.numel() // 4


# ------------------------------------------------------------------ evaluation

def eval_polynomial_device(d_poly: torch.Tensor, point: int, length: int | None = None) -> np.ndarray:
    """polynomial_ops.rs:441-497: p(point) for a device poly; the result (Montgomery
    limbs, shape (4,)) comes back to the host."""
    lib = _lib.load()
    n = _len(d_poly) if length is None else length
    ws = int(lib._halo2_eval_polynomial_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    pt = _scalar_host(point)
    res = ctypes.c_void_p()
    _lib.check(lib._halo2_eval_polynomial(dv.ptr(d_poly), ctypes.byref(_hobj(pt)), ctypes.byref(res), n,
                                          dv.ptr(scratch), ws, dv.stream()), "_halo2_eval_polynomial")
    off = (res.value - dv.ptr(scratch)) // 8  # the result lives inside the scratch
    return scratch.view(torch.int64)[off:off + 4].cpu().numpy().view(np.uint64).copy()


def eval_polynomial_gpu(poly, point: int) -> np.ndarray:
    """polynomial_ops.rs:25-29: stage a host poly and evaluate on the device."""
    return eval_polynomial_device(dv.to_device(_host(poly)), point)


def batch_eval_polynomial_device_out(d_polys: list[torch.Tensor], points: list[int]) -> torch.Tensor:
    """polynomial_ops.rs:142-222: device polys in, a device buffer of results out
    (slot i = d_polys[i](points[i])), all polynomials in one batched launch sequence
    (`h2amd_eval_polynomial_batch_device`)."""
    assert len(d_polys) == len(points) and d_polys, "empty batch"
    lib = _lib.load()
    count = len(d_polys)
    out = dv.empty_u64(4 * count)
    lens = np.ascontiguousarray([_len(p) for p in d_polys], dtype=np.uint64)
    pts = np.ascontiguousarray(np.stack([mont_limbs(x % R_MOD) for x in points]))
    ptrs = (ctypes.c_void_p * count)(*[dv.ptr(p) for p in d_polys])
    ws = int(lib.h2amd_eval_polynomial_batch_device_workspace_size(lens.ctypes.data, count))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib.h2amd_eval_polynomial_batch_device(ptrs, lens.ctypes.data, pts.ctypes.data, count, dv.ptr(out),
                                                      dv.ptr(scratch), ws, dv.stream()),
               "h2amd_eval_polynomial_batch_device")
    return out


def batch_eval_polynomial_d2h(d_polys: list[torch.Tensor], points: list[int]) -> np.ndarray:
    """polynomial_ops.rs:224-257: device polys, one batched copy of the results to the host."""
    if not d_polys:
        return np.zeros((0, 4), dtype=np.uint64)
    return dv.to_host(batch_eval_polynomial_device_out(d_polys, points))


def basic_batch_eval_polynomial_gpu(polys: list[np.ndarray], points: list[int], poly_offset: int,
                                    poly_length: int) -> np.ndarray:
    """polynomial_ops.rs:32-67: host polys, coefficients [offset, offset + length) of each,
    evaluated as a polynomial of degree < length (results (batch, 4))."""
    lib = _lib.load()
    polys = [_host(p) for p in polys]
    batch = len(polys)
    assert batch == len(points)
    pts = np.ascontiguousarray(np.stack([mont_limbs(x % R_MOD) for x in points]))
    res = np.zeros((batch, 4), dtype=np.uint64)
    ws = int(lib._halo2_eval_polynomial_batch_workspace_size(poly_length, batch))
    scratch = dv.empty_bytes(ws)
    arr = _lib.obj_array([p.ctypes.data for p in polys])
    _lib.check(lib._halo2_eval_polynomial_batch(arr, ctypes.byref(_hobj(pts)), ctypes.byref(_hobj(res)),
                                                poly_offset, poly_length, batch, dv.ptr(scratch), ws, dv.stream()),
               "_halo2_eval_polynomial_batch")
    return res


def _chunk_plan(poly_length: int, max_length: int):
    if max_length == 0:
        raise _lib.Halo2GpuError(2, "not enough device memory for one coefficient per polynomial")
    num_chunks = -(-poly_length // max_length)
    for c in range(num_chunks):
        off = c * max_length
        yield off, (poly_length - off) if c == num_chunks - 1 else max_length


def batch_eval_polynomial_gpu(polys: list[np.ndarray], points: list[int], poly_length: int,
                              free_bytes: int | None = None) -> np.ndarray:
    """polynomial_ops.rs:70-127: memory-aware batch evaluation; when the batch does not fit,
    coefficient chunks are evaluated separately, scaled by point^offset and summed."""
    lib = _lib.load()
    batch = len(polys)
    free = query_device_free_bytes_for_chunking() if free_bytes is None else free_bytes
    max_length = int(lib._halo2_eval_poly_batch_max_len(poly_length, batch, free))
    if poly_length <= max_length:
        return basic_batch_eval_polynomial_gpu(polys, points, 0, poly_length)
    acc = [0] * batch
    for off, ln in _chunk_plan(poly_length, max_length):
        part = basic_batch_eval_polynomial_gpu(polys, points, off, ln)
        for i in range(batch):
            acc[i] = (acc[i] + _value(part[i]) * pow(points[i], off, R_MOD)) % R_MOD
    return np.stack([mont_limbs(v) for v in acc])


# ------------------------------------------------------------------ multiopen

def batch_multiopen_poly_calculation_gpu(polys: list[np.ndarray], poly_acc: np.ndarray, poly_offset: int,
                                         poly_length: int, challenge_point: list[int],
                                         evaluate_point: list[int]) -> np.ndarray:
    """multiopen.rs:14-55: over coefficients [offset, offset + length), writes
    poly_acc = sum_i challenge_i * poly_i into that slice of `poly_acc` (host, in place)
    and returns the slice evaluations (batch, 4)."""
    lib = _lib.load()
    assert poly_acc.dtype == np.uint64 and poly_acc.flags["C_CONTIGUOUS"]
    polys = [_host(p) for p in polys]
    batch = len(polys)
    ch = np.ascontiguousarray(np.stack([mont_limbs(x % R_MOD) for x in challenge_point]))
    ep = np.ascontiguousarray(np.stack([mont_limbs(x % R_MOD) for x in evaluate_point]))
    res = np.zeros((batch, 4), dtype=np.uint64)
    ws = int(lib._halo2_multiopen_poly_calculation_workspace_size(poly_length, batch))
    scratch = dv.empty_bytes(ws)
    arr = _lib.obj_array([p.ctypes.data for p in polys])
    _lib.check(lib._halo2_multiopen_poly_calculation(arr, ctypes.byref(_hobj(poly_acc)), poly_offset, poly_length,
                                                     batch, ctypes.byref(_hobj(ch)), ctypes.byref(_hobj(ep)),
                                                     ctypes.byref(_hobj(res)), dv.ptr(scratch), ws, dv.stream()),
               "_halo2_multiopen_poly_calculation")
    return res


def multiopen_poly_calculation_gpu(polys: list[np.ndarray], challenge_point: list[int], poly_acc: np.ndarray,
                                   evaluate_point: list[int], poly_length: int,
                                   free_bytes: int | None = None) -> np.ndarray:
    """multiopen.rs:57-125: memory-aware multiopen fold + evaluation (chunks combined with
    point^offset as in the batch evaluation)."""
    lib = _lib.load()
    batch = len(polys)
    free = query_device_free_bytes_for_chunking() if free_bytes is None else free_bytes
    max_length = int(lib._halo2_multiopen_poly_max_len(poly_length, batch, free))
    if poly_length <= max_length:
        return batch_multiopen_poly_calculation_gpu(polys, poly_acc, 0, poly_length, challenge_point,
                                                    evaluate_point)
    acc = [0] * batch
    for off, ln in _chunk_plan(poly_length, max_length):
        part = batch_multiopen_poly_calculation_gpu(polys, poly_acc, off, ln, challenge_point, evaluate_point)
        for i in range(batch):
            acc[i] = (acc[i] + _value(part[i]) * pow(evaluate_point[i], off, R_MOD)) % R_MOD
    return np.stack([mont_limbs(v) for v in acc])


# ------------------------------------------------------------ elementwise ops

def poly_multiply_add_device_with_d_scalar(d_acc: torch.Tensor, d_in: torch.Tensor, d_scalar: torch.Tensor,
                                           length: int | None = None) -> None:
    """polynomial_ops.rs:348-372: d_acc[i] += *d_scalar * d_in[i]."""
    n = _len(d_acc) if length is None else length
    _lib.check(_lib.load()._halo2_poly_multiply_add(dv.ptr(d_acc), dv.ptr(d_in), dv.ptr(d_scalar), n, dv.stream()),
               "_halo2_poly_multiply_add")


def poly_multiply_add_device(d_acc: torch.Tensor, d_in: torch.Tensor, scalar: int) -> None:
    """polynomial_ops.rs:326-346."""
    assert _len(d_acc) == _len(d_in)
    poly_multiply_add_device_with_d_scalar(d_acc, d_in, _scalar_dev(scalar))


def poly_multiply_add_many_device(d_acc: torch.Tensor, polys: list, d_scalars: torch.Tensor) -> None:
    """d_acc += sum_j d_scalars[j] * polys[j] in one pass per 16 polynomials
    (h2amd_poly_multiply_add_many); a polynomial shorter than d_acc counts as zero-padded.
    The SHPLONK folds of prover.rs:183-268 and 331-386, one call per fold instead of one
    _halo2_poly_multiply_add per polynomial."""
    n = _len(d_acc)
    count = len(polys)
    assert _len(d_scalars) >= count, "fewer scalars than polynomials"
    if count == 0:
        return
    lens = [_len(p) for p in polys]
    assert max(lens) <= n, "a polynomial longer than the accumulator"
    ptrs = (ctypes.c_void_p * count)(*[dv.ptr(p) for p in polys])
    lengths = (ctypes.c_uint64 * count)(*lens)
    _lib.check(_lib.load().h2amd_poly_multiply_add_many(dv.ptr(d_acc), ptrs, lengths, dv.ptr(d_scalars), count, n,
                                                         dv.stream()), "h2amd_poly_multiply_add_many")


def poly_multiply_add_device_at_lut_offset(d_acc: torch.Tensor, d_in: torch.Tensor, d_lut: torch.Tensor,
                                           lut_offset: int) -> None:
    """polynomial_ops.rs:397-439: the scalar is entry `lut_offset` of a device table."""
    assert _len(d_acc) == _len(d_in)
    assert lut_offset < _len(d_lut), "offset out of range"
    _lib.check(_lib.load()._halo2_poly_multiply_add(dv.ptr(d_acc), dv.ptr(d_in), dv.ptr(d_lut) + 32 * lut_offset,
                                                    _len(d_acc), dv.stream()), "_halo2_poly_multiply_add")


def poly_scale_device_with_d_s_minus_one(d_buf: torch.Tensor, d_s_minus_one: torch.Tensor) -> None:
    """polynomial_ops.rs:499-539: d_buf *= s through d_buf += (s - 1) * d_buf."""
    poly_multiply_add_device_with_d_scalar(d_buf, d_buf, d_s_minus_one)


def poly_multiply_add_single_gpu(poly_acc: np.ndarray, poly_in, scalar: int) -> np.ndarray:
    """polynomial_ops.rs:259-324: host slices, returns the new accumulator."""
    d_acc = dv.to_device(_host(poly_acc))
    poly_multiply_add_device(d_acc, dv.to_device(_host(poly_in)), scalar)
    return dv.to_host(d_acc)


def poly_elementwise_multiply_device(d_out: torch.Tensor, d_a: torch.Tensor, d_b: torch.Tensor) -> None:
    """culib.rs:276-283: d_out[i] = d_a[i] * d_b[i]."""
    _lib.check(_lib.load()._halo2_poly_elementwise_multiply(dv.ptr(d_out), dv.ptr(d_a), dv.ptr(d_b), _len(d_out),
                                                            dv.stream()), "_halo2_poly_elementwise_multiply")


def poly_fill_one_device(d_out: torch.Tensor) -> None:
    """polynomial_ops.rs:374-395: every entry becomes one."""
    if _len(d_out) == 0:
        return
    _lib.check(_lib.load()._halo2_poly_fill_one(dv.ptr(d_out), _len(d_out), dv.stream()), "_halo2_poly_fill_one")


def poly_sub_short_in_place_device(d_acc: torch.Tensor, d_short: torch.Tensor) -> None:
    """polynomial_ops.rs:694-728."""
    short_len = _len(d_short)
    if short_len == 0:
        return
    assert _len(d_acc) >= short_len, "short_len exceeds acc len"
    _lib.check(_lib.load()._halo2_poly_sub_short_inplace(dv.ptr(d_acc), dv.ptr(d_short), short_len, dv.stream()),
               "_halo2_poly_sub_short_inplace")


def poly_sub_short_out_of_place_device(d_out: torch.Tensor, d_long: torch.Tensor, d_short: torch.Tensor | None) -> None:
    """polynomial_ops.rs:730-772."""
    long_len = _len(d_long)
    short_len = 0 if d_short is None else _len(d_short)
    assert _len(d_out) == long_len and short_len <= long_len
    if long_len == 0:
        return
    _lib.check(_lib.load()._halo2_poly_sub_short_out_of_place(dv.ptr(d_out), dv.ptr(d_long),
                                                              dv.ptr(d_short) if short_len else None, short_len,
                                                              long_len, dv.stream()),
               "_halo2_poly_sub_short_out_of_place")


def poly_sub_scalar_at_zero_device(d_buf: torch.Tensor, d_scalar: torch.Tensor) -> None:
    """polynomial_ops.rs:774-804: d_buf[0] -= *d_scalar."""
    _lib.check(_lib.load()._halo2_poly_sub_scalar_at_zero(dv.ptr(d_buf), dv.ptr(d_scalar), dv.stream()),
               "_halo2_poly_sub_scalar_at_zero")


# ------------------------------------------------------------- kate division

@functools.lru_cache(maxsize=64)
def _kate_workspace(n: int) -> int:
    return int(_lib.load()._halo2_kate_division_workspace_size(n))


def _kate(d_a: torch.Tensor, d_root: torch.Tensor, out_len: int | None) -> torch.Tensor:
    lib = _lib.load()
    n = _len(d_a)
    q_len = max(0, n - 1)
    if out_len is None:
        if q_len == 0:
            return dv.empty_u64(0)[:0]
        out = dv.empty_u64(4 * q_len)
    else:
        assert out_len >= q_len, f"out_len {out_len} < n-1 {q_len}"
        if out_len == 0:
            return dv.empty_u64(0)[:0]
        out = dv.empty_u64(4 * out_len)
    ws = _kate_workspace(n)
    scratch = dv.empty_bytes(ws)
    if out_len is None:
        err = lib._halo2_kate_division_device(dv.ptr(d_a), dv.ptr(out), dv.ptr(d_root), n, dv.ptr(scratch), ws,
                                              dv.stream())
    else:
        err = lib._halo2_kate_division_device_padded(dv.ptr(d_a), dv.ptr(out), dv.ptr(d_root), n, out_len,
                                                     dv.ptr(scratch), ws, dv.stream())
    _lib.check(err, "_halo2_kate_division_device")
    return out


def kate_division_device(d_a: torch.Tensor, root: int) -> torch.Tensor:
    """polynomial_ops.rs:541-583: (a(X) - a(root)) / (X - root), n - 1 device coefficients."""
    return _kate(d_a, _scalar_dev(root), None)


def kate_division_device_with_d_root(d_a: torch.Tensor, d_root: torch.Tensor) -> torch.Tensor:
    """polynomial_ops.rs:585-639."""
    return _kate(d_a, d_root, None)


def kate_division_device_padded_with_d_root(d_a: torch.Tensor, d_root: torch.Tensor, out_len: int) -> torch.Tensor:
    """polynomial_ops.rs:641-692: quotient then zeros up to out_len (>= n - 1)."""
    return _kate(d_a, d_root, out_len)


# ----------------------------------------------------------- Assigned decode

def decode_assigned_to_num_denom_device(raw: np.ndarray, n: int, layout: _lib.AssignedLayout):
    """polynomial_ops.rs:822-867: `raw` holds n Assigned<F> elements as bytes (stride
    layout.stride_bytes); returns device (numerators, denominators)."""
    if n == 0:
        return dv.empty_u64(0)[:0], dv.empty_u64(0)[:0]
    d_raw = torch.from_numpy(np.ascontiguousarray(raw, dtype=np.uint8).reshape(-1).copy()).to(dv.device())
    nums, dens = dv.empty_u64(4 * n), dv.empty_u64(4 * n)
    _lib.check(_lib.load()._halo2_decode_assigned(dv.ptr(nums), dv.ptr(dens), dv.ptr(d_raw), n, layout, dv.stream()),
               "_halo2_decode_assigned")
    dv.sync()
    return nums, dens


def decode_assigned_into_denom_slice_device(raw: np.ndarray, n: int, layout: _lib.AssignedLayout,
                                            d_denoms_concat: torch.Tensor, denom_dst_offset: int) -> torch.Tensor:
    """polynomial_ops.rs:869-920: denominators written into d_denoms_concat at an offset;
    numerators returned."""
    assert denom_dst_offset + n <= _len(d_denoms_concat), "dst range out of bounds"
    if n == 0:
        return dv.empty_u64(0)[:0]
    d_raw = torch.from_numpy(np.ascontiguousarray(raw, dtype=np.uint8).reshape(-1).copy()).to(dv.device())
    nums = dv.empty_u64(4 * n)
    _lib.check(_lib.load()._halo2_decode_assigned(dv.ptr(nums), dv.ptr(d_denoms_concat) + 32 * denom_dst_offset,
                                                  dv.ptr(d_raw), n, layout, dv.stream()), "_halo2_decode_assigned")
    dv.sync()
    return nums


# ------------------------------------------------------------- omega tables

def power_of_omega_gpu(omega: int, log_n: int, pow_: int) -> tuple[np.ndarray, np.ndarray]:
    """omega.rs:15-49: (omega^pow, the log table [1, omega, omega^2, omega^4, ...])."""
    lib = _lib.load()
    res = np.zeros((1, 4), dtype=np.uint64)
    lut = np.zeros((log_n + 1, 4), dtype=np.uint64)
    om = _scalar_host(omega)
    ws = int(lib._halo2_power_of_omega_workspace_size(log_n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_power_of_omega(ctypes.byref(_hobj(res)), ctypes.byref(_hobj(lut)), ctypes.byref(_hobj(om)),
                                         log_n, pow_, dv.ptr(scratch), ws, dv.stream()), "_halo2_power_of_omega")
    return res[0], lut


def generate_omega_lut_gpu(omega: int, log_n: int) -> np.ndarray:
    """omega.rs:105-139: omega^i for i < 2^10, then omega^(i * 2^10) for i < 2^(log_n - 10)."""
    lib = _lib.load()
    lo = 1 << DENSE_POWER_DEGREE
    hi = 1 << max(0, log_n - DENSE_POWER_DEGREE)
    lut = np.zeros((lo + hi, 4), dtype=np.uint64)
    om = _scalar_host(omega)
    ws = int(lib._halo2_generate_omega_lut_workspace_size(log_n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_generate_omega_lut(ctypes.byref(_hobj(lut)), ctypes.byref(_hobj(om)), log_n,
                                             dv.ptr(scratch), ws, dv.stream()), "_halo2_generate_omega_lut")
    return lut


def generate_omega_powers_gpu(omega: int, log_n: int, output_num: int) -> np.ndarray:
    """omega.rs:51-103: the first output_num powers omega^i, to the host."""
    assert output_num <= 1 << log_n
    lib = _lib.load()
    out = dv.empty_u64(4 << log_n)
    d_omega = _scalar_dev(omega)
    ws = int(lib._halo2_generate_omega_powers_workspace_size(log_n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_generate_omega_powers(dv.ptr(out), dv.ptr(d_omega), log_n, dv.ptr(scratch), ws,
                                                dv.stream()), "_halo2_generate_omega_powers")
    return dv.to_host(out)[:output_num]


def generate_omegadelta_device(mapping: np.ndarray, omega: int, delta: int, log_n: int, omega_start: int,
                               omega_end: int, column_num: int, column_offset: int = 0) -> torch.Tensor:
    """culib.rs:469-490 / omega.cu:163-210: out[i - start] = omega^row * delta^col for
    (col, row) = mapping[i], i in [omega_start, omega_end]; `mapping` is (rows, 2) u32."""
    lib = _lib.load()
    m = np.ascontiguousarray(np.asarray(mapping, dtype=np.uint32).reshape(-1, 2))
    assert omega_end < len(m) and omega_start <= omega_end
    assert int(m[:, 0].max(initial=0)) < max(1, column_num) and int(m[:, 1].max(initial=0)) < (1 << log_n)
    d_map = torch.from_numpy(m.view(np.int32).reshape(-1).copy()).to(dv.device())
    out = dv.empty_u64(4 * (omega_end - omega_start + 1))
    ws = int(lib._halo2_generate_omegadelta_workspace_size(log_n, omega_start, omega_end, column_num, column_offset))
    scratch = dv.empty_bytes(ws)
    om, de = _scalar_host(omega), _scalar_host(delta)
    _lib.check(lib._halo2_generate_omegadelta(dv.ptr(out), dv.ptr(d_map), om.ctypes.data, de.ctypes.data, log_n,
                                              omega_start, omega_end, column_num, column_offset, dv.ptr(scratch), ws,
                                              dv.stream()), "_halo2_generate_omegadelta")
    dv.sync()
    return out
