"""NTT wrappers over the C ABI, mirroring src/cuda/funcs/ntt.rs of the reference
(workspace sizing, scratch allocation, device/host residency variants).

Device polynomials are int64 torch tensors of 4*n words (Montgomery Fr limbs);
host polynomials are (n, 4) uint64 numpy arrays. Scalars are Python ints.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import R_MOD, mont_bytes, mont_limbs

FFT, IFFT, COSET_FFT, ICOSET_FFT, IFFT_COSET_FFT, COSET_FFT_PART = 1, 2, 3, 4, 5, 6


def _scalar_dev(x: int) -> torch.Tensor:
    return dv.to_device(mont_limbs(x).reshape(1, 4))


def fft_normal_device(ntt_type: int, log_n: int, d_in: torch.Tensor, d_out: torch.Tensor, omega: int,
                      divisor: int, stream: int | None = None) -> None:
    """src/cuda/funcs/ntt.rs:176-207 (device input/output; in == out allowed)."""
    lib = _lib.load()
    ws = int(lib._halo2_fft_normal_to_device_workspace_size(ntt_type, log_n, log_n))
    scratch = dv.empty_bytes(ws)
    d_omega = _scalar_dev(omega)
    div = ctypes.create_string_buffer(mont_bytes(divisor), 32)
    err = lib._halo2_fft_normal_to_device(ntt_type, log_n, dv.ptr(d_in), dv.ptr(d_out), dv.ptr(d_omega), div,
                                          dv.ptr(scratch), ws, stream if stream is not None else dv.stream())
    _lib.check(err, "_halo2_fft_normal_to_device")


def fft_gpu(ntt_type: int, a: np.ndarray, log_n: int, omega: int, divisor: int) -> np.ndarray:
    """src/cuda/funcs/ntt.rs:209-227 (host in/out; returns a new array)."""
    lib = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    out = np.empty_like(a)
    ws = int(lib._halo2_fft_normal_workspace_size(ntt_type, log_n, log_n))
    scratch = dv.empty_bytes(ws)
    om = ctypes.create_string_buffer(mont_bytes(omega), 32)
    div = ctypes.create_string_buffer(mont_bytes(divisor), 32)
    err = lib._halo2_fft_normal(ntt_type, log_n, a.ctypes.data, out.ctypes.data, om, div, dv.ptr(scratch), ws,
                                dv.stream())
    _lib.check(err, "_halo2_fft_normal")
    return out


def cosetfft_gpu(a: np.ndarray, log_n: int, extend_log_n: int, omega: int) -> np.ndarray:
    """src/cuda/funcs/ntt.rs:229-259 (host in 2^log_n -> host out 2^extend_log_n)."""
    lib = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    out = np.empty((1 << extend_log_n, 4), dtype=np.uint64)
    ws = int(lib._halo2_fft_normal_workspace_size(COSET_FFT, log_n, extend_log_n))
    scratch = dv.empty_bytes(ws)
    om = ctypes.create_string_buffer(mont_bytes(omega), 32)
    div = ctypes.create_string_buffer(mont_bytes(1), 32)
    err = lib._halo2_cosetfft(COSET_FFT, log_n, extend_log_n, a.ctypes.data, out.ctypes.data, om, div,
                              dv.ptr(scratch), ws, dv.stream())
    _lib.check(err, "_halo2_cosetfft")
    return out


def fft_many_host(ntt_type: int, inputs: list[np.ndarray], log_n: int, extend_log_n: int, omega: int,
                  divisor: int) -> list[np.ndarray]:
    """src/cuda/funcs/ntt.rs:261-300 (host inputs and outputs through `_halo2_fft_many`)."""
    lib = _lib.load()
    ins = [np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4) for x in inputs]
    out_len = 1 << (extend_log_n if ntt_type in (COSET_FFT, ICOSET_FFT) else log_n)
    outs = [np.empty((out_len, 4), dtype=np.uint64) for _ in ins]
    ws = int(lib._halo2_fft_many_workspace_size(ntt_type, log_n, extend_log_n))
    scratch = dv.empty_bytes(ws)
    d_omega = _scalar_dev(omega)
    div_host = np.ascontiguousarray(mont_limbs(divisor))
    div_obj = _lib.obj(div_host.ctypes.data)
    err = lib._halo2_fft_many(ntt_type, len(ins), log_n, extend_log_n,
                              _lib.obj_array([x.ctypes.data for x in ins]),
                              _lib.obj_array([x.ctypes.data for x in outs]), dv.ptr(d_omega),
                              ctypes.byref(div_obj), dv.ptr(scratch), ws, dv.stream())
    _lib.check(err, "_halo2_fft_many")
    return outs


def fft_many_to_device(ntt_type: int, inputs, log_n: int, extend_log_n: int, omega: int, divisor: int,
                       outputs: list[torch.Tensor] | None = None) -> list[torch.Tensor]:
    """src/cuda/funcs/ntt.rs:302-440 (`_halo2_fft_many_to_device`): inputs are device tensors or
    host arrays; outputs are device tensors (allocated when not given)."""
    lib = _lib.load()
    on_dev = len(inputs) > 0 and isinstance(inputs[0], torch.Tensor)
    if on_dev:
        in_ptrs = [dv.ptr(x) for x in inputs]
        keep = inputs
    else:
        keep = [np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 4) for x in inputs]
        in_ptrs = [x.ctypes.data for x in keep]
    out_len = 1 << (extend_log_n if ntt_type in (COSET_FFT, ICOSET_FFT) else log_n)
    if outputs is None:
        outputs = [dv.empty_u64(4 * out_len) for _ in inputs]
    ws = int(lib._halo2_fft_many_to_device_workspace_size(ntt_type, log_n, extend_log_n))
    scratch = dv.empty_bytes(ws)
    d_omega = _scalar_dev(omega)
    div_host = np.ascontiguousarray(mont_limbs(divisor))
    div_obj = _lib.obj(div_host.ctypes.data)
    err = lib._halo2_fft_many_to_device(ntt_type, len(in_ptrs), log_n, extend_log_n, _lib.obj_array(in_ptrs),
                                        _lib.obj_array([dv.ptr(x) for x in outputs]), dv.ptr(d_omega),
                                        ctypes.byref(div_obj), on_dev, dv.ptr(scratch), ws, dv.stream())
    _lib.check(err, "_halo2_fft_many_to_device")
    del keep
    return outputs


def fft_batch_to_device(ntt_type: int, inputs: list[torch.Tensor], log_ns: list[int], extend_log_ns: list[int],
                        omegas: list[int], divisors: list[int],
                        outputs: list[torch.Tensor] | None = None) -> list[torch.Tensor]:
    """`h2amd_fft_batch_to_device`: transforms of uneven sizes (each as
    `_halo2_fft_many_to_device` would run it alone), every pass of all of them in one
    launch. Device inputs and outputs; outputs are allocated when not given."""
    lib = _lib.load()
    count = len(inputs)
    if count == 0:
        return []
    coset = ntt_type in (COSET_FFT, ICOSET_FFT)
    if outputs is None:
        outputs = [dv.empty_u64(4 << (e if coset else k)) for k, e in zip(log_ns, extend_log_ns)]
    ln = np.ascontiguousarray(log_ns, dtype=np.uint32)
    ex = np.ascontiguousarray(extend_log_ns, dtype=np.uint32)
    d_omegas = {}
    for w in omegas:  # one device word per distinct omega, so equal sizes share their tables
        if w % R_MOD not in d_omegas:
            d_omegas[w % R_MOD] = _scalar_dev(w)
    om = (ctypes.c_void_p * count)(*[dv.ptr(d_omegas[w % R_MOD]) for w in omegas])
    ins = (ctypes.c_void_p * count)(*[dv.ptr(x) for x in inputs])
    outs = (ctypes.c_void_p * count)(*[dv.ptr(x) for x in outputs])
    divs = [np.ascontiguousarray(mont_limbs(d)) for d in divisors]
    div_objs = _lib.obj_array([d.ctypes.data for d in divs])
    ws = int(lib.h2amd_fft_batch_to_device_workspace_size(ntt_type, count, ln.ctypes.data, ex.ctypes.data))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib.h2amd_fft_batch_to_device(ntt_type, count, ln.ctypes.data, ex.ctypes.data, ins, outs, om,
                                             div_objs, dv.ptr(scratch), ws, dv.stream()), "h2amd_fft_batch_to_device")
    del divs
    return outputs


def distribute_powers_zeta_device(d_a: torch.Tensor, coset_powers: list[int]) -> None:
    """src/cuda/funcs/ntt.rs:114-138."""
    lib = _lib.load()
    cp = dv.to_device(np.stack([mont_limbs(c) for c in coset_powers]))
    n = d_a.numel() // 4
    _lib.check(lib._halo2_distribute_powers_zeta(dv.ptr(d_a), dv.ptr(cp), len(coset_powers), n, dv.stream()),
               "_halo2_distribute_powers_zeta")


def divide_by_vanishing_poly_device(d_poly: torch.Tensor, d_t_evals: torch.Tensor) -> None:
    """src/cuda/funcs/ntt.rs:140-162."""
    lib = _lib.load()
    _lib.check(lib._halo2_divide_by_vanishing_poly(dv.ptr(d_poly), dv.ptr(d_t_evals), d_t_evals.numel() // 4,
                                                   d_poly.numel() // 4, dv.stream()),
               "_halo2_divide_by_vanishing_poly")


def extended_from_lagrange_vec_device(d_out: torch.Tensor, d_parts: list[torch.Tensor], n: int) -> None:
    """src/cuda/funcs/ntt.rs:80-110 (pointer table uploaded to the device)."""
    lib = _lib.load()
    table = dv.to_device(np.array([dv.ptr(p) for p in d_parts], dtype=np.uint64))
    _lib.check(lib._halo2_extended_from_lagrange_vec_device(dv.ptr(d_out), dv.ptr(table), len(d_parts), n,
                                                            dv.stream()),
               "_halo2_extended_from_lagrange_vec_device")
