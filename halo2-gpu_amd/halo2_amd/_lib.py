"""ctypes loader for libhalo2_gpu.so, the gfx950 C-ABI library.

The library is loaded after ``torch`` so that both share one HIP runtime
(same soname), which lets torch-allocated device buffers and streams cross the
boundary. There is no fallback: if the library is missing, every entry point
raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhalo2_gpu.so")
PKG_ROOT = os.path.dirname(_HERE)  # halo2-gpu_amd/


class RustError(ctypes.Structure):
    """cuda/include/common/halo2_ffi.h:30-48: {int32 code; char* message}."""

    _fields_ = [("code", ctypes.c_int32), ("message", ctypes.c_void_p)]


class FFITraitObject(ctypes.Structure):
    """cuda/include/common/halo2_ffi.h:51-60: {uint64_t* ptr}."""

    _fields_ = [("ptr", ctypes.c_void_p)]


class AssignedLayout(ctypes.Structure):
    """src/cuda/culib.rs:20-26: byte layout of one Assigned<F> element."""

    _fields_ = [("stride_bytes", ctypes.c_uint32), ("num_offset", ctypes.c_uint32),
                ("denom_offset", ctypes.c_uint32)]


class Halo2GpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]
_libc.free.restype = None

_lib = None

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
I32 = ctypes.c_int32
BOOL = ctypes.c_bool
OBJP = ctypes.POINTER(FFITraitObject)

# name -> (restype, argtypes); the C declarations are in include/halo2_gpu.h
SIGNATURES = {
    "_halo2_msm_max_length": (SZ, [U64]),
    "_halo2_multiexp_workspace_size": (U64, [U64]),
    "_halo2_multiexp": (RustError, [OBJP, OBJP, OBJP, SZ, P, U64, P]),
    "_halo2_multiexp_device_bases_workspace_size": (U64, [U64]),
    "_halo2_multiexp_device_bases": (RustError, [OBJP, P, OBJP, SZ, P, U64, P]),
    "_halo2_multiexp_device_scalars_device_bases_workspace_size": (U64, [U64]),
    "_halo2_multiexp_device_scalars_device_bases": (RustError, [P, P, OBJP, SZ, P, U64, P]),
    "h2amd_multiexp_noalloc_max_length": (U64, []),
    "h2amd_multiexp_device_scalars_device_bases_noalloc": (RustError, [P, P, OBJP, SZ, P]),
    "_halo2_fft_normal_workspace_size": (U64, [U32, U32, U32]),
    "_halo2_fft_normal": (RustError, [U32, U32, P, P, P, P, P, U64, P]),
    "_halo2_fft_normal_check_memory": (BOOL, [U32, P, U32, U32]),
    "_halo2_fft_normal_to_device_workspace_size": (U64, [U32, U32, U32]),
    "_halo2_fft_normal_to_device": (RustError, [U32, U32, P, P, P, P, P, U64, P]),
    "_halo2_cosetfft": (RustError, [U32, U32, U32, P, P, P, P, P, U64, P]),
    "_halo2_fft_many_workspace_size": (U64, [U32, U32, U32]),
    "_halo2_fft_many_to_device_workspace_size": (U64, [U32, U32, U32]),
    "_halo2_fft_many": (RustError, [U32, U32, U32, U32, OBJP, OBJP, P, OBJP, P, U64, P]),
    "_halo2_fft_many_to_device": (RustError, [U32, U32, U32, U32, OBJP, OBJP, P, OBJP, BOOL, P, U64, P]),
    "h2amd_fft_batch_to_device_workspace_size": (U64, [U32, U32, P, P]),
    "h2amd_fft_batch_to_device": (RustError, [U32, U32, P, P, P, P, P, OBJP, P, U64, P]),
    "_halo2_fft_many_check_memory": (BOOL, [U32, U32, U32]),
    "_halo2_get_fft_split_radix": (I32, [U32, U32, U32, U64]),
    "_halo2_extended_from_lagrange_vec_device": (RustError, [P, P, U32, U64, P]),
    "_halo2_divide_by_vanishing_poly": (RustError, [P, P, U32, U64, P]),
    "_halo2_distribute_powers_zeta": (RustError, [P, P, U32, U64, P]),
    "_halo2_batch_invert": (RustError, [P, U32, SZ, P]),
    "_halo2_commit_product_max_len": (SZ, [SZ, U64]),
    "_halo2_commit_product_workspace_size": (U64, [U64]),
    "_halo2_commit_product": (RustError, [OBJP, OBJP, OBJP, OBJP, OBJP, P, P, SZ, SZ, P, U64, P]),
    "_halo2_commit_product_device_inputs": (RustError, [OBJP, OBJP, OBJP, OBJP, OBJP, P, P, SZ, SZ, P]),
    "_halo2_permute_expression_pair_workspace_size": (U64, [U64, U64]),
    "_halo2_permute_expression_pair": (RustError, [P, P, P, P, U64, U64, P, U64, P]),
    "_halo2_permutation_product_max_len": (SZ, [SZ, U64]),
    "_halo2_permutation_product_workspace_size": (U64, [U64]),
    "_halo2_permutation_product": (RustError, [OBJP, OBJP, OBJP, P, P, P, P, P, SZ, SZ, SZ, P, U64, P]),
    "_halo2_permutation_product_device_inputs_workspace_size": (U64, [U64]),
    "_halo2_permutation_product_device_inputs": (RustError,
                                                 [OBJP, OBJP, OBJP, P, P, P, P, P, SZ, SZ, SZ, P, U64, P]),
    "h2amd_eval_polynomial_batch_device_workspace_size": (U64, [P, U32]),
    "h2amd_eval_polynomial_batch_device": (RustError, [P, P, P, U32, P, P, U64, P]),
    "_halo2_grand_product_max_len": (SZ, [SZ, U64]),
    "_halo2_grand_product_workspace_size": (U64, [U64]),
    "_halo2_grand_product": (RustError, [OBJP, OBJP, OBJP, SZ, SZ, P, U64, P]),
    "_halo2_grand_product_device_inputs": (RustError, [OBJP, OBJP, SZ, SZ, P]),
    "_halo2_quotient_lookups": (RustError, [P, P, P, P, P, P, P, P, P, P, P, SZ, P]),
    "_halo2_quotient_permutation": (RustError, [P, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                                U64, U64, U64, I32, I32, I32, U64, P]),
    "_halo2_evaluate_h_max_rows": (U64, [P] + [U64] * 12),
    "_halo2_quotient_workspace_size": (U64, [P] + [U64] * 11),
    "_halo2_quotient": (RustError, [OBJP, U64, OBJP, U64, OBJP, U64, FFITraitObject, U64, FFITraitObject, U64,
                                    FFITraitObject, P, U64, P, U64, FFITraitObject, U64, U64, U64, U64, P, U64, P]),
    "_halo2_quotient_device_columns": (RustError, [P, U64, P, U64, P, U64, FFITraitObject, U64, FFITraitObject, U64,
                                                   FFITraitObject, P, U64, P, U64, FFITraitObject, U64, U64, U64,
                                                   U64, P, U64, P]),
    "_halo2_quotient_device_columns_device_out": (RustError, [P, U64, P, U64, P, U64, FFITraitObject, U64,
                                                              FFITraitObject, U64, FFITraitObject, P, U64, P, U64,
                                                              P, U64, P, U64, U64, U64, P, U64, P]),
    "_halo2_generate_omega_powers_workspace_size": (U64, [U32]),
    "_halo2_generate_omega_powers": (RustError, [P, P, U32, P, U64, P]),
    "_halo2_power_of_omega_workspace_size": (U64, [U32]),
    "_halo2_power_of_omega": (RustError, [OBJP, OBJP, OBJP, U32, U32, P, U64, P]),
    "_halo2_generate_omega_lut_workspace_size": (U64, [U32]),
    "_halo2_generate_omega_lut": (RustError, [OBJP, OBJP, U32, P, U64, P]),
    "_halo2_generate_omegadelta_workspace_size": (U64, [U32, U32, U32, U32, U32]),
    "_halo2_generate_omegadelta": (RustError, [P, P, P, P, U32, U32, U32, U32, U32, P, U64, P]),
    "_halo2_eval_polynomial_workspace_size": (U64, [U64]),
    "_halo2_eval_polynomial": (RustError, [P, OBJP, ctypes.POINTER(P), U64, P, U64, P]),
    "_halo2_eval_poly_batch_max_len": (SZ, [SZ, SZ, U64]),
    "_halo2_eval_polynomial_batch_workspace_size": (U64, [U64, U64]),
    "_halo2_eval_polynomial_batch": (RustError, [OBJP, OBJP, OBJP, SZ, SZ, SZ, P, U64, P]),
    "_halo2_multiopen_poly_max_len": (SZ, [SZ, SZ, U64]),
    "_halo2_multiopen_poly_calculation_workspace_size": (U64, [U64, U64]),
    "_halo2_multiopen_poly_calculation": (RustError, [OBJP, OBJP, SZ, SZ, SZ, OBJP, OBJP, OBJP, P, U64, P]),
    "_halo2_poly_multiply_add": (RustError, [P, P, P, SZ, P]),
    "h2amd_poly_multiply_add_many": (RustError, [P, P, P, P, U32, U64, P]),
    "_halo2_poly_elementwise_multiply": (RustError, [P, P, P, U64, P]),
    "_halo2_poly_fill_one": (RustError, [P, U64, P]),
    "_halo2_poly_sub_short_inplace": (RustError, [P, P, U64, P]),
    "_halo2_poly_sub_short_out_of_place": (RustError, [P, P, P, U64, U64, P]),
    "_halo2_poly_sub_scalar_at_zero": (RustError, [P, P, P]),
    "_halo2_kate_division_workspace_size": (U64, [U64]),
    "_halo2_kate_division_device": (RustError, [P, P, P, U64, P, U64, P]),
    "_halo2_kate_division_device_padded": (RustError, [P, P, P, U64, U64, P, U64, P]),
    "_halo2_decode_assigned": (RustError, [P, P, P, U64, AssignedLayout, P]),
    "h2amd_msm_table_bytes": (U64, [U64]),
    "h2amd_msm_precompute_workspace_size": (U64, [U64]),
    "h2amd_msm_precompute": (RustError, [P, P, U64, P, U64, P]),
    "h2amd_multiexp_precomputed_workspace_size": (U64, [U64, U64]),
    "h2amd_multiexp_precomputed": (RustError, [P, P, U64, OBJP, U64, P, U64, P]),
    "h2amd_multiexp_batch_workspace_size": (U64, [P, U32]),
    "h2amd_g1_flags_workspace_size": (U64, []),
    "h2amd_g1_to_lagrange_workspace_size": (U64, [U32]),
    "h2amd_msm_set_widths": (RustError, [P, U32]),
    "h2amd_msm_width": (U32, [U64]),
    "h2amd_msm_table_width": (U32, [U64]),
    "h2amd_ntt_set_tiles": (RustError, [P, U32]),
    "h2amd_g1_to_lagrange": (RustError, [P, P, U32, P, P, U64, P]),
    "h2amd_g1_validate": (RustError, [P, U64, P, P, P, U64, P]),
    "h2amd_g1_compress": (RustError, [P, P, U64, P]),
    "h2amd_g1_decompress": (RustError, [P, P, U64, P, P, P, U64, P]),
    "h2amd_multiexp_batch": (RustError, [P, P, U32, P, P, P, U64, P]),
    "h2amd_multiexp_batch_precomputed_workspace_size": (U64, [P, U32, U64]),
    "h2amd_multiexp_batch_precomputed": (RustError, [P, P, U32, P, U64, P, P, U64, P]),
    "h2amd_random_fr": (RustError, [P, U64, U64, U64, P]),
    "h2amd_g1_mul_generator_workspace_size": (U64, [U64]),
    "h2amd_g1_mul_generator": (RustError, [P, P, U64, P, U64, P]),
    "h2amd_profile_enable": (None, [ctypes.c_int]),
    "h2amd_profile_reset": (None, []),
    "h2amd_profile_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_double)]),
    "h2amd_mad_peak_probe": (RustError, [P, U32, U32, ctypes.POINTER(ctypes.c_double), P]),
}


def profile_read(name: str) -> tuple[int, float]:
    """(launches, total milliseconds) recorded for a kernel name since the last reset."""
    lib = load()
    cnt = U64(0)
    ms = ctypes.c_double(0.0)
    lib.h2amd_profile_read(name.encode(), ctypes.byref(cnt), ctypes.byref(ms))
    return int(cnt.value), float(ms.value)


def load(path: str | None = None):
    """Load the library (no fallback: raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (share torch's HIP runtime)

    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run `make -C {PKG_ROOT}` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(err: RustError, what: str = "") -> None:
    if err.code != 0:
        msg = ""
        if err.message:
            msg = ctypes.string_at(err.message).decode(errors="replace")
            _libc.free(err.message)
        raise Halo2GpuError(err.code, f"{what}: {msg}" if what else msg)


def obj(ptr: int) -> FFITraitObject:
    return FFITraitObject(ctypes.c_void_p(ptr))


def obj_array(ptrs) -> ctypes.Array:
    arr = (FFITraitObject * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i].ptr = p
    return arr
