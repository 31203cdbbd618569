"""CPU checks of the C ABI: the library loads, exports every symbol that
include/halo2_gpu.h declares, and the pure-host sizing entry points answer
without a GPU (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "halo2_gpu.h")
LIB = os.path.join(ROOT, "halo2-gpu_amd", "halo2_amd", "libhalo2_gpu.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(_halo2_\w+|h2amd_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def cdll():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (run __graft_entry__.build())")
    import torch  # noqa: F401  share the HIP runtime soname with torch

    return ctypes.CDLL(LIB)


def test_header_declares_reference_surface():
    names = declared_functions()
    for must in ("_halo2_multiexp", "_halo2_fft_normal_to_device", "_halo2_fft_many_to_device",
                 "_halo2_batch_invert", "_halo2_grand_product_device_inputs", "_halo2_quotient_lookups",
                 "_halo2_permutation_product_device_inputs", "_halo2_commit_product_device_inputs"):
        assert must in names


def test_every_declared_symbol_is_exported(cdll):
    missing = [n for n in declared_functions() if not hasattr(cdll, n)]
    assert not missing, f"declared in halo2_gpu.h but not exported: {missing}"


def test_python_signatures_cover_header():
    import sys

    sys.path.insert(0, os.path.join(ROOT, "halo2-gpu_amd"))
    from halo2_amd import _lib

    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_host_sizing_entry_points(cdll):
    from halo2_amd import _lib

    lib = _lib.load(LIB)
    for t in (1, 2, 6):
        assert lib._halo2_fft_normal_to_device_workspace_size(t, 20, 20) >= 32 << 20
    assert lib._halo2_fft_normal_workspace_size(3, 16, 18) >= 32 * ((1 << 16) + 2 * (1 << 18))
    assert lib._halo2_fft_many_to_device_workspace_size(6, 0, 0) == 0
    assert lib._halo2_fft_normal_check_memory(1, None, 10, 10)
    assert lib._halo2_fft_many_check_memory(1, 10, 10)
    assert lib._halo2_get_fft_split_radix(1, 20, 20, 1 << 40) == 0
    assert lib._halo2_get_fft_split_radix(1, 20, 20, 1) == -1
    assert lib._halo2_grand_product_max_len(1 << 20, 1 << 40) == 1 << 20
    assert lib._halo2_grand_product_max_len(1 << 20, 1024) == 16
    assert lib._halo2_commit_product_max_len(1 << 20, 0) == 0
    assert lib._halo2_permutation_product_workspace_size(1000) >= 4 * 32 * 1000


def test_opening_helper_sizing(cdll):
    """Host-only sizing of the opening-phase helpers (culib.rs:231-251, 319, 401-410, 430-475)."""
    from halo2_amd import _lib

    lib = _lib.load(LIB)
    ws = lib._halo2_eval_polynomial_batch_workspace_size(1 << 20, 8)
    assert ws >= 2 * 32 << 20  # two staged polynomials
    assert lib._halo2_eval_poly_batch_max_len(1 << 20, 8, ws + 1) == 1 << 20
    assert lib._halo2_eval_poly_batch_max_len(1 << 20, 8, ws // 2 + 4096) == 1 << 19
    assert lib._halo2_eval_poly_batch_max_len(1 << 20, 8, 0) == 0
    mws = lib._halo2_multiopen_poly_calculation_workspace_size(1 << 20, 8)
    assert mws >= 3 * 32 << 20  # two staged polynomials and the accumulator
    assert lib._halo2_multiopen_poly_max_len(1 << 20, 8, mws + 1) == 1 << 20
    assert lib._halo2_kate_division_workspace_size(0) == 0 and lib._halo2_kate_division_workspace_size(1) == 0
    assert 0 < lib._halo2_kate_division_workspace_size(1 << 20) < 32 << 20  # no full-length state copy
    assert lib._halo2_eval_polynomial_workspace_size(1 << 24) < 1 << 20
    assert lib._halo2_generate_omega_lut_workspace_size(20) >= 32 * ((1 << 10) + (1 << 10))
    assert lib._halo2_power_of_omega_workspace_size(31) >= 32 * 32
