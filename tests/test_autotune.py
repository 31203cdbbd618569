"""The tuned MSM width and NTT tile tables (h2amd_msm_set_widths, h2amd_ntt_set_tiles) and
halo2_amd.autotune: every table entry gives the same results (MSM against the oracle's
closed form, NTT against the default plan), the width reaches the workspace sizing, bad
entries are refused, and the tuner returns its choices and restores the default tables."""
import numpy as np
import pytest

from conftest import rand_fr


def test_width_table_host_side():
    """Pure host: the table changes the width and the MSM workspace, and is validated."""
    import torch  # noqa: F401  one HIP runtime with torch

    from halo2_amd import _lib
    from halo2_amd.autotune import set_msm_widths

    lib = _lib.load()
    n = 1 << 18
    try:
        c0, c17 = int(lib.h2amd_msm_width(n)), int(lib.h2amd_msm_width(n // 2))
        ws0 = int(lib._halo2_multiexp_device_scalars_device_bases_workspace_size(n))
        set_msm_widths({18: c0 + 2})
        assert int(lib.h2amd_msm_width(n)) == c0 + 2 and int(lib.h2amd_msm_width(n + 5)) == c0 + 2
        assert int(lib.h2amd_msm_width(n // 2)) == c17  # other sizes keep the default rule
        assert int(lib._halo2_multiexp_device_scalars_device_bases_workspace_size(n)) != ws0
        for bad in (3, 21):
            with pytest.raises(_lib.Halo2GpuError):
                set_msm_widths({18: bad})
        t = np.zeros(65, dtype=np.uint8)
        with pytest.raises(_lib.Halo2GpuError):
            _lib.check(lib.h2amd_msm_set_widths(t.ctypes.data, 65))
        # precomputed tables: their own entries (up to 22), the table size follows the width
        tw0, tb0 = int(lib.h2amd_msm_table_width(n)), int(lib.h2amd_msm_table_bytes(n))
        set_msm_widths({}, {18: 16})
        assert int(lib.h2amd_msm_table_width(n)) == 16 and int(lib.h2amd_msm_width(n)) == c0
        assert int(lib.h2amd_msm_table_bytes(n)) == 64 * n * ((255 + 15) // 16) != tb0
        set_msm_widths({}, {18: 22})
        with pytest.raises(_lib.Halo2GpuError):
            set_msm_widths({}, {18: 23})
        set_msm_widths({})
        assert int(lib.h2amd_msm_table_width(n)) == tw0
        with pytest.raises(_lib.Halo2GpuError):
            _lib.check(lib.h2amd_ntt_set_tiles(np.full(5, 13, dtype=np.uint8).ctypes.data, 5))
    finally:
        set_msm_widths({})
    assert int(lib.h2amd_msm_width(n)) == c0


@pytest.mark.gpu
def test_msm_widths_same_result(lib, co):
    from halo2_amd import device as dv, msm
    from halo2_amd.autotune import set_msm_widths

    n = 5000
    s, c = rand_fr(co, n, 1), rand_fr(co, n, 2)
    pts = co.g1_mul_gen(s)
    exp = co.g1_mul(co.g1_generator(), co.fr_inner_product(c, s))
    ds, dp = dv.to_device(c), dv.to_device(pts)
    try:
        for w in (4, 7, 9, 13, 16, 20):
            set_msm_widths({12: w})
            assert int(lib.h2amd_msm_width(n)) == w
            assert (msm.multiexp_device_scalars_device_bases(ds, dp, n) == exp).all(), f"width {w}"
    finally:
        set_msm_widths({})


@pytest.mark.gpu
def test_table_widths_same_result(lib, co):
    """A fixed-base table built at any width gives the plain MSM's result."""
    from halo2_amd import device as dv, msm
    from halo2_amd.autotune import set_msm_widths

    n = 3000
    ds = dv.to_device(rand_fr(co, n, 41))
    dp = msm.g1_mul_generator(dv.to_device(rand_fr(co, n, 42)), n)
    exp = msm.multiexp_device_scalars_device_bases(ds, dp, n)
    try:
        for w in (6, 9, 11, 12, 14):
            set_msm_widths({}, {11: w})
            assert int(lib.h2amd_msm_table_width(n)) == w
            table = msm.msm_precompute(dp, n)
            assert (msm.multiexp_precomputed(ds, table, n, n) == exp).all(), f"table width {w}"
    finally:
        set_msm_widths({})


@pytest.mark.gpu
def test_ntt_tiles_same_result(lib, co):
    from halo2_amd import device as dv, fields as F, ntt
    from halo2_amd.autotune import set_ntt_tiles

    for lg in (10, 14, 17):
        a = dv.to_device(rand_fr(co, 1 << lg, 10 + lg))
        ref = dv.empty_u64(4 << lg)
        ntt.fft_normal_device(6, lg, a, ref, F.omega(lg), F.ZETA)
        try:
            for t in (6, 8, 10, 11, 12):
                set_ntt_tiles({lg: t})
                o = dv.empty_u64(4 << lg)
                ntt.fft_normal_device(6, lg, a, o, F.omega(lg), F.ZETA)
                assert bool((o == ref).all()), f"2^{lg} tile {t}"
        finally:
            set_ntt_tiles({})


@pytest.mark.gpu
def test_tuner_round_trip(lib, co, tmp_path):
    from halo2_amd import autotune

    res = autotune.tune(msm_logs=(12, 14), ntt_logs=(12, 14), table_logs=(12, 14), apply=True)
    try:
        for lg in (12, 14):
            m = res["msm"][lg]
            assert m["best"] in m["ms"] and m["default"] in m["ms"]
            assert int(lib.h2amd_msm_width(1 << lg)) == m["best"]
            t = res["msm_table"][lg]
            assert t["best"] in t["ms"] and t["default"] in t["ms"]
            assert int(lib.h2amd_msm_table_width(1 << lg)) == t["best"]
            assert res["ntt"][lg]["best"] in res["ntt"][lg]["ms"]
        path = str(tmp_path / "tuned.json")
        autotune.save(res, path)
        autotune.set_msm_widths({})
        back = autotune.load(path)
        assert int(lib.h2amd_msm_width(1 << 14)) == back["msm"]["14"]["best"]
    finally:
        autotune.set_msm_widths({})
        autotune.set_ntt_tiles({})
