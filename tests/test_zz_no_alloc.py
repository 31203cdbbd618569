"""The scratch-free MSM entry point (h2amd_multiexp_device_scalars_device_bases_noalloc)
through the C ABI against the workspace entry point it mirrors
(_halo2_multiexp_device_scalars_device_bases): two calls in a row, the second with a
smaller input, so stale contents of the shared built-in scratch would show; the
maximum length; lengths above it rejected with an error before anything runs."""
import ctypes

import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _noalloc(lib, d_scalars, d_bases, n):
    from halo2_amd import _lib, device as dv

    out = np.zeros(12, dtype=np.uint64)
    obj = _lib.obj(out.ctypes.data)
    err = lib.h2amd_multiexp_device_scalars_device_bases_noalloc(dv.ptr(d_scalars), dv.ptr(d_bases),
                                                                 ctypes.byref(obj), n, dv.stream())
    _lib.check(err, "h2amd_multiexp_device_scalars_device_bases_noalloc")
    return out


def test_noalloc_msm_matches_workspace_entry_point(lib, co):
    from halo2_amd import device as dv, msm

    nmax = int(lib.h2amd_multiexp_noalloc_max_length())
    assert nmax == 1 << 16
    ds = msm.random_fr(nmax, 501)
    pts = msm.g1_mul_generator(ds, nmax)
    for n, seed in ((nmax, 502), (1000, 503), (nmax, 504), (1, 505), (3, 506)):
        print(f"noalloc msm n={n}", flush=True)
        dc = dv.to_device(rand_fr(co, n, seed))
        got = _noalloc(lib, dc, pts, n)
        exp = msm.multiexp_device_scalars_device_bases(dc, pts[: 8 * n], n)
        assert (got == exp).all(), n
    # zero length: the identity, as the workspace entry point gives
    dc = dv.to_device(rand_fr(co, 1, 507))
    assert (_noalloc(lib, dc, pts, 0) == msm.multiexp_device_scalars_device_bases(dc, pts, 0)).all()


def test_noalloc_msm_rejects_oversize(lib):
    from halo2_amd import _lib, device as dv, msm

    nmax = int(lib.h2amd_multiexp_noalloc_max_length())
    d = dv.empty_u64(8)
    with pytest.raises(_lib.Halo2GpuError, match="maximum"):
        _noalloc(lib, d, d, nmax + 1)
    # the built-in scratch still works after the refusal
    ds = msm.random_fr(64, 508)
    pts = msm.g1_mul_generator(ds, 64)
    dc = msm.random_fr(64, 509)
    assert (_noalloc(lib, dc, pts, 64) == msm.multiexp_device_scalars_device_bases(dc, pts, 64)).all()
