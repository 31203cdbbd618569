"""GPU parity of the lookup and permutation quotient identities
(src/cuda/tests.rs:1570-1707 and :1895-2000 of the reference) against the CPU
restatements of quotient_lookups_cpu / permutation_quotient_cpu_chunk."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import rand_fr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64, 1 << 14, (1 << 18) + 3])
def test_quotient_lookups(lib, co, n):
    from halo2_amd import _lib, device as dv

    print(f"quotient_lookups n={n}", flush=True)
    arrs = [rand_fr(co, n, 100 + i) for i in range(8)]
    sc = rand_fr(co, 3, 108)
    exp = co.quotient_lookups(*arrs, sc[0], sc[1], sc[2])
    d = [dv.to_device(x) for x in arrs]
    ds = [dv.to_device(sc[i:i + 1]) for i in range(3)]
    _lib.check(lib._halo2_quotient_lookups(*[dv.ptr(x) for x in d], *[dv.ptr(x) for x in ds], n, dv.stream()))
    assert (dv.to_host(d[0]) == exp).all()


@pytest.mark.parametrize("n,n_cols,chunk_len", [(256, 3, 3), (4096, 7, 3), ((1 << 16) + 4, 5, 2), (1024, 1, 1)])
def test_quotient_permutation(lib, co, n, n_cols, chunk_len):
    from halo2_amd import _lib, device as dv

    print(f"quotient_permutation n={n} cols={n_cols} chunk={chunk_len}", flush=True)
    n_sets = (n_cols + chunk_len - 1) // chunk_len
    zs = [rand_fr(co, n, 200 + i) for i in range(n_sets)]
    perms = [rand_fr(co, n, 300 + i) for i in range(n_cols)]
    cols = [rand_fr(co, n, 400 + i) for i in range(n_cols)]
    values, l0, ll, la = (rand_fr(co, n, 500 + i) for i in range(4))
    sc = rand_fr(co, 7, 600)
    beta, gamma, y, delta, delta_start, cur, omega = sc
    last_rotation, rot_scale, isize = -3, 4, n
    exp = co.permutation_quotient(values, zs, perms, cols, l0, ll, la, beta, gamma, y, delta, delta_start, cur,
                                  omega, chunk_len, last_rotation, rot_scale, isize)
    d_vals = dv.to_device(values)
    dz, dp, dc = ([dv.to_device(x) for x in group] for group in (zs, perms, cols))
    tables = [torch.tensor([dv.ptr(x) for x in group], dtype=torch.int64, device=dv.device()) for group in (dz, dp, dc)]
    dl = [dv.to_device(x) for x in (l0, ll, la)]
    host_sc = [np.ascontiguousarray(x) for x in sc]
    _lib.check(lib._halo2_quotient_permutation(
        dv.ptr(d_vals), *[dv.ptr(x) for x in dl], *[dv.ptr(t) for t in tables],
        *[ctypes.c_void_p(x.ctypes.data) for x in host_sc], n_sets, chunk_len, n_cols, last_rotation, rot_scale,
        isize, n, dv.stream()))
    assert (dv.to_host(d_vals) == exp).all()
