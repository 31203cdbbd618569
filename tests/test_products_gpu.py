"""GPU parity of batch inversion, the grand-product scan and the permutation /
lookup products (src/cuda/tests.rs:978-1567 of the reference), bit-exact against
the CPU oracle, through the C ABI."""
import ctypes

import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _dev_scalar(x_limbs):
    from halo2_amd import device as dv

    return dv.to_device(np.asarray(x_limbs, dtype=np.uint64).reshape(1, 4))


@pytest.mark.parametrize("n", [1, 7, 255, 2048, 2049, 4096 + 129, 65536 + 3, 1 << 20, 1 << 24, (1 << 24) + 2049])
def test_batch_invert_fr(lib, co, n):
    from halo2_amd import _lib, device as dv

    print(f"batch_invert n={n}", flush=True)
    a = rand_fr(co, n, 10 + n)
    a[:: max(1, n // 5)] = 0  # zeros stay zero
    d = dv.to_device(a)
    _lib.check(lib._halo2_batch_invert(dv.ptr(d), 0, n, dv.stream()))
    got = dv.to_host(d)
    if n <= 1 << 16:
        assert (got == co.batch_invert(0, a)).all()
    else:
        # size-independent: x * x^-1 == 1 on every non-zero row (oracle multiply), zeros preserved
        from halo2_amd import fields as F

        prod = co.field_mul(0, a, got)
        nz = a.any(axis=1)
        assert (prod[nz] == F.mont_limbs(1)).all()
        assert not got[~nz].any()


def test_batch_invert_fq(lib, co):
    from halo2_amd import _lib, device as dv, fields as F

    n = 3000
    vals = np.array([F.mont_limbs(int(x) % F.P_MOD, F.P_MOD) for x in
                     np.random.default_rng(3).integers(0, 2**62, n)], dtype=np.uint64)
    vals[10] = 0
    d = dv.to_device(vals)
    _lib.check(lib._halo2_batch_invert(dv.ptr(d), 1, n, dv.stream()))
    assert (dv.to_host(d) == co.batch_invert(1, vals)).all()
    e = lib._halo2_batch_invert(dv.ptr(d), 2, n, dv.stream())
    assert e.code != 0


@pytest.mark.parametrize("n", [1, 5, 2047, 2048, 2049, 4096 + 17, 1 << 16, (1 << 22) + 5])
def test_grand_product_device(lib, co, n):
    from halo2_amd import _lib, device as dv

    print(f"grand_product n={n}", flush=True)
    a = rand_fr(co, n, 20 + n)
    pre = rand_fr(co, 1, 99)
    exp = co.grand_product(a, pre[0])
    d = dv.to_device(a)
    dp = dv.to_device(pre)
    _lib.check(lib._halo2_grand_product_device_inputs(ctypes.byref(_lib.obj(dv.ptr(d))),
                                                      ctypes.byref(_lib.obj(dv.ptr(dp))), n, 0, dv.stream()))
    assert (dv.to_host(d) == exp).all()


def test_grand_product_chunked_carry_and_host(lib, co):
    """Multi-chunk with a device-carried prefix (tests.rs:1413-1467) and the host entry point."""
    from halo2_amd import _lib, device as dv

    n = 1 << 14
    a = rand_fr(co, n, 31)
    pre = rand_fr(co, 1, 32)
    exp = co.grand_product(a, pre[0])
    d = dv.to_device(a)
    chunk = n // 4
    prefix = dv.to_device(pre)
    for off in range(0, n, chunk):
        _lib.check(lib._halo2_grand_product_device_inputs(ctypes.byref(_lib.obj(dv.ptr(d))),
                                                          ctypes.byref(_lib.obj(dv.ptr(prefix))), chunk, off,
                                                          dv.stream()))
        prefix = d[4 * (off + chunk - 1): 4 * (off + chunk)].clone()
    assert (dv.to_host(d) == exp).all()
    # host arm with offsets, chunk by chunk
    out = np.zeros_like(a)
    ws = int(lib._halo2_grand_product_workspace_size(chunk))
    scratch = dv.empty_bytes(ws)
    p = pre[0].copy()
    for off in range(0, n, chunk):
        _lib.check(lib._halo2_grand_product(ctypes.byref(_lib.obj(out.ctypes.data)),
                                            ctypes.byref(_lib.obj(a.ctypes.data)),
                                            ctypes.byref(_lib.obj(p.ctypes.data)), chunk, off, dv.ptr(scratch), ws,
                                            dv.stream()))
        p = out[off + chunk - 1].copy()
    assert (out == exp).all()


def _perm_inputs(co, n, batch, seed):
    cols = [rand_fr(co, n, seed + i) for i in range(batch)]
    perms = [rand_fr(co, n, seed + 100 + i) for i in range(batch)]
    sc = rand_fr(co, 5, seed + 200)
    return cols, perms, sc


# up to 4 columns go through one launch: 5 and 9 columns take two and three, the delta
# power carried between them
@pytest.mark.parametrize("n,batch", [(1000, 1), (1 << 14, 3), ((1 << 18) + 1, 4), (4099, 5), (1 << 12, 9)])
def test_permutation_product_device(lib, co, n, batch):
    from halo2_amd import _lib, device as dv

    print(f"permutation_product n={n} batch={batch}", flush=True)
    cols, perms, sc = _perm_inputs(co, n, batch, 40)
    beta, gamma, delta, omega, dw = sc
    mod0 = rand_fr(co, n, 45)
    exp, dw_exp = co.permutation_product(mod0, perms, cols, beta, gamma, delta, omega, dw)
    d_mod = dv.to_device(mod0)
    d_cols = [dv.to_device(c) for c in cols]
    d_perms = [dv.to_device(p) for p in perms]
    d_sc = [_dev_scalar(x) for x in sc]
    ws = int(lib._halo2_permutation_product_device_inputs_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_permutation_product_device_inputs(
        ctypes.byref(_lib.obj(dv.ptr(d_mod))), _lib.obj_array([dv.ptr(x) for x in d_perms]),
        _lib.obj_array([dv.ptr(x) for x in d_cols]), *[dv.ptr(x) for x in d_sc], n, 0, batch, dv.ptr(scratch), ws,
        dv.stream()))
    assert (dv.to_host(d_mod) == exp).all()
    assert (dv.to_host(d_sc[4]) == dw_exp).all()


def test_permutation_product_host_offset(lib, co):
    from halo2_amd import _lib, device as dv

    n_full, off, n = 5000, 1200, 3000
    cols, perms, sc = _perm_inputs(co, n_full, 2, 60)
    beta, gamma, delta, omega, dw = sc
    mod0 = rand_fr(co, n, 66)
    exp, dw_exp = co.permutation_product(mod0, [p[off:off + n] for p in perms], [c[off:off + n] for c in cols],
                                         beta, gamma, delta, omega, dw)
    mod = mod0.copy()
    d_sc = [_dev_scalar(x) for x in sc]
    ws = int(lib._halo2_permutation_product_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_permutation_product(
        ctypes.byref(_lib.obj(mod.ctypes.data)), _lib.obj_array([p.ctypes.data for p in perms]),
        _lib.obj_array([c.ctypes.data for c in cols]), *[dv.ptr(x) for x in d_sc], n, off, 2, dv.ptr(scratch), ws,
        dv.stream()))
    assert (mod == exp).all()
    assert (dv.to_host(d_sc[4]) == dw_exp).all()


@pytest.mark.parametrize("n", [100, 1 << 16, (1 << 20) + 7])
def test_lookup_product(lib, co, n):
    from halo2_amd import _lib, device as dv

    print(f"lookup_product n={n}", flush=True)
    arrs = [rand_fr(co, n, 70 + i) for i in range(4)]
    sc = rand_fr(co, 2, 75)
    exp = co.lookup_product(*arrs, sc[0], sc[1])
    d = [dv.to_device(x) for x in arrs]
    out = dv.empty_u64(4 * n)
    db, dg = _dev_scalar(sc[0]), _dev_scalar(sc[1])
    objs = [_lib.obj(dv.ptr(x)) for x in [out] + d]
    _lib.check(lib._halo2_commit_product_device_inputs(*[ctypes.byref(o) for o in objs], dv.ptr(db), dv.ptr(dg), n,
                                                       0, dv.stream()))
    assert (dv.to_host(out) == exp).all()
    if n <= 1 << 16:
        # host arm at an offset inside longer host buffers
        off = 13
        full = [np.concatenate([rand_fr(co, off, 80 + i), x]) for i, x in enumerate(arrs)]
        lp = np.zeros((off + n, 4), dtype=np.uint64)
        ws = int(lib._halo2_commit_product_workspace_size(n))
        scratch = dv.empty_bytes(ws)
        objs = [_lib.obj(x.ctypes.data) for x in [lp] + full]
        _lib.check(lib._halo2_commit_product(*[ctypes.byref(o) for o in objs], dv.ptr(db), dv.ptr(dg), n, off,
                                             dv.ptr(scratch), ws, dv.stream()))
        assert (lp[off:] == exp).all() and not lp[:off].any()
