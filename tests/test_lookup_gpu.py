"""Lookup permutation pair and lookup product on the device against the oracle
(src/cuda/tests.rs:978-1141 and 2340-2548 of the reference): all-distinct,
mixed-duplicate, all-same and small-value (keys equal in the high words) inputs,
usable_rows = n - blinding - 1 with garbage in the blinding rows, sizes across
the 2048-key sort tile and many merge passes; bit-exact."""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _case(co, U, pattern, seed):
    from halo2_amd.fields import mont_limbs

    rng = np.random.default_rng(seed)
    if U == 0:
        z = np.zeros((0, 4), dtype=np.uint64)
        return z, z
    if pattern == "small":
        alpha = min(U, 1000)
        alphabet = np.stack([mont_limbs(v) for v in range(alpha)])
    else:
        alpha = {"distinct": U, "mixed": max(1, U // 3), "same": 1}[pattern]
        alphabet = rand_fr(co, alpha, seed)
    table = np.concatenate([alphabet, alphabet[rng.integers(0, alpha, U - alpha)]])
    table = table[rng.permutation(U)]
    if pattern == "distinct":
        inp = alphabet[rng.permutation(U)]
    else:
        inp = alphabet[rng.integers(0, alpha, U)]
    return np.ascontiguousarray(inp), np.ascontiguousarray(table)


CASES = [(4, 5, "distinct"), (4, 0, "same"), (10, 5, "mixed"), (11, 5, "small"), (12, 0, "distinct"),
         (12, 3, "mixed"), (13, 7, "same"), (16, 5, "mixed"), (16, 5, "small"), (20, 5, "distinct"),
         (20, 5, "mixed")]


@pytest.mark.parametrize("log_n,blinding,pattern", CASES)
def test_permute_expression_pair(lib, co, log_n, blinding, pattern):
    from halo2_amd import device as dv
    from halo2_amd.lookup import permute_expression_pair_device

    n = 1 << log_n
    U = n - blinding - 1 if blinding else n
    print(f"permute n={n} U={U} {pattern}", flush=True)
    inp, table = _case(co, U, pattern, 31 * log_n + blinding)
    pad_i = np.concatenate([inp, rand_fr(co, n - U, 7)])
    pad_t = np.concatenate([table, rand_fr(co, n - U, 8)])
    exp_i, exp_t = co.permute_expression_pair(pad_i, pad_t, n, U)
    d_pi = dv.to_device(np.full((n, 4), 0x5A, dtype=np.uint64))
    d_pt = dv.to_device(np.full((n, 4), 0x5A, dtype=np.uint64))
    permute_expression_pair_device(dv.to_device(pad_i), dv.to_device(pad_t), d_pi, d_pt, n, U)
    assert (dv.to_host(d_pi) == exp_i).all()
    assert (dv.to_host(d_pt) == exp_t).all()


def test_permute_expression_pair_edges(lib, co):
    """usable_rows 0 and 1 (outputs all zero / one row), and usable_rows > n rejected."""
    from halo2_amd import _lib, device as dv
    from halo2_amd.lookup import permute_expression_pair_device

    n = 64
    a, t = rand_fr(co, n, 1), rand_fr(co, n, 2)
    t[0] = a[0]
    for U in (0, 1):
        exp_i, exp_t = co.permute_expression_pair(a, t, n, U)
        d_pi, d_pt = dv.to_device(np.ones((n, 4), dtype=np.uint64)), dv.to_device(np.ones((n, 4), dtype=np.uint64))
        permute_expression_pair_device(dv.to_device(a), dv.to_device(t), d_pi, d_pt, n, U)
        assert (dv.to_host(d_pi) == exp_i).all() and (dv.to_host(d_pt) == exp_t).all()
    d = dv.to_device(a)
    err = lib._halo2_permute_expression_pair(dv.ptr(d), dv.ptr(d), dv.ptr(d), dv.ptr(d), 4, 5, None, 0, dv.stream())
    with pytest.raises(_lib.Halo2GpuError):
        _lib.check(err)


@pytest.mark.parametrize("free_bytes", [None, 200_000])
def test_lookup_product_mirrors(lib, co, free_bytes):
    from halo2_amd import device as dv, fields as F
    from halo2_amd.lookup import lookup_product_device, lookup_product_gpu

    n = 9000
    pi, pt, ci, ct = (rand_fr(co, n, 40 + i) for i in range(4))
    beta, gamma = 0xB37A, 0x6A33A
    exp = co.lookup_product(pi, pt, ci, ct, F.mont_limbs(beta), F.mont_limbs(gamma))
    assert (lookup_product_gpu(pi, pt, ci, ct, beta, gamma, free_bytes=free_bytes) == exp).all()
    got = lookup_product_device(*(dv.to_device(x) for x in (pi, pt, ci, ct)), beta, gamma)
    assert (dv.to_host(got) == exp).all()
