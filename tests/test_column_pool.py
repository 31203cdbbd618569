"""ColumnPool and compress_expressions (the lookup-compress use of the quotient rule
interpreter), mirroring src/cuda/tests.rs:2035-2330 of the reference:
device arm vs a direct evaluation folded by theta, host-output vs device-output
entry points, borrowed device columns, and the all-at-once memory gate."""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _expressions():
    from halo2_amd.quotient import Advice, Fixed, Instance, Negated, Product, Scaled, Sum

    f0, f1, a0, a1, i0 = Fixed(0), Fixed(1), Advice(0), Advice(1), Instance(0)
    return [Sum(f0, a0), Sum(Product(f1, a1), Negated(i0)), Scaled(a0, 7)]


def _direct(cols, theta):
    """sum_e acc * theta + e over the three expressions, from canonical values."""
    from halo2_amd import fields as F

    R = F.R_MOD
    v = {k: [F.value_of(x) for x in c] for k, c in cols.items()}
    out = []
    for r in range(len(v["f0"])):
        es = [(v["f0"][r] + v["a0"][r]) % R, (v["f1"][r] * v["a1"][r] - v["i0"][r]) % R, 7 * v["a0"][r] % R]
        acc = 0
        for e in es:
            acc = (acc * theta + e) % R
        out.append(acc)
    return out


@pytest.mark.parametrize("k", [10, 16])
def test_compress_expressions_vs_direct(lib, co, k):
    import torch

    from halo2_amd import device as dv, fields as F
    from halo2_amd.quotient import ColumnPool, compress_expressions_device, compress_expressions_in_place_device, \
        for_compress

    n = 1 << k
    print(f"compress expressions k={k}", flush=True)
    cols = {"f0": rand_fr(co, n, 10), "f1": rand_fr(co, n, 11), "a0": rand_fr(co, n, 12), "a1": rand_fr(co, n, 13),
            "i0": rand_fr(co, n, 14)}
    theta = 0x1234567890ABCDEF1234567890ABCDEF % F.R_MOD
    pool = ColumnPool(n)
    assert not pool.is_initialized()
    pool.try_init([cols["f0"], cols["f1"]], [cols["a0"], cols["a1"]], [cols["i0"]])
    assert pool.is_initialized() and (pool.num_fixed(), pool.num_advice(), pool.num_instance()) == (2, 2, 1)
    host = compress_expressions_device(_expressions(), theta, n, 1, pool, [])
    rows = range(n) if k <= 10 else np.random.default_rng(k).choice(n, 512, replace=False)
    if k <= 10:
        assert [F.value_of(x) for x in host] == _direct(cols, theta)
    else:
        sub = {c: v[rows] for c, v in cols.items()}
        assert [F.value_of(x) for x in host[rows]] == _direct(sub, theta)
    # device output through a pre-built graph and explicit expression constants
    d_out = dv.empty_u64(4 * n)
    compress_expressions_in_place_device(for_compress(_expressions()), [0, 1, F.R_MOD - 1, 2, theta], n, 1, pool,
                                         [], d_out)
    assert (dv.to_host(d_out) == host).all()
    # borrowed device columns give the same bytes
    d = {c: dv.to_device(v) for c, v in cols.items()}
    pool2 = ColumnPool(n)
    pool2.try_init_device([d["f0"], d["f1"]], [], [d["a0"], d["a1"]], [d["i0"]])
    assert (compress_expressions_device(_expressions(), theta, n, 1, pool2, []) == host).all()
    torch.cuda.synchronize()


def test_column_pool_memory_gate(lib):
    """The gate runs before any upload: empty inputs pass; an absurd request raises
    InsufficientGpuMemory and leaves the pool uninitialised (tests.rs:2279-2330)."""
    from halo2_amd.quotient import ColumnPool, InsufficientGpuMemory

    absurd = 1 << 56
    pool = ColumnPool(absurd)
    pool.try_init([], [], [])
    assert pool.is_initialized()
    assert ColumnPool.estimate_resident_bytes(absurd, 64, 0, 0) == 32 * absurd * 64
    oom = ColumnPool(absurd)
    dummy = np.zeros((0, 4), dtype=np.uint64)
    with pytest.raises(InsufficientGpuMemory) as ei:
        oom.try_init([dummy] * 64, [], [])
    assert ei.value.magnitude == 64
    assert not oom.is_initialized()
