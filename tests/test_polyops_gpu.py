"""Opening-phase and key-generation helpers on the device against the oracle, on the
reference's own cases (src/cuda/tests.rs:53-68, 632-686, 2002-2037, 2830-3100;
src/poly/domain.rs:1227-1350; tests/eval_polynomial_horner_device_equivalence.rs,
tests/batch_eval_polynomial_device_out_equivalence.rs,
tests/poly_multiply_add_device_equivalence.rs): evaluation at a point (single,
batched, chunked), the multiopen fold, multiply-add / scale / elementwise / fill,
short subtracts, division by (X - u) (plain, chained, padded, truncated), Assigned
decoding, and the omega tables. Every comparison is bit-exact."""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _fr(co, x):
    from halo2_amd.fields import mont_limbs

    return mont_limbs(x)


def _rand_int(seed):
    import random

    from halo2_amd.fields import R_MOD

    return random.Random(seed).randrange(R_MOD)


EVAL_LENGTHS = [1, 2, 63, 255, 256, 4095, 4096, 4097, 16383, 16384, 16385, 1 << 16, (1 << 20) + 3,
                (1 << 22) + 4096 * 256 + 1]


@pytest.mark.parametrize("n", EVAL_LENGTHS)
def test_eval_polynomial_device(lib, co, n):
    from halo2_amd import device as dv
    from halo2_amd.polyops import eval_polynomial_device

    print(f"eval n={n}", flush=True)
    a = rand_fr(co, n, 100 + n % 9973)
    d = dv.to_device(a)
    for u in (0, 1, _rand_int(n)):
        assert (eval_polynomial_device(d, u) == co.eval_poly(a, _fr(co, u))).all(), u


def test_eval_polynomial_batch_device_out(lib, co):
    from halo2_amd import device as dv
    from halo2_amd.polyops import batch_eval_polynomial_d2h

    lens = [1, 300, 4096, 70000, 5]
    polys = [rand_fr(co, n, 200 + i) for i, n in enumerate(lens)]
    pts = [_rand_int(300 + i) for i in range(len(lens))]
    got = batch_eval_polynomial_d2h([dv.to_device(p) for p in polys], pts)
    exp = np.stack([co.eval_poly(p, _fr(co, x)) for p, x in zip(polys, pts)])
    assert (got == exp).all()


def test_eval_polynomial_batch_device_uneven(lib, co):
    """One batched launch sequence over 41 polynomials (more than one launch's 32) of
    uneven lengths, empty ones included (p(u) = 0), points 0, 1 and random."""
    import torch

    from halo2_amd import device as dv
    from halo2_amd.polyops import batch_eval_polynomial_d2h

    base = [0, 1, 2, 255, 256, 257, 4095, 4096, 4097, 8193, 1 << 20, 300000, 0, 7]
    lens = (base * 3)[:41]
    polys = [rand_fr(co, n, 900 + i) if n else np.zeros((0, 4), np.uint64) for i, n in enumerate(lens)]
    pts = [[0, 1][i] if i < 2 else _rand_int(950 + i) for i in range(len(lens))]
    d = [dv.to_device(p) if len(p) else torch.zeros(0, dtype=torch.int64, device=dv.device()) for p in polys]
    got = batch_eval_polynomial_d2h(d, pts)
    for i, (p, x) in enumerate(zip(polys, pts)):
        exp = co.eval_poly(p, _fr(co, x)) if len(p) else np.zeros(4, np.uint64)
        assert (got[i] == exp).all(), f"poly {i} n={lens[i]}"


@pytest.mark.parametrize("free_bytes", [None, 3 * 32 * 1000])
def test_batch_eval_polynomial_gpu(lib, co, free_bytes):
    """Host polys; the small free_bytes forces the chunked path (chunks scaled by point^offset)."""
    from halo2_amd.polyops import basic_batch_eval_polynomial_gpu, batch_eval_polynomial_gpu

    n = 5000
    polys = [rand_fr(co, n, 400 + i) for i in range(4)]
    pts = [0, 1, _rand_int(1), _rand_int(2)]
    got = batch_eval_polynomial_gpu(polys, pts, n, free_bytes=free_bytes)
    exp = np.stack([co.eval_poly(p, _fr(co, x)) for p, x in zip(polys, pts)])
    assert (got == exp).all()
    part = basic_batch_eval_polynomial_gpu(polys, pts, 1000, 777)
    exp = np.stack([co.eval_poly(p[1000:1777], _fr(co, x)) for p, x in zip(polys, pts)])
    assert (part == exp).all()


@pytest.mark.parametrize("free_bytes", [None, 5 * 32 * 1024])
def test_multiopen_poly_calculation(lib, co, free_bytes):
    from halo2_amd.polyops import multiopen_poly_calculation_gpu

    n, batch = 3000, 5
    polys = [rand_fr(co, n, 500 + i) for i in range(batch)]
    ch = [_rand_int(600 + i) for i in range(batch)]
    ep = [_rand_int(700 + i) for i in range(batch)]
    acc = np.full((n, 4), 7, dtype=np.uint64)  # overwritten
    res = multiopen_poly_calculation_gpu(polys, ch, acc, ep, n, free_bytes=free_bytes)
    exp_acc = np.zeros((n, 4), dtype=np.uint64)
    for p, c in zip(polys, ch):
        exp_acc = co.field_add(0, exp_acc, co.field_mul(0, p, np.tile(_fr(co, c), (n, 1))))
    assert (acc == exp_acc).all()
    exp = np.stack([co.eval_poly(p, _fr(co, x)) for p, x in zip(polys, ep)])
    assert (res == exp).all()


def test_multiply_add_scale_elementwise_fill(lib, co):
    from halo2_amd import device as dv
    from halo2_amd.fields import R_MOD
    from halo2_amd.polyops import (poly_elementwise_multiply_device, poly_fill_one_device,
                                   poly_multiply_add_device, poly_multiply_add_device_at_lut_offset,
                                   poly_multiply_add_single_gpu, poly_scale_device_with_d_s_minus_one)

    n = (1 << 16) + 5
    acc, x, lut = rand_fr(co, n, 801), rand_fr(co, n, 802), rand_fr(co, 9, 803)
    s = _rand_int(804)
    tile = lambda v: np.tile(v, (n, 1))  # noqa: E731
    exp = co.field_add(0, acc, co.field_mul(0, x, tile(_fr(co, s))))
    d_acc = dv.to_device(acc)
    poly_multiply_add_device(d_acc, dv.to_device(x), s)
    assert (dv.to_host(d_acc) == exp).all()
    assert (poly_multiply_add_single_gpu(acc, x, s) == exp).all()
    d_acc = dv.to_device(acc)
    poly_multiply_add_device_at_lut_offset(d_acc, dv.to_device(x), dv.to_device(lut), 6)
    assert (dv.to_host(d_acc) == co.field_add(0, acc, co.field_mul(0, x, tile(lut[6])))).all()
    d_buf = dv.to_device(acc)  # in-place scale: acc += (s - 1) acc
    poly_scale_device_with_d_s_minus_one(d_buf, dv.to_device(_fr(co, (s - 1) % R_MOD).reshape(1, 4)))
    assert (dv.to_host(d_buf) == co.field_mul(0, acc, tile(_fr(co, s)))).all()
    d_a, d_b = dv.to_device(acc), dv.to_device(x)
    d_out = dv.empty_u64(4 * n)
    poly_elementwise_multiply_device(d_out, d_a, d_b)
    assert (dv.to_host(d_out) == co.field_mul(0, acc, x)).all()
    poly_elementwise_multiply_device(d_a, d_a, d_b)  # aliasing the output with an input
    assert (dv.to_host(d_a) == co.field_mul(0, acc, x)).all()
    poly_fill_one_device(d_out)
    assert (dv.to_host(d_out) == tile(_fr(co, 1))).all()


@pytest.mark.parametrize("count", [1, 2, 3, 16, 17, 33])
def test_multiply_add_many(lib, co, count):
    """h2amd_poly_multiply_add_many (the SHPLONK folds in one pass per 16 polynomials)
    against the oracle's acc + sum_j s_j p_j, with polynomials shorter than the
    accumulator counted as zero-padded (lengths n, n - 1, n / 2, 1, ...)."""
    from halo2_amd import device as dv
    from halo2_amd.polyops import poly_multiply_add_many_device

    print(f"multiply_add_many count={count}", flush=True)
    n = (1 << 14) + 3
    lens = [[n, n - 1, n // 2, 1, 4097][j % 5] for j in range(count)]
    acc = rand_fr(co, n, 900 + count)
    polys = [rand_fr(co, lens[j], 910 + 7 * count + j) for j in range(count)]
    scal = rand_fr(co, count, 990 + count)
    exp = acc.copy()
    for j in range(count):
        term = co.field_mul(0, polys[j], np.tile(scal[j], (lens[j], 1)))
        exp[:lens[j]] = co.field_add(0, exp[:lens[j]], term)
    d_acc = dv.to_device(acc)
    poly_multiply_add_many_device(d_acc, [dv.to_device(p) for p in polys], dv.to_device(scal))
    assert (dv.to_host(d_acc) == exp).all()


def test_sub_short_and_scalar_at_zero(lib, co):
    from halo2_amd import device as dv
    from halo2_amd.polyops import (poly_sub_scalar_at_zero_device, poly_sub_short_in_place_device,
                                   poly_sub_short_out_of_place_device)

    n = 1 << 14
    for short_len in (0, 1, 5, 256, n):
        acc = rand_fr(co, n, 900 + short_len)
        short = rand_fr(co, short_len, 950 + short_len)
        exp = acc.copy()
        if short_len:
            exp[:short_len] = co.field_sub(0, acc[:short_len], short)
        d_short = dv.to_device(short) if short_len else None
        d_out = dv.empty_u64(4 * n)
        poly_sub_short_out_of_place_device(d_out, dv.to_device(acc), d_short)
        assert (dv.to_host(d_out) == exp).all(), short_len
        if short_len:
            d_acc = dv.to_device(acc)
            poly_sub_short_in_place_device(d_acc, d_short)
            assert (dv.to_host(d_acc) == exp).all(), short_len
    a = rand_fr(co, n, 990)
    s = rand_fr(co, 1, 991)
    exp = a.copy()
    exp[0] = co.field_sub(0, a[:1], s)[0]
    d = dv.to_device(a)
    poly_sub_scalar_at_zero_device(d, dv.to_device(s))
    assert (dv.to_host(d) == exp).all()


# one tile is 2048 outputs: 2049 / 2050 are one tile and a tile plus one output; 2^19 + 2
# needs 257 tiles (two per thread of the carries pass), 2^22 + 3 needs 2049 (sixteen)
KATE_LENGTHS = [2, 3, 32, 64, 65, 1024, 2049, 2050, 4097, 1 << 16, (1 << 19) + 2, 1 << 20, (1 << 22) + 3]


@pytest.mark.parametrize("n", KATE_LENGTHS)
def test_kate_division_device(lib, co, n):
    from halo2_amd import device as dv
    from halo2_amd.polyops import kate_division_device

    print(f"kate n={n}", flush=True)
    a = rand_fr(co, n, 1100 + n % 9973)
    d = dv.to_device(a)
    for u in (1, 0, _rand_int(n), _rand_int(n + 1)):
        q = kate_division_device(d, u)
        assert (dv.to_host(q) == co.kate_division(a, _fr(co, u))).all(), u


def test_kate_division_chained_and_padded(lib, co):
    """Three chained divisions (the SHPLONK div_by_vanishing pattern), the padded variant
    with pad 0 / 1 / 256, a truncated output and the n = 1 edge cases."""
    from halo2_amd import _lib, device as dv
    from halo2_amd.polyops import (kate_division_device, kate_division_device_padded_with_d_root,
                                   kate_division_device_with_d_root)

    n = 1 << 18
    a = rand_fr(co, n, 1200)
    roots = [_rand_int(1201), _rand_int(1202), _rand_int(1203)]
    exp = a
    d = dv.to_device(a)
    for r in roots:
        exp = co.kate_division(exp, _fr(co, r))
        d = kate_division_device(d, r)
    assert (dv.to_host(d) == exp).all()
    for m in (32, 1024, 1 << 16):
        a = rand_fr(co, m, 1300 + m)
        r = _rand_int(m)
        d_root = dv.to_device(_fr(co, r).reshape(1, 4))
        q = co.kate_division(a, _fr(co, r))
        for pad in (0, 1, 256):
            got = dv.to_host(kate_division_device_padded_with_d_root(dv.to_device(a), d_root, m - 1 + pad))
            assert (got[:m - 1] == q).all() and not got[m - 1:].any(), (m, pad)
        # out_len < n - 1 at the C boundary keeps the first out_len coefficients
        lib_ = _lib.load()
        cut = m // 3
        ws = int(lib_._halo2_kate_division_workspace_size(m))
        scratch = dv.empty_bytes(ws)
        out = dv.to_device(np.full((cut + 4, 4), 0xAB, dtype=np.uint64))
        _lib.check(lib_._halo2_kate_division_device_padded(dv.ptr(dv.to_device(a)), dv.ptr(out), dv.ptr(d_root), m,
                                                           cut, dv.ptr(scratch), ws, dv.stream()))
        got = dv.to_host(out)
        assert (got[:cut] == q[:cut]).all() and (got[cut:] == 0xAB).all()
    one = dv.to_device(rand_fr(co, 1, 1400))
    d_root = dv.to_device(_fr(co, 5).reshape(1, 4))
    for out_len in (1, 4, 1024):
        assert not dv.to_host(kate_division_device_padded_with_d_root(one, d_root, out_len)).any()
    assert kate_division_device(one, 5).numel() == 0
    assert kate_division_device_with_d_root(one, d_root).numel() == 0


def test_decode_assigned(lib, co):
    """Assigned<F> bytes as Rust lays them out (#[repr(C, u8)]: tag at 0, payloads at 8
    and 40, stride 72): Zero -> (0, 1), Trivial(x) -> (x, 1), Rational(n, d) -> (n, d);
    also the denominator-slice variant."""
    from halo2_amd import _lib, device as dv
    from halo2_amd.polyops import decode_assigned_into_denom_slice_device, decode_assigned_to_num_denom_device

    n = 5000
    rng = np.random.default_rng(7)
    tags = rng.integers(0, 3, n).astype(np.uint8)
    num, den = rand_fr(co, n, 1500), rand_fr(co, n, 1501)
    raw = rng.integers(0, 256, (n, 72)).astype(np.uint8)  # padding bytes hold garbage
    raw[:, 0] = tags
    raw[:, 8:40] = num.view(np.uint8).reshape(n, 32)
    raw[:, 40:72] = den.view(np.uint8).reshape(n, 32)
    one = np.tile(_fr(co, 1), (n, 1))
    exp_num = np.where((tags == 0)[:, None], 0, num).astype(np.uint64)
    exp_den = np.where((tags == 2)[:, None], den, one).astype(np.uint64)
    layout = _lib.AssignedLayout(72, 8, 40)
    d_num, d_den = decode_assigned_to_num_denom_device(raw, n, layout)
    assert (dv.to_host(d_num) == exp_num).all() and (dv.to_host(d_den) == exp_den).all()
    concat = dv.to_device(np.zeros((n + 100, 4), dtype=np.uint64))
    d_num2 = decode_assigned_into_denom_slice_device(raw, n, layout, concat, 60)
    host = dv.to_host(concat)
    assert (dv.to_host(d_num2) == exp_num).all() and (host[60:60 + n] == exp_den).all()
    assert not host[:60].any() and not host[60 + n:].any()
    with pytest.raises(_lib.Halo2GpuError):
        decode_assigned_to_num_denom_device(raw, n, _lib.AssignedLayout(72, 8, 48))


def test_omega_tables(lib, co):
    """power_of_omega (domain.rs:1227-1253), the dense/sparse omega table (domain.rs:1318-1350,
    generate_omega_lut_cpu in src/cpu/arithmetic.rs:199-220), omega powers (tests.rs:70-108)
    and the omega * delta mapping, all against Python big-integer powers."""
    from halo2_amd import _lib, device as dv
    from halo2_amd.fields import DELTA, R_MOD, mont_limbs, omega
    from halo2_amd.polyops import (generate_omega_lut_gpu, generate_omega_powers_gpu, generate_omegadelta_device,
                                   power_of_omega_gpu)

    M = lambda vals: np.stack([mont_limbs(v % R_MOD) for v in vals])  # noqa: E731
    for log_n in (1, 10, 17, 27):
        w = omega(log_n)
        for p in (0, 1, (1 << log_n) - 1, 12345 % (1 << log_n)):
            res, lut = power_of_omega_gpu(w, log_n, p)
            assert (res == mont_limbs(pow(w, p, R_MOD))).all()
            assert (lut == M([1] + [pow(w, 1 << i, R_MOD) for i in range(log_n)])).all()
    with pytest.raises(_lib.Halo2GpuError):
        power_of_omega_gpu(omega(5), 5, 32)
    with pytest.raises(_lib.Halo2GpuError):
        power_of_omega_gpu(omega(5), 32, 1)
    for log_n in (10, 14, 20):
        w = omega(log_n)
        lut = generate_omega_lut_gpu(w, log_n)
        w10 = pow(w, 1 << 10, R_MOD)
        exp = M([pow(w, i, R_MOD) for i in range(1 << 10)] + [pow(w10, i, R_MOD) for i in range(1 << (log_n - 10))])
        assert (lut == exp).all(), log_n
    with pytest.raises(_lib.Halo2GpuError):
        generate_omega_lut_gpu(omega(9), 9)
    w = omega(12)
    assert (generate_omega_powers_gpu(w, 12, 3000) == M([pow(w, i, R_MOD) for i in range(3000)])).all()
    log_n, cols = 13, 5
    rng = np.random.default_rng(3)
    rows = 1 << log_n
    mapping = np.stack([rng.integers(0, cols, rows), rng.integers(0, rows, rows)], axis=1).astype(np.uint32)
    w = omega(log_n)
    start, end = 100, rows - 7
    got = dv.to_host(generate_omegadelta_device(mapping, w, DELTA, log_n, start, end, cols, 3))
    exp = M([pow(w, int(mapping[i, 1]), R_MOD) * pow(DELTA, int(mapping[i, 0]), R_MOD) for i in range(start, end + 1)])
    assert (got == exp).all()
