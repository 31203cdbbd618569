"""GPU parity of the NTT family against the CPU oracle (bit-exact), through the C ABI.

Mirrors the reference's FFT tests (src/cuda/tests.rs:110-387,
src/poly/domain.rs:1363-1866): forward/inverse/coset/coset-part transforms on
seeded inputs compared element by element with the oracle, in-place and
out-of-place, single and batched, host and device residency, plus the
size-independent round-trip and linearity properties at 2^24.
"""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _oracle_kind(co, po, kind, a, k, ext_k=None, c=None):
    """Oracle result of one NTT kind (composed from the oracle's FFT and twists)."""
    from halo2_amd import fields as F

    if kind == 1:
        return co.fft(a, k, F.mont_limbs(F.omega(k)))
    if kind == 2:
        out = co.fft(a, k, F.mont_limbs(F.inv(F.omega(k))))
        return co.scale_inplace(out, F.mont_limbs(F.inv(1 << k)))
    if kind == 6:
        t = co.distribute_powers_inplace(np.array(a, copy=True), F.mont_limbs(c))
        return co.fft(t, k, F.mont_limbs(F.omega(k)))
    if kind == 3:
        n_ext = 1 << ext_k
        pad = np.zeros((n_ext, 4), dtype=np.uint64)
        pad[: len(a)] = a
        z = F.mont_limbs(F.ZETA)
        pad = co.distribute_powers_zeta(pad, np.stack([z, F.mont_limbs(F.ZETA * F.ZETA % F.R_MOD)]))
        return co.fft(pad, ext_k, F.mont_limbs(F.omega(ext_k)))
    raise ValueError(kind)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 7, 8, 11, 12, 13, 14, 16, 18])
def test_fft_normal_to_device_kinds(lib, co, po, k):
    from halo2_amd import device as dv, fields as F, ntt

    n = 1 << k
    print(f"fft_normal_to_device k={k}", flush=True)
    a = rand_fr(co, n, 1000 + k)
    c = F.ZETA * pow(F.omega(k + 2), 3, F.R_MOD) % F.R_MOD
    for kind, omega, divisor in ((1, F.omega(k), 1), (2, F.inv(F.omega(k)), F.inv(n)), (6, F.omega(k), c)):
        exp = _oracle_kind(co, po, kind, a, k, c=c)
        d_in = dv.to_device(a)
        d_out = dv.empty_u64(4 * n)
        ntt.fft_normal_device(kind, k, d_in, d_out, omega, divisor)
        assert (dv.to_host(d_out) == exp).all(), f"kind {kind} out-of-place k={k}"
        assert (dv.to_host(d_in) == a).all(), "input must not be modified"
        ntt.fft_normal_device(kind, k, d_in, d_in, omega, divisor)  # in place
        assert (dv.to_host(d_in) == exp).all(), f"kind {kind} in-place k={k}"


@pytest.mark.parametrize("k", [20, 22])
def test_fft_large_vs_oracle(lib, co, po, k):
    from halo2_amd import device as dv, fields as F, ntt

    n = 1 << k
    print(f"fft large k={k}", flush=True)
    a = rand_fr(co, n, 2000 + k)
    c = F.ZETA
    exp = _oracle_kind(co, po, 6, a, k, c=c)
    d_in = dv.to_device(a)
    d_out = dv.empty_u64(4 * n)
    ntt.fft_normal_device(6, k, d_in, d_out, F.omega(k), c)
    assert (dv.to_host(d_out) == exp).all()


def test_fft_three_pass_size(lib, co, po):
    """2^25 uses the three-pass plan: in place, against the oracle."""
    from halo2_amd import device as dv, fields as F, ntt

    k = 25
    print("fft three-pass k=25", flush=True)
    n = 1 << k
    a = rand_fr(co, n, 77)
    exp = co.fft(a, k, F.mont_limbs(F.omega(k)))
    d = dv.to_device(a)
    ntt.fft_normal_device(1, k, d, d, F.omega(k), 1)
    assert (dv.to_host(d) == exp).all()


def test_round_trip_and_linearity_2_24(lib):
    """Size-independent checks at 2^24: iFFT(FFT(a)) == a, and CosetFFT_Part is linear
    (T(a + b) == T(a) + T(b) on sampled rows)."""
    import torch

    from halo2_amd import _lib, device as dv, fields as F, ntt

    k = 24
    n = 1 << k
    print("round trip k=24", flush=True)
    a = dv.empty_u64(4 * n)
    b = dv.empty_u64(4 * n)
    _lib.check(lib.h2amd_random_fr(dv.ptr(a), n, 5, 0, dv.stream()))
    _lib.check(lib.h2amd_random_fr(dv.ptr(b), n, 6, 0, dv.stream()))
    fa = dv.empty_u64(4 * n)
    ntt.fft_normal_device(1, k, a, fa, F.omega(k), 1)
    back = dv.empty_u64(4 * n)
    ntt.fft_normal_device(2, k, fa, back, F.inv(F.omega(k)), F.inv(n))
    assert torch.equal(back, a)
    del fa, back
    c = F.ZETA
    ta, tb, ts = dv.empty_u64(4 * n), dv.empty_u64(4 * n), dv.empty_u64(4 * n)
    ntt.fft_normal_device(6, k, a, ta, F.omega(k), c)
    ntt.fft_normal_device(6, k, b, tb, F.omega(k), c)
    s = dv.to_device(_limb_add_mod(dv.to_host(a), dv.to_host(b)))
    ntt.fft_normal_device(6, k, s, ts, F.omega(k), c)
    hta, htb, hts = dv.to_host(ta), dv.to_host(tb), dv.to_host(ts)
    for r in np.random.default_rng(1).choice(n, 256, replace=False):
        assert (F.value_of(hta[r]) + F.value_of(htb[r])) % F.R_MOD == F.value_of(hts[r])


@pytest.mark.parametrize("k", [1, 6, 12, 13, 20, 24, 25])
def test_fft_largest_stored_inputs(lib, co, po, k):
    """Every input word = r - 1, the largest reduced representative (the worst case for
    lazily reduced butterflies). Montgomery storage is linear, so the transform of a
    constant has the closed form: FFT -> (N (r - 1), 0, ..., 0) and iFFT -> (r - 1, 0, ..., 0)
    in stored words; CosetFFT_Part at small k is checked against the oracle."""
    import torch

    from halo2_amd import device as dv, fields as F, ntt

    n = 1 << k
    print(f"largest stored inputs k={k}", flush=True)
    top = F.limbs(F.R_MOD - 1).astype(np.uint64)
    d = torch.tensor(np.tile(top, n).view(np.int64), device="cuda")
    out = dv.empty_u64(4 * n)
    ntt.fft_normal_device(1, k, d, out, F.omega(k), 1)
    h = dv.to_host(out)
    assert F.from_limbs(h[0]) == (n * (F.R_MOD - 1)) % F.R_MOD
    assert not h[1:].any()
    ntt.fft_normal_device(2, k, d, out, F.inv(F.omega(k)), F.inv(n))
    h = dv.to_host(out)
    assert F.from_limbs(h[0]) == F.R_MOD - 1
    assert not h[1:].any()
    if k <= 13:
        a = np.tile(top, n).reshape(n, 4)
        c = F.ZETA
        ntt.fft_normal_device(6, k, d, out, F.omega(k), c)
        assert (dv.to_host(out) == _oracle_kind(co, po, 6, a, k, c=c)).all()


def _limb_add_mod(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Vectorised (x + y) mod r on Montgomery limbs (Montgomery form is additive)."""
    from halo2_amd import fields as F

    mod = F.limbs(F.R_MOD).astype(np.uint64)
    out = np.zeros_like(x)
    carry = np.zeros(len(x), dtype=np.uint64)
    for i in range(4):
        s = x[:, i] + y[:, i]
        c1 = (s < x[:, i]).astype(np.uint64)
        s2 = s + carry
        c2 = (s2 < s).astype(np.uint64)
        out[:, i] = s2
        carry = c1 | c2
    # subtract r where out >= r
    ge = np.zeros(len(x), dtype=bool)
    decided = np.zeros(len(x), dtype=bool)
    for i in (3, 2, 1, 0):
        gt = (out[:, i] > mod[i]) & ~decided
        lt = (out[:, i] < mod[i]) & ~decided
        ge |= gt
        decided |= gt | lt
    ge |= ~decided  # equal
    borrow = np.zeros(len(x), dtype=np.uint64)
    res = out.copy()
    for i in range(4):
        d = out[:, i] - mod[i]
        b1 = (out[:, i] < mod[i]).astype(np.uint64)
        d2 = d - borrow
        b2 = (d < borrow).astype(np.uint64)
        res[:, i] = d2
        borrow = b1 | b2
    return np.where(ge[:, None], res, out)


def test_fft_normal_host(lib, co, po):
    from halo2_amd import fields as F, ntt

    for k in (3, 10, 14):
        n = 1 << k
        a = rand_fr(co, n, 300 + k)
        for kind, omega, divisor in ((1, F.omega(k), 1), (2, F.inv(F.omega(k)), F.inv(n)),
                                     (6, F.omega(k), F.ZETA)):
            exp = _oracle_kind(co, po, kind, a, k, c=divisor)
            assert (ntt.fft_gpu(kind, a, k, omega, divisor) == exp).all(), (k, kind)
        # cosetFFT at the same size (the reference's _fft_normal runs kinds at 2^log_n)
        exp = _oracle_kind(co, po, 3, a, k, ext_k=k)
        assert (ntt.fft_gpu(3, a, k, F.omega(k), 1) == exp).all()
        # icosetFFT inverts cosetFFT
        fwd = ntt.fft_gpu(3, a, k, F.omega(k), 1)
        assert (ntt.fft_gpu(4, fwd, k, F.inv(F.omega(k)), F.inv(n)) == a).all()


@pytest.mark.parametrize("k,ext", [(3, 5), (10, 12), (14, 16), (16, 18)])
def test_cosetfft_extension(lib, co, po, k, ext):
    from halo2_amd import fields as F, ntt

    a = rand_fr(co, 1 << k, 400 + k)
    exp = _oracle_kind(co, po, 3, a, k, ext_k=ext)
    assert (ntt.cosetfft_gpu(a, k, ext, F.omega(ext)) == exp).all()


def test_fft_many_residencies(lib, co, po):
    from halo2_amd import device as dv, fields as F, ntt

    k, ext = 11, 13
    polys = [rand_fr(co, 1 << k, 500 + i) for i in range(5)]
    # coset extension many (host -> host)
    outs = ntt.fft_many_host(3, polys, k, ext, F.omega(ext), 1)
    for p, o in zip(polys, outs):
        assert (o == _oracle_kind(co, po, 3, p, k, ext_k=ext)).all()
    # CosetFFT_Part many: host inputs -> device outputs, and device inputs -> device outputs
    c = F.ZETA * F.omega(ext) % F.R_MOD
    exps = [_oracle_kind(co, po, 6, p, k, c=c) for p in polys]
    for o, e in zip(ntt.fft_many_to_device(6, polys, k, k, F.omega(k), c), exps):
        assert (dv.to_host(o) == e).all()
    d_in = [dv.to_device(p) for p in polys]
    for o, e in zip(ntt.fft_many_to_device(6, d_in, k, k, F.omega(k), c), exps):
        assert (dv.to_host(o) == e).all()
    # aliased device batch (in == out) takes the scratch path one job at a time
    ntt.fft_many_to_device(6, d_in, k, k, F.omega(k), c, outputs=d_in)
    for o, e in zip(d_in, exps):
        assert (dv.to_host(o) == e).all()
    # iFFT many device
    ds = [dv.to_device(p) for p in polys]
    outs = ntt.fft_many_to_device(2, ds, k, k, F.inv(F.omega(k)), F.inv(1 << k))
    for p, o in zip(polys, outs):
        assert (dv.to_host(o) == _oracle_kind(co, po, 2, p, k)).all()


def test_many_batches_beyond_one_launch(lib, co, po):
    """More polynomials than one launch carries (32) at a two-pass size."""
    from halo2_amd import device as dv, fields as F, ntt

    k = 14
    polys = [rand_fr(co, 1 << k, 900 + i) for i in range(35)]
    c = F.ZETA
    outs = ntt.fft_many_to_device(6, [dv.to_device(p) for p in polys], k, k, F.omega(k), c)
    for i in (0, 17, 31, 32, 34):
        assert (dv.to_host(outs[i]) == _oracle_kind(co, po, 6, polys[i], k, c=c)).all(), i


def test_domain_methods(lib, co, po):
    """EvaluationDomain (j = 5 -> quotient degree 4, extended_k = k + 2)."""
    from halo2_amd import device as dv, fields as F
    from halo2_amd.domain import EvaluationDomain

    k = 10
    dom = EvaluationDomain(5, k)
    assert dom.extended_k == k + 2 and dom.quotient_poly_degree == 4
    a = rand_fr(co, 1 << k, 600)
    # parts of the extended transform interleave into coeff_to_extended
    ext = dom.coeff_to_extended_many([a])[0]
    assert (ext == _oracle_kind(co, po, 3, a, k, ext_k=k + 2)).all()
    parts = []
    for p in range(4):
        parts += dom.coeff_to_extended_part_many_device([dv.to_device(a)], pow(dom.extended_omega, p, F.R_MOD))
    inter = dv.to_host(dom.extended_from_lagrange_vec_device(parts))
    assert (inter == ext).all()
    # extended_to_coeff inverts coeff_to_extended (quotient degree 4 keeps all 4n coefficients)
    back = dv.to_host(dom.extended_to_coeff_device(dv.to_device(ext)))
    assert (back[: 1 << k] == a).all() and not back[1 << k:].any()
    # divide_by_vanishing: t_evaluations cycle with period 4
    d = dv.to_device(ext)
    dom.divide_by_vanishing_poly_device(d)
    t = np.stack([F.mont_limbs(v) for v in dom.t_evaluations])
    assert (dv.to_host(d) == co.divide_by_vanishing(ext, t)).all()
    # t_evaluations[i] = 1 / ((zeta * ext_omega^i)^n - 1)
    for i, v in enumerate(dom.t_evaluations):
        x = pow(F.ZETA * pow(dom.extended_omega, i, F.R_MOD) % F.R_MOD, 1 << k, F.R_MOD)
        assert v * (x - 1) % F.R_MOD == 1


def test_lagrange_to_coeff(lib, co, po):
    from halo2_amd import device as dv, fields as F
    from halo2_amd.domain import EvaluationDomain

    k = 12
    dom = EvaluationDomain(4, k)
    polys = [rand_fr(co, 1 << k, 700 + i) for i in range(3)]
    outs = dom.lagrange_to_coeff_many_device(polys)
    for p, o in zip(polys, outs):
        assert (dv.to_host(o) == _oracle_kind(co, po, 2, p, k)).all()
    one = dom.lagrange_to_coeff_device_input(dv.to_device(polys[0]))
    assert (dv.to_host(one) == dv.to_host(outs[0])).all()


def test_ntt_error_paths(lib):
    import ctypes

    from halo2_amd import _lib, device as dv

    one = ctypes.create_string_buffer(32)
    scratch = dv.empty_bytes(1 << 20)
    buf = dv.empty_u64(64)
    e = lib._halo2_fft_normal_to_device(3, 4, dv.ptr(buf), dv.ptr(buf), dv.ptr(buf), one, dv.ptr(scratch), 1 << 20,
                                        dv.stream())
    assert e.code != 0
    _libc_free(e)
    e = lib._halo2_fft_normal_to_device(1, 0, dv.ptr(buf), dv.ptr(buf), dv.ptr(buf), one, dv.ptr(scratch), 1 << 20,
                                        dv.stream())
    assert e.code != 0
    _libc_free(e)
    # workspace too small is an error, not a fault
    e = lib._halo2_fft_normal_to_device(1, 16, dv.ptr(buf), dv.ptr(buf), dv.ptr(buf), one, dv.ptr(scratch), 1024,
                                        dv.stream())
    assert e.code != 0
    _libc_free(e)
    with pytest.raises(_lib.Halo2GpuError) as ei:
        _lib.check(lib._halo2_fft_many_to_device(4, 1, 4, 4, _lib.obj_array([dv.ptr(buf)]),
                                                 _lib.obj_array([dv.ptr(buf)]), dv.ptr(buf),
                                                 ctypes.byref(_lib.obj(ctypes.addressof(one))), True,
                                                 dv.ptr(scratch), 1 << 20, dv.stream()))
    assert ei.value.code == 2
    with pytest.raises(_lib.Halo2GpuError) as ei:
        _lib.check(lib._halo2_fft_normal(5, 4, None, None, None, None, None, 0, dv.stream()))
    assert ei.value.code == -1


def _libc_free(e):
    from halo2_amd import _lib

    if e.message:
        _lib._libc.free(e.message)


def test_fft_batch_uneven(lib, co, po):
    """Uneven batches (one- and two-pass sizes side by side, 2^1 .. 2^18) of FFT, iFFT,
    CosetFFT_Part and cosetFFT, each job against the oracle."""
    from halo2_amd import device as dv, fields as F, ntt

    ks = [1, 3, 5, 12, 13, 14, 16, 18, 14, 7]
    print(f"fft batch sizes {ks}", flush=True)
    polys = [rand_fr(co, 1 << k, 1300 + i) for i, k in enumerate(ks)]
    d_in = [dv.to_device(p) for p in polys]
    cs = [F.ZETA * pow(F.omega(k + 2), 3, F.R_MOD) % F.R_MOD for k in ks]
    cs[8] = cs[5] + 1  # two jobs of one size with different coset factors
    for kind in (1, 2, 6):
        omegas = [F.omega(k) if kind != 2 else F.inv(F.omega(k)) for k in ks]
        divs = {1: [1] * len(ks), 2: [F.inv(1 << k) for k in ks], 6: cs}[kind]
        outs = ntt.fft_batch_to_device(kind, d_in, ks, ks, omegas, divs)
        for i, k in enumerate(ks):
            exp = _oracle_kind(co, po, kind, polys[i], k, c=divs[i])
            assert (dv.to_host(outs[i]) == exp).all(), f"kind {kind} job {i} k={k}"
        for p, d in zip(polys, d_in):
            assert (dv.to_host(d) == p).all(), "inputs must not be modified"
    pairs = [(3, 5), (10, 12), (14, 16), (16, 18), (1, 3)]
    cpolys = [rand_fr(co, 1 << k, 1400 + k) for k, _ in pairs]
    outs = ntt.fft_batch_to_device(3, [dv.to_device(p) for p in cpolys], [k for k, _ in pairs],
                                   [e for _, e in pairs], [F.omega(e) for _, e in pairs], [1] * len(pairs))
    for (k, e), p, o in zip(pairs, cpolys, outs):
        assert (dv.to_host(o) == _oracle_kind(co, po, 3, p, k, ext_k=e)).all(), f"coset k={k} ext={e}"


def test_fft_batch_three_pass_and_refusals(lib, co, po):
    """A three-pass transform (2^25, scratch in the workspace) in a batch with small ones
    matches the single-transform entry point; an in-place two-pass job is refused."""
    from halo2_amd import _lib, device as dv, fields as F, ntt, msm

    ks = [25, 4, 14]
    print("fft batch with 2^25", flush=True)
    d_in = [msm.random_fr(1 << k, 1500 + k) for k in ks]
    outs = ntt.fft_batch_to_device(1, d_in, ks, ks, [F.omega(k) for k in ks], [1] * len(ks))
    for k, d, o in zip(ks, d_in, outs):
        ref = dv.empty_u64(4 << k)
        ntt.fft_normal_device(1, k, d, ref, F.omega(k), 1)
        assert bool((ref == o).all()), f"k={k}"
    small = rand_fr(co, 1 << 4, 3)
    assert (dv.to_host(outs[1]) == _oracle_kind(co, po, 1, dv.to_host(d_in[1]), 4)).all()
    d = dv.to_device(rand_fr(co, 1 << 14, 4))
    with pytest.raises(_lib.Halo2GpuError):
        ntt.fft_batch_to_device(1, [d], [14], [14], [F.omega(14)], [1], outputs=[d])
    # a one-pass transform may run in place
    d2 = dv.to_device(small)
    ntt.fft_batch_to_device(1, [d2], [4], [4], [F.omega(4)], [1], outputs=[d2])
    assert (dv.to_host(d2) == _oracle_kind(co, po, 1, small, 4)).all()


def test_coeff_to_extended_batch(lib, co, po):
    """Domains of k = 4 .. 16 (extended k + 2): the batched coeff_to_extended equals each
    domain's own."""
    from halo2_amd import device as dv
    from halo2_amd.domain import EvaluationDomain, coeff_to_extended_batch_device

    ks = [16, 4, 9, 12, 13, 10, 16]
    doms = [EvaluationDomain(5, k) for k in ks]
    polys = [dv.to_device(rand_fr(co, 1 << k, 1600 + i)) for i, k in enumerate(ks)]
    outs = coeff_to_extended_batch_device(doms, polys)
    for dom, p, o in zip(doms, polys, outs):
        assert bool((dom.coeff_to_extended_many_device([p])[0] == o).all()), f"k={dom.k}"
