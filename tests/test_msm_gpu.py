"""GPU parity of the BN254 G1 MSM against the CPU oracle (bit-exact affine output),
through the C ABI. Mirrors tests/multiexp_dense_bucket_edge_equivalence.rs of the
reference (random, dense constant prefix, zero scalars + identity bases, all-zero,
single non-zero, ten dense buckets per window, sub-tile dense chunk; chunked and
unchunked) and src/cuda/tests.rs:390-585, plus closed forms at 2^20+:
sum c_i (s_i G) = (sum c_i s_i) G and the KZG identity sum c_i tau^i G = f(tau) G."""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _dense_constant():
    from halo2_amd import fields as F

    v = F.from_limbs([0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9, 0x94D049BB133111EB, 0x022B5F6A3C1D0E8F])
    return F.mont_limbs(v)


def _ten_group_scalar(g):
    from halo2_amd import fields as F

    b = bytes((((g + j) % 250) + 1) for j in range(31)) + b"\x00"
    return F.mont_limbs(int.from_bytes(b, "little"))


def _check_all_paths(co, scalars, points, label, chunked=True):
    """GPU device/device, device-bases, host/host and chunked vs the oracle MSM."""
    from halo2_amd import device as dv, msm

    n = len(scalars)
    exp = co.msm(scalars, points)
    ds, dp = dv.to_device(scalars), dv.to_device(points)
    got = msm.multiexp_device_scalars_device_bases(ds, dp, n)
    assert (got == exp).all(), f"{label}: device scalars/bases"
    assert (dv.to_host(ds) == scalars).all(), "scalars must not be modified"
    assert (msm.multiexp_device_bases(scalars, dp) == exp).all(), f"{label}: device bases"
    assert (msm.multiexp(scalars, points) == exp).all(), f"{label}: host"
    if chunked and n >= 8:
        assert (msm.multiexp_gpu_device_scalars_device_bases_chunked(ds, dp, n, n // 4 + 1) == exp).all(), \
            f"{label}: chunked"


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 17, 31, 32, 100, 1000, 1 << 14])
def test_msm_random_small(lib, co, n):
    print(f"msm random n={n}", flush=True)
    s = rand_fr(co, n, 3000 + n)
    c = rand_fr(co, n, 4000 + n)
    pts = co.g1_mul_gen(s)
    _check_all_paths(co, c, pts, f"random n={n}")


def test_msm_edge_cases(lib, co):
    from halo2_amd import fields as F

    n = 1 << 14
    k = _dense_constant()
    s = rand_fr(co, n, 11)
    c = rand_fr(co, n, 12)
    pts = co.g1_mul_gen(s)
    zero = np.zeros(4, dtype=np.uint64)
    print("msm edge: zeros + identity bases", flush=True)
    cz = c.copy()
    for i in range(n):
        if i % 3 == 0:
            cz[i] = zero
        elif i % 7 == 0:
            cz[i] = k
    pz = pts.copy()
    pz[::5] = 0  # identity bases (0, 0)
    _check_all_paths(co, cz, pz, "zeros + identity bases")
    print("msm edge: all zero", flush=True)
    allz = np.zeros_like(c)
    _check_all_paths(co, allz, pts, "all-zero scalars")
    from halo2_amd import msm, device as dv

    out = msm.multiexp_device_scalars_device_bases(dv.to_device(allz), dv.to_device(pts), n)
    assert (out[:4] == 0).all() and (out[4:8] == F.mont_limbs(1, F.P_MOD)).all() and (out[8:] == 0).all()
    print("msm edge: single nonzero", flush=True)
    single = np.zeros_like(c)
    single[0] = k
    _check_all_paths(co, single, pts, "single nonzero")
    print("msm edge: dense prefix", flush=True)
    dense = c.copy()
    dense[: n // 4] = k
    _check_all_paths(co, dense, pts, "dense 25% prefix")
    print("msm edge: all equal (one bucket per window)", flush=True)
    _check_all_paths(co, np.tile(k, (n, 1)), pts, "all equal")
    print("msm edge: repeated points and P + (-P)", flush=True)
    rep = np.tile(pts[5], (n, 1))
    _check_all_paths(co, c, rep, "same base everywhere (doubling inside buckets)")
    negs = pts.copy()
    negs[1::2] = pts[0::2]
    csame = np.tile(c[0], (n, 1))
    _check_all_paths(co, csame, negs, "pairs of equal points, equal scalars")
    # pairs (P, -P) with equal scalars: every pair cancels inside a bucket
    inv = pts.copy()
    for i in range(1, n, 2):
        y = F.from_limbs(pts[i - 1, 4:8])
        inv[i, :4] = pts[i - 1, :4]
        inv[i, 4:8] = F.limbs((F.P_MOD - y) % F.P_MOD)
    _check_all_paths(co, csame, inv, "pairs (P, -P) cancel")


def test_msm_ten_dense_buckets_and_subtile(lib, co):
    print("msm ten dense buckets per window", flush=True)
    groups, per = 10, 6554
    n = groups * per
    c = np.concatenate([np.tile(_ten_group_scalar(g), (per, 1)) for g in range(groups)])
    pts = co.g1_mul_gen(rand_fr(co, n, 21))
    _check_all_paths(co, c, pts, "ten dense buckets / window", chunked=True)
    print("msm sub-tile dense n=20", flush=True)
    c20 = np.tile(_dense_constant(), (20, 1))
    _check_all_paths(co, c20, co.g1_mul_gen(rand_fr(co, 20, 22)), "sub-tile dense", chunked=False)


def test_msm_65536_vs_oracle(lib, co):
    print("msm n=65536 vs oracle Pippenger", flush=True)
    n = 1 << 16
    s = rand_fr(co, n, 31)
    c = rand_fr(co, n, 32)
    pts = co.g1_mul_gen(s)
    _check_all_paths(co, c, pts, "2^16", chunked=False)


def test_g1_mul_generator_matches_oracle(lib, co):
    from halo2_amd import device as dv, msm

    n = 3000
    s = rand_fr(co, n, 41)
    s[7] = 0
    got = dv.to_host(msm.g1_mul_generator(dv.to_device(s), n), cols=8)
    assert (got == co.g1_mul_gen(s)).all()


@pytest.mark.parametrize("k", [20, 22, 24])
def test_msm_closed_form_large(lib, co, k):
    """sum c_i (s_i G) = (sum c_i s_i) G with device-generated bases and scalars. At 2^24
    the plain path runs 20-bit windows (13 windows of 2^19 signed buckets) without GLV."""
    from halo2_amd import device as dv, msm

    n = 1 << k
    print(f"msm closed form n=2^{k}", flush=True)
    ds = msm.random_fr(n, 51 + k)
    dc = msm.random_fr(n, 61 + k)
    pts = msm.g1_mul_generator(ds, n)
    got = msm.multiexp_device_scalars_device_bases(dc, pts, n)
    ip = co.fr_inner_product(dv.to_host(dc), dv.to_host(ds))
    exp = co.g1_mul(co.g1_generator(), ip)
    assert (got == exp).all()


def test_msm_closed_form_dense_2_20(lib, co):
    """Every scalar equal (one bucket per window holding all 2^20 entries, merged by the
    work-group tree) and a half-zero / half-constant mix: closed form (c sum s_i) G."""
    import torch

    from halo2_amd import device as dv, msm

    n = 1 << 20
    print("msm dense closed form n=2^20", flush=True)
    ds = msm.random_fr(n, 71)
    pts = msm.g1_mul_generator(ds, n)
    k = _dense_constant()
    dc = torch.tensor(np.tile(k, n).view(np.int64), device=ds.device)
    got = msm.multiexp_device_scalars_device_bases(dc, pts, n)
    hs = dv.to_host(ds)
    exp = co.g1_mul(co.g1_generator(), co.fr_inner_product(np.tile(k, (n, 1)), hs))
    assert (got == exp).all()
    half = np.tile(k, (n, 1))
    half[::2] = 0
    got = msm.multiexp_device_scalars_device_bases(dv.to_device(half), pts, n)
    assert (got == co.g1_mul(co.g1_generator(), co.fr_inner_product(half, hs))).all()


def test_kzg_commit_srs_identity(lib, co):
    """ParamsKZG with a seeded tau: commit(f) = f(tau) G, resident SRS, device and host polys."""
    from halo2_amd import device as dv, fields as F, msm

    k = 18
    n = 1 << k
    tau = 0x1234567890ABCDEF1234567890ABCDEF % F.R_MOD
    print(f"kzg commit k={k}", flush=True)
    params = msm.ParamsKZG(k, tau)
    poly = rand_fr(co, n, 71)
    pows = co.fr_powers(F.mont_limbs(tau), n)
    # the SRS itself: g[i] = tau^i G on a sample
    g = dv.to_host(params.g_device, cols=8)
    idx = [0, 1, 2, 12345, n - 1]
    assert (g[idx] == co.g1_mul_gen(pows[idx])).all()
    exp = co.g1_mul(co.g1_generator(), co.fr_inner_product(poly, pows))
    assert (params.commit_device(dv.to_device(poly)) == exp).all()
    assert (params.commit(poly) == exp).all()
    # a shorter polynomial commits against the SRS prefix
    m = 1000
    exp_m = co.g1_mul(co.g1_generator(), co.fr_inner_product(poly[:m], pows[:m]))
    assert (params.commit(poly[:m]) == exp_m).all()


@pytest.mark.parametrize("k", [22, 24])
def test_kzg_commit_srs_identity_large(lib, co, k):
    """The KZG commitment of the largest sizes, 2^22 and 2^24 points: f(tau) G through the
    fixed-base table of the resident SRS (20-bit table windows: 13 rows of 2^k points) and
    through the plain device MSM on the same bases, both against the closed form."""
    import torch

    from halo2_amd import device as dv, fields as F, msm

    n = 1 << k
    tau = (0x5EED0000 + k) * 0x9E3779B97F4A7C15 % F.R_MOD
    print(f"kzg commit closed form k={k}", flush=True)
    params = msm.ParamsKZG(k, tau)
    dc = msm.random_fr(n, 101 + k)
    pows = co.fr_powers(F.mont_limbs(tau), n)
    hc = dv.to_host(dc)
    exp = co.g1_mul(co.g1_generator(), co.fr_inner_product(hc, pows))
    assert (params.commit_device(dc) == exp).all(), "fixed-base table"
    assert (msm.multiexp_device_scalars_device_bases(dc, params.g_device, n) == exp).all(), "plain"
    m = n - 12345  # a prefix of the table
    exp_m = co.g1_mul(co.g1_generator(), co.fr_inner_product(hc[:m], pows[:m]))
    assert (params.commit_device(dc[: 4 * m]) == exp_m).all(), "prefix"
    del params
    torch.cuda.empty_cache()


def test_kzg_commit_lagrange(lib, co):
    """The Lagrange basis (commitment.rs:157-183): g_lagrange[i] = L_i(tau) G with
    L_i(tau) = (tau^n - 1)/n * w^i/(tau - w^i) on a sample of i, and commit_lagrange(v)
    equals commit(iFFT(v)) — the basis is the inverse DFT of the powers basis."""
    from halo2_amd import device as dv, fields as F, msm
    from halo2_amd.domain import EvaluationDomain

    k = 14
    n = 1 << k
    R = F.R_MOD
    tau = 0x77AA55 % R
    params = msm.ParamsKZG(k, tau)
    w = F.omega(k)
    gl = dv.to_host(params.g_lagrange_device, cols=8)
    mult = (pow(tau, n, R) - 1) * pow(n, -1, R) % R
    idx = [0, 1, 777, n - 1]
    scal = np.stack([F.mont_limbs(mult * pow(w, i, R) * pow(tau - pow(w, i, R), -1, R) % R) for i in idx])
    assert (gl[idx] == co.g1_mul_gen(scal)).all()
    values = rand_fr(co, n, 72)
    coeffs = EvaluationDomain(3, k).lagrange_to_coeff_device_input(dv.to_device(values))
    exp = params.commit_device(coeffs)
    assert (params.commit_lagrange_device(dv.to_device(values)) == exp).all()
    assert (params.commit_lagrange(values) == exp).all()
    pows = co.fr_powers(F.mont_limbs(tau), n)
    assert (exp == co.g1_mul(co.g1_generator(), co.fr_inner_product(dv.to_host(coeffs), pows))).all()


def _plan_width(n):
    """Window width of the fixed-base table for n bases: clamp(log2 n, 4, 22), then the
    narrowest width with the same window count."""
    c = max(4, min(22, n.bit_length() - 1))
    w0 = -(-255 // c)
    return -(-255 // w0)


@pytest.mark.parametrize("n", [1, 2, 17, 1000, 1 << 16, (1 << 20) + 3, 1 << 22])
def test_msm_fixed_base_table(lib, co, n):
    """The fixed-base path equals the plain device MSM bit for bit: full length, prefixes
    of the table, identity bases and edge scalars; table rows are 2^(c w) P."""
    from halo2_amd import device as dv, fields as F, msm

    print(f"fixed-base n={n}", flush=True)
    pts = msm.g1_mul_generator(msm.random_fr(n, seed=90 + n % 1000), n)
    if n >= 17:  # identity bases (0, 0)
        pts.view(-1, 8)[3].zero_()
        pts.view(-1, 8)[n - 1].zero_()
    sc = rand_fr(co, n, 91)
    if n >= 17:
        sc[5] = 0
        sc[6] = F.mont_limbs(F.R_MOD - 1)
        sc[7] = F.mont_limbs(1)
    d_sc = dv.to_device(sc)
    table = msm.msm_precompute(pts, n)
    for m in sorted({n, max(1, n // 3), max(1, n - 1)}):
        exp = msm.multiexp_device_scalars_device_bases(d_sc[: 4 * m], pts[: 8 * m], m)
        got = msm.multiexp_precomputed(d_sc[: 4 * m], table, n, m)
        assert (got == exp).all(), m
    if n <= 1000:
        c = _plan_width(n)
        W = -(-255 // c)
        rows = dv.to_host(table, cols=8)
        base = dv.to_host(pts, cols=8)
        i = n // 2
        for w in (1, W - 1):
            s = F.mont_limbs(pow(2, c * w, F.R_MOD))
            exp_pt = co.g1_mul(base[i], s)
            # the table holds the 2^261-scale form: compare through the oracle's affine bytes
            got_aff = rows[w * n + i]
            conv = lambda v: F.mont_limbs(F.value_of(v, F.P_MOD) * pow(2, -5, F.P_MOD) % F.P_MOD, F.P_MOD)  # noqa: E731
            assert (np.concatenate([conv(got_aff[:4]), conv(got_aff[4:])]) == exp_pt[:8]).all(), w


def test_kzg_fixed_base_commit(lib, co):
    """ParamsKZG commits through the fixed-base tables (default) and the plain path agree."""
    from halo2_amd import device as dv, msm

    k = 12
    params = msm.ParamsKZG(k, 0xABCDEF)
    plain = msm.ParamsKZG(k, 0xABCDEF, fixed_base=False)
    poly = rand_fr(co, 1 << k, 93)
    assert (params.commit_device(dv.to_device(poly)) == plain.commit_device(dv.to_device(poly))).all()
    assert (params.commit(poly[:100]) == plain.commit(poly[:100])).all()
    assert (params.commit_lagrange(poly) == plain.commit_lagrange(poly)).all()


def _batch_case(co, lengths, seed, nmax):
    """Device bases s_i G (i < nmax) and per-job scalars: expected results from the
    closed form (sum_i c_i s_i) G on the oracle."""
    from halo2_amd import device as dv, msm

    ds = msm.random_fr(nmax, seed)
    pts = msm.g1_mul_generator(ds, nmax)
    hs = dv.to_host(ds)
    scal, exp = [], []
    for j, n in enumerate(lengths):
        c = rand_fr(co, n, seed * 100 + j) if n else np.zeros((0, 4), dtype=np.uint64)
        if j % 5 == 3 and n:  # a dense job: every scalar equal (one bucket per window)
            c = np.tile(_dense_constant(), (n, 1))
        if j % 7 == 6 and n:  # zero scalars
            c[: n // 2] = 0
        scal.append(c)
        exp.append(co.g1_mul(co.g1_generator(), co.fr_inner_product(c, hs[:n])) if n else None)
    return pts, scal, exp


def _identity():
    from halo2_amd import fields as F

    return np.concatenate([np.zeros(4, np.uint64), F.mont_limbs(1, F.P_MOD), np.zeros(4, np.uint64)])


def test_msm_batch_uneven(lib, co):
    """MSMs of uneven lengths (empty, tile edges, 2^16 + a few, 2^18) in one batched
    pipeline: each equals its closed form and the single-MSM result."""
    import torch

    from halo2_amd import device as dv, msm

    lengths = [0, 1, 3, 4095, 4096, 4097, 12345, 1 << 16, (1 << 16) + 5, 1, 1 << 18, 77, 0, 5000]
    nmax = max(lengths)
    print(f"msm batch {len(lengths)} jobs, max {nmax}", flush=True)
    pts, scal, exp = _batch_case(co, lengths, 81, nmax)
    d = [dv.to_device(c) if len(c) else torch.zeros(4, dtype=torch.int64, device=pts.device) for c in scal]
    got = msm.multiexp_batch_device(d, pts, lengths)
    for j, n in enumerate(lengths):
        if n == 0:
            assert (got[j] == _identity()).all(), f"job {j} empty"
            continue
        assert (got[j] == exp[j]).all(), f"job {j} n={n}"
        if n in (4097, 1 << 16):
            assert (got[j] == msm.multiexp_device_scalars_device_bases(d[j], pts, n)).all()


def test_msm_batch_precomputed_uneven(lib, co):
    """The batch over a fixed-base table: uneven lengths (empty, tile edges, dense,
    half-zero, the table's full length) and lookup-style small scalars, each equal to
    its closed form; a job longer than the table is refused."""
    import torch

    from halo2_amd import _lib, device as dv, fields as F, msm

    lengths = [0, 1, 4095, 4096, 4097, 12345, 1 << 16, 1, 77, 0, 5000, (1 << 16) - 3]
    nmax = 1 << 16
    print(f"msm batch precomputed {len(lengths)} jobs, table {nmax}", flush=True)
    pts, scal, exp = _batch_case(co, lengths, 83, nmax)
    small = np.stack([F.mont_limbs(int(v)) for v in np.random.default_rng(84).integers(0, 1 << 16, 3000)])
    scal.append(small)
    exp.append(co.msm(small, dv.to_host(pts, cols=8)[:3000]))
    lengths.append(3000)
    table = msm.msm_precompute(pts, nmax)
    d = [dv.to_device(c) if len(c) else torch.zeros(4, dtype=torch.int64, device=pts.device) for c in scal]
    got = msm.multiexp_batch_precomputed(d, table, nmax, lengths)
    for j, n in enumerate(lengths):
        if n == 0:
            assert (got[j] == _identity()).all(), f"job {j} empty"
            continue
        assert (got[j] == exp[j]).all(), f"job {j} n={n}"
    with pytest.raises(_lib.Halo2GpuError):
        msm.multiexp_batch_precomputed([msm.random_fr(nmax + 1, 85)], table, nmax)


def test_msm_batch_many_jobs_and_limits(lib, co):
    """64 small jobs (the most one batch takes) against closed forms; 65 is refused and
    an all-empty batch gives identities."""
    import torch

    from halo2_amd import _lib, device as dv, msm

    rng = np.random.default_rng(5)
    lengths = [int(x) for x in rng.integers(1, 3000, 64)]
    pts, scal, exp = _batch_case(co, lengths, 82, max(lengths))
    d = [dv.to_device(c) for c in scal]
    got = msm.multiexp_batch_device(d, pts)
    for j in range(64):
        assert (got[j] == exp[j]).all(), f"job {j}"
    with pytest.raises(_lib.Halo2GpuError):
        msm.multiexp_batch_device(d + d[:1], pts)
    z = torch.zeros(4, dtype=torch.int64, device=pts.device)
    got = msm.multiexp_batch_device([z, z], pts, [0, 0])
    assert (got == _identity()).all()


def test_kzg_commit_many(lib, co):
    """ParamsKZG.commit_many_device: polynomials of several lengths committed in one
    batch equal commit_device one by one and f(tau) G."""
    from halo2_amd import device as dv, fields as F, msm

    k = 14
    n = 1 << k
    tau = 0xABCDEF0123456789 % F.R_MOD
    params = msm.ParamsKZG(k, tau, fixed_base=False)
    pows = co.fr_powers(F.mont_limbs(tau), n)
    lens = [n, 100, n // 2 + 1, 4096, 1]
    polys = [rand_fr(co, m, 90 + i) for i, m in enumerate(lens)]
    got = params.commit_many_device([dv.to_device(p) for p in polys])
    for j, p in enumerate(polys):
        exp = co.g1_mul(co.g1_generator(), co.fr_inner_product(p, pows[: len(p)]))
        assert (got[j] == exp).all(), f"poly {j}"
        assert (params.commit_device(dv.to_device(p)) == exp).all()


@pytest.mark.parametrize("k", [0, 1, 2, 5, 8, 11])
def test_g_to_lagrange_matches_setup(lib, co, k):
    """g_to_lagrange(g) on the device equals the Lagrange bases the setup builds
    directly from tau (L_i(tau) G, commitment.rs:157-183): two independent routes."""
    from halo2_amd import device as dv, fields as F, msm

    print(f"g_to_lagrange k={k}", flush=True)
    params = msm.ParamsKZG(k, 0x7A0_5EED % F.R_MOD, fixed_base=False)
    got = msm.g_to_lagrange_device(params.g_device, k)
    assert (dv.to_host(got, cols=8) == dv.to_host(params.g_lagrange_device, cols=8)).all()


def test_g_to_lagrange_identities_and_downsize(lib, co):
    """Bases (G, O, O, ...): every output is n^-1 G (identity inputs through every stage);
    downsize(k) equals a fresh setup at k; from_parts without g_lagrange computes it."""
    from halo2_amd import device as dv, fields as F, msm

    k = 6
    n = 1 << k
    g = np.zeros((n, 8), dtype=np.uint64)
    g[0] = co.g1_generator()
    out = dv.to_host(msm.g_to_lagrange_device(dv.to_device(g), k), cols=8)
    exp = co.g1_mul(co.g1_generator(), F.mont_limbs(F.inv(n)))
    for i in range(n):
        assert (out[i] == exp[:8]).all(), i
    tau = 0x1234_5678_9ABC % F.R_MOD
    big = msm.ParamsKZG(12, tau, fixed_base=False)
    big.downsize(9)
    small = msm.ParamsKZG(9, tau, fixed_base=False)
    assert big.n == 512 and (dv.to_host(big.g_device, cols=8) == dv.to_host(small.g_device, cols=8)).all()
    assert (dv.to_host(big.g_lagrange_device, cols=8) == dv.to_host(small.g_lagrange_device, cols=8)).all()
    parts = msm.ParamsKZG.from_parts(9, small.g_device, None, small.g2, small.s_g2, fixed_base=False)
    assert (dv.to_host(parts.g_lagrange_device, cols=8) == dv.to_host(small.g_lagrange_device, cols=8)).all()
    poly = rand_fr(co, 512, 3)
    assert (parts.commit_lagrange(poly) == small.commit_lagrange(poly)).all()


def test_msm_small_scalars_empty_windows(lib, co):
    """Lookup-style scalars (values below 2^16, sorted, most of a table column zero)
    leave every window above the second empty: long runs of empty buckets between
    filled ones, single and batched (where the empty runs span jobs), against the oracle."""
    from halo2_amd import device as dv, fields as F, msm

    n = 1 << 14
    rng = np.random.default_rng(515)
    small = np.sort(rng.integers(0, 1 << 16, n))
    table = np.zeros(n, dtype=np.int64)
    table[-(n // 10):] = np.arange(n // 10)
    cols = [np.stack([F.mont_limbs(int(v)) for v in vals]) for vals in (small, table)]
    ds = msm.random_fr(n, 516)
    pts = msm.g1_mul_generator(ds, n)
    hp = dv.to_host(pts, cols=8)
    print("msm small scalars: single", flush=True)
    for c in cols:
        assert (msm.multiexp_device_scalars_device_bases(dv.to_device(c), pts, n) == co.msm(c, hp)).all()
    print("msm small scalars: batch", flush=True)
    jobs = [rand_fr(co, n, 517), cols[0], cols[1], rand_fr(co, n, 518), cols[0]]
    got = msm.multiexp_batch_device([dv.to_device(c) for c in jobs], pts)
    for j, c in enumerate(jobs):
        assert (got[j] == co.msm(c, hp)).all(), f"job {j}"


def test_kzg_table_width_guard(lib, co):
    """A fixed-base table is only valid under the width it was built with: after the
    table width changes, ParamsKZG refuses to commit through the old table."""
    from halo2_amd import autotune, msm

    k = 10
    params = msm.ParamsKZG(k, 4242)
    poly = msm.random_fr(1 << k, 4243)
    ref = params.commit_device(poly)
    assert (msm.ParamsKZG(k, 4242, fixed_base=False).commit_device(poly) == ref).all()
    try:
        autotune.set_msm_widths({}, {k: 7})
        with pytest.raises(RuntimeError):
            params.commit_device(poly)
        fresh = msm.ParamsKZG(k, 4242)  # built under width 7: the same commitment
        assert (fresh.commit_device(poly) == ref).all()
        assert (fresh.commit_many_device([poly, poly[: 4 * 100]])[0] == ref).all()
    finally:
        autotune.set_msm_widths({})
