"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong constant, a swapped operand, a wrong window) fails:
* BN254 constants from the curve seed u; Montgomery constants as powers of 2;
* the reference-held zeta constants (zeta.h:19-30) are cube roots of unity;
* the 2^28-th root of unity has exact order 2^28;
* C field ops vs Python big integers on random and edge values;
* the C radix-2 FFT vs the O(n^2) DFT in big integers, and vs the C naive DFT;
* the C G1 arithmetic vs affine big-integer formulas, r*G = O;
* the C Pippenger MSM vs the closed form (sum c_i s_i) G and vs per-point
  double-and-add, on the reference's own edge cases
  (tests/multiexp_dense_bucket_edge_equivalence.rs:48-281);
* products, scans and quotient identities vs independent big-integer loops;
* ChaCha20 vs the RFC 8439 section 2.3.2 known-answer block.
"""
import random

import numpy as np
import pytest

from conftest import rand_fr


def test_bn254_parameters(po):
    u = po.BN_U
    assert po.P == 36 * u**4 + 36 * u**3 + 24 * u**2 + 6 * u + 1
    assert po.R == po.P + 1 - (6 * u * u + 1)  # r = p + 1 - t
    for m in (po.P, po.R):
        for a in (2, 3, 5, 7, 11):
            assert pow(a, m - 1, m) == 1  # Fermat probable prime
    # Montgomery constants held by the reference (alt_bn128.hpp:19-47), recomputed
    assert (1 << 256) % po.P == 0x0E0A77C19A07DF2F666EA36F7879462C0A78EB28F5C70B3DD35D438DC58F0D9D
    assert (1 << 512) % po.P == 0x06D89F71CAB8351F47AB1EFF0A417FF6B5E71911D44501FBF32CFC5B538AFA89
    assert (1 << 256) % po.R == 0x0E0A77C19A07DF2F666EA36F7879462E36FC76959F60CD29AC96341C4FFFFFFB
    assert (1 << 512) % po.R == 0x0216D0B17F4E44A58C49833D53BB808553FE3AB1E35C59E31BB8E645AE216DA7
    assert (-pow(po.P, -1, 1 << 64)) % (1 << 64) == 0x87D20782E4866389
    assert (-pow(po.R, -1, 1 << 64)) % (1 << 64) == 0xC2E1F593EFFFFFFF


def test_roots_of_unity(po):
    z = po.zeta()
    assert z != 1 and pow(z, 3, po.R) == 1
    z2 = po.from_mont(po.limbs_to_int(po.ZETA_2_MONT_LIMBS))
    assert z2 == z * z % po.R
    w = po.root_of_unity()
    assert pow(w, 1 << 28, po.R) == 1 and pow(w, 1 << 27, po.R) == po.R - 1
    assert (po.R - 1) % (1 << 28) == 0 and ((po.R - 1) >> 28) % 2 == 1


def test_field_ops_vs_bigint(co, po):
    rng = random.Random(7)
    for m, fld in ((po.R, 0), (po.P, 1)):
        vals = [0, 1, 2, m - 1, m - 2, (1 << 253), m // 2] + [rng.randrange(m) for _ in range(50)]
        wals = vals[::-1]
        a = np.array([po.int_to_limbs(po.to_mont(v, m)) for v in vals], dtype=np.uint64)
        b = np.array([po.int_to_limbs(po.to_mont(v, m)) for v in wals], dtype=np.uint64)
        dec = lambda arr: [po.from_mont(po.limbs_to_int(r), m) for r in arr]
        assert dec(co.field_mul(fld, a, b)) == [x * y % m for x, y in zip(vals, wals)]
        assert dec(co.field_add(fld, a, b)) == [(x + y) % m for x, y in zip(vals, wals)]
        assert dec(co.field_sub(fld, a, b)) == [(x - y) % m for x, y in zip(vals, wals)]
        inv = dec(co.field_inv(fld, a))
        assert inv == [pow(x, -1, m) if x else 0 for x in vals]
        # every output is fully reduced
        for arr in (co.field_mul(fld, a, b), co.field_add(fld, a, b), co.field_sub(fld, a, b)):
            assert all(po.limbs_to_int(r) < m for r in arr)


def test_from_u512(co, po):
    rng = random.Random(3)
    raw = bytes(rng.getrandbits(8) for _ in range(64 * 20))
    out = co.fr_from_u512(raw)
    for i in range(20):
        v = int.from_bytes(raw[64 * i: 64 * i + 64], "little") % po.R
        assert po.fr_values(out[i])[0] == v


def test_chacha20_rfc8439_blocks(co):
    # RFC 8439 appendix A.1, test vectors #1 and #2: all-zero key and nonce, block
    # counters 0 and 1 (the nonce-0 case is where the rand_chacha counter layout and
    # RFC 8439 coincide).
    assert co.chacha_block([0] * 8, 0) == bytes.fromhex(
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert co.chacha_block([0] * 8, 1) == bytes.fromhex(
        "9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
        "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")


def test_fft_vs_naive_dft(co, po):
    for k in (1, 2, 3, 5, 6):
        n = 1 << k
        vals = [random.Random(k).randrange(po.R) for _ in range(n)]
        w = po.omega(k)
        got = po.fr_values(co.fft(po.fr_array(vals), k, po.fr_array([w])[0], 2))
        assert got == po.dft(vals, w)
    # larger: C FFT vs C naive DFT (independent loops)
    k = 9
    a = rand_fr(co, 1 << k, 5)
    w = po.fr_array([po.omega(k)])[0]
    assert (co.fft(a, k, w, 4) == co.dft_naive(a, k, w)).all()


def test_coset_forms(co, po):
    # coeff_to_extended and extended_to_coeff round trip, and parts == extended transform
    k, ext_k = 3, 5
    vals = [random.Random(11).randrange(po.R) for _ in range(1 << k)]
    ext = po.coeff_to_extended(vals, k, ext_k)
    back = po.extended_to_coeff(ext, ext_k, 1 << ext_k)
    assert back[: 1 << k] == vals and all(v == 0 for v in back[1 << k:])
    # CosetFFT_Part p equals every (2^(ext_k-k))-th... the parts interleave the extended values
    w_ext = po.omega(ext_k)
    m = 1 << (ext_k - k)
    for p in range(m):
        part = po.coset_part(vals, k, po.zeta() * pow(w_ext, p, po.R) % po.R)
        assert part == [ext[i * m + p] for i in range(1 << k)]


def test_g1_arithmetic(co, po):
    g = co.g1_generator()
    assert co.g1_is_on_curve(g)
    assert (g == po.g1_affine_array([po.G1_GEN])[0]).all()
    # r * G = O via repeated doubling in big ints, independent of the scalar reduction
    acc, add, s = None, po.G1_GEN, po.R
    while s:
        if s & 1:
            acc = po.g1_add(acc, add)
        add = po.g1_add(add, add)
        s >>= 1
    assert acc is None
    rng = random.Random(2)
    ss = [rng.randrange(po.R) for _ in range(6)] + [0, 1, po.R - 1]
    pts = co.g1_mul_gen(po.fr_array(ss), 2)
    exp = po.g1_affine_array([po.g1_mul(po.G1_GEN, s) for s in ss])
    assert (pts == exp).all()
    for p in pts:
        assert co.g1_is_on_curve(p)


def _dense_constant(po):
    # tests/multiexp_dense_bucket_edge_equivalence.rs:48-55 (canonical limbs)
    return po.limbs_to_int([0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9, 0x94D049BB133111EB, 0x022B5F6A3C1D0E8F])


def _ten_group_scalar(g):
    # multiexp_dense_bucket_edge_equivalence.rs:223-230
    b = bytes((((g + j) % 250) + 1) for j in range(31)) + b"\x00"
    return int.from_bytes(b, "little")


def test_msm_closed_form_and_edges(co, po):
    rng = random.Random(9)
    # random: sum c_i (s_i G) == (sum c_i s_i) G
    n = 300
    s = [rng.randrange(po.R) for _ in range(n)]
    c = [rng.randrange(po.R) for _ in range(n)]
    pts = co.g1_mul_gen(po.fr_array(s), 4)
    got = co.msm(po.fr_array(c), pts, 4)
    exp = po.g1_out12(po.g1_mul(po.G1_GEN, sum(x * y for x, y in zip(c, s)) % po.R))
    assert (got == exp).all()
    assert (co.msm(po.fr_array(c), pts, 1) == exp).all()
    # tiny n against per-point double-and-add (bucket widths c = 1 and 3)
    for m in (1, 3, 17):
        assert (co.msm(po.fr_array(c[:m]), pts[:m], 1) == co.msm_naive(po.fr_array(c[:m]), pts[:m])).all()
    # zero scalars + identity bases; all zero; single nonzero; dense constant; ten groups
    k = _dense_constant(po)
    cases = {
        "zeros_identity": ([0 if i % 3 == 0 else (k if i % 7 == 0 else c[i]) for i in range(n)],
                           [None if i % 5 == 0 else s[i] for i in range(n)]),
        "all_zero": ([0] * n, s),
        "single": ([k] + [0] * (n - 1), s),
        "dense": ([k] * (n // 4) + c[n // 4:], s),
        "ten_groups": ([_ten_group_scalar(g) for g in range(10) for _ in range(20)], s[:200]),
    }
    for name, (cs, ss) in cases.items():
        points = po.g1_affine_array([None if x is None else po.g1_mul(po.G1_GEN, x) for x in ss])
        total = sum(a * (b or 0) for a, b in zip(cs, ss)) % po.R
        exp = po.g1_out12(po.g1_mul(po.G1_GEN, total) if total else None)
        assert (co.msm(po.fr_array(cs), points, 3) == exp).all(), name


def test_products_vs_bigint(co, po):
    rng = random.Random(4)
    n = 64
    vals = [rng.randrange(po.R) for _ in range(n)]
    vals[5] = 0
    vals[17] = 0
    inv = po.fr_values(co.batch_invert(0, po.fr_array(vals)))
    assert inv == [pow(v, -1, po.R) if v else 0 for v in vals]
    pre = rng.randrange(po.R)
    gp = po.fr_values(co.grand_product(po.fr_array(vals), po.fr_array([pre])[0]))
    acc = pre
    exp = []
    for v in vals:
        acc = acc * v % po.R
        exp.append(acc)
    assert gp == exp
    # lookup product: (A + b)(S + g) / ((A' + b)(S' + g))
    A, S, Ap, Sp = ([rng.randrange(po.R) for _ in range(n)] for _ in range(4))
    b, g = rng.randrange(po.R), rng.randrange(po.R)
    got = po.fr_values(co.lookup_product(*(po.fr_array(x) for x in (Ap, Sp, A, S)), po.fr_array([b])[0],
                                         po.fr_array([g])[0]))
    exp = [(A[i] + b) * (S[i] + g) * pow((Ap[i] + b) * (Sp[i] + g), -1, po.R) % po.R for i in range(n)]
    assert got == exp
    # permutation product over 3 columns with deltaomega advanced by delta per column
    cols = [[rng.randrange(po.R) for _ in range(n)] for _ in range(3)]
    perms = [[rng.randrange(po.R) for _ in range(n)] for _ in range(3)]
    mod0 = [rng.randrange(po.R) for _ in range(n)]
    delta, omega, dw = (rng.randrange(po.R) for _ in range(3))
    got, dw_out = co.permutation_product(po.fr_array(mod0), [po.fr_array(p) for p in perms],
                                         [po.fr_array(v) for v in cols], *(po.fr_array([x])[0] for x in
                                                                           (b, g, delta, omega, dw)))
    exp = []
    for i in range(n):
        den = mod0[i]
        num = 1
        d = dw
        for c in range(3):
            den = den * (b * perms[c][i] + g + cols[c][i]) % po.R
            num = num * (d * pow(omega, i, po.R) * b + g + cols[c][i]) % po.R
            d = d * delta % po.R
        exp.append(num * pow(den, -1, po.R) % po.R)
    assert po.fr_values(got) == exp
    assert po.fr_values(dw_out)[0] == dw * pow(delta, 3, po.R) % po.R


def test_quotient_lookups_vs_bigint(co, po):
    rng = random.Random(12)
    n = 32
    arrs = [[rng.randrange(po.R) for _ in range(n)] for _ in range(8)]
    v0, table, z, a, s, l0, ll, la = arrs
    b, g, y = (rng.randrange(po.R) for _ in range(3))
    got = po.fr_values(co.quotient_lookups(*(po.fr_array(x) for x in arrs), *(po.fr_array([x])[0] for x in (b, g, y))))
    R = po.R
    exp = []
    for i in range(n):
        rn, rp = (i + 1) % n, (i - 1) % n
        v = v0[i]
        v = (v * y + (1 - z[i]) * l0[i]) % R
        v = (v * y + (z[i] * z[i] - z[i]) * ll[i]) % R
        v = (v * y + (z[rn] * (a[i] + b) * (s[i] + g) - z[i] * table[i]) * la[i]) % R
        v = (v * y + (a[i] - s[i]) * l0[i]) % R
        v = (v * y + (a[i] - s[i]) * (a[i] - a[rp]) * la[i]) % R
        exp.append(v)
    assert got == exp


def test_eval_poly_closed_forms(co, po):
    """Horner evaluation (halo2-axiom arithmetic::eval_polynomial): the geometric series
    sum u^i = (u^n - 1)/(u - 1), u = 0 gives a_0, u = 1 gives sum a_i, and a product of
    linear factors vanishes at each root while matching prod (z - r_j) elsewhere."""
    R = po.R
    rng = random.Random(11)
    for n in (1, 2, 7, 64, 257):
        u = rng.randrange(2, R)
        ones = po.fr_array([1] * n)
        got = po.fr_values(co.eval_poly(ones, po.fr_array([u])[0]).reshape(1, 4))[0]
        assert got == (pow(u, n, R) - 1) * pow(u - 1, -1, R) % R
        vals = [rng.randrange(R) for _ in range(n)]
        a = po.fr_array(vals)
        assert po.fr_values(co.eval_poly(a, po.fr_array([0])[0]).reshape(1, 4))[0] == vals[0]
        assert po.fr_values(co.eval_poly(a, po.fr_array([1])[0]).reshape(1, 4))[0] == sum(vals) % R
    assert po.fr_values(co.eval_poly(np.zeros((0, 4), dtype=np.uint64), po.fr_array([5])[0]).reshape(1, 4))[0] == 0
    roots = [rng.randrange(R) for _ in range(9)]
    coeffs = [1]
    for r in roots:  # multiply by (X - r)
        coeffs = [((coeffs[i - 1] if i > 0 else 0) - r * (coeffs[i] if i < len(coeffs) else 0)) % R
                  for i in range(len(coeffs) + 1)]
    a = po.fr_array(coeffs)
    for r in roots:
        assert po.fr_values(co.eval_poly(a, po.fr_array([r])[0]).reshape(1, 4))[0] == 0
    z = rng.randrange(R)
    expect = 1
    for r in roots:
        expect = expect * (z - r) % R
    assert po.fr_values(co.eval_poly(a, po.fr_array([z])[0]).reshape(1, 4))[0] == expect


def test_kate_division_identity(co, po):
    """kate_division (halo2-axiom arithmetic::kate_division): q (X - u) + a(u) = a(X)
    coefficient by coefficient, and X^n - u^n divides to sum u^(n-1-i) X^i."""
    R = po.R
    rng = random.Random(12)
    for n in (2, 3, 17, 200):
        for u in (0, 1, rng.randrange(R)):
            vals = [rng.randrange(R) for _ in range(n)]
            q = po.fr_values(co.kate_division(po.fr_array(vals), po.fr_array([u])[0]))
            assert len(q) == n - 1
            au = sum(v * pow(u, i, R) for i, v in enumerate(vals)) % R
            assert (au - u * q[0]) % R == vals[0]
            for i in range(1, n - 1):
                assert (q[i - 1] - u * q[i]) % R == vals[i]
            assert q[n - 2] == vals[n - 1]
        u = rng.randrange(R)
        xn = [(-pow(u, n, R)) % R] + [0] * (n - 1) + [1]
        q = po.fr_values(co.kate_division(po.fr_array(xn), po.fr_array([u])[0]))
        assert q == [pow(u, n - 1 - i, R) for i in range(n)]
    assert co.kate_division(po.fr_array([5]), po.fr_array([3])[0]).shape == (0, 4)


def _perm_case(rng, R, U, pattern):
    if U == 0:
        return [], []
    alpha = {"distinct": U, "mixed": max(1, U // 3), "same": 1, "small": min(U, 7)}[pattern]
    if pattern == "small":  # small canonical values: keys differ only in the low word
        alphabet = list(range(alpha))
    else:
        alphabet = [rng.randrange(R) for _ in range(alpha)]
    table = alphabet + [rng.choice(alphabet) for _ in range(U - alpha)]
    rng.shuffle(table)
    if pattern == "distinct":
        inp = alphabet[:]
        rng.shuffle(inp)
    else:
        inp = [rng.choice(alphabet) for _ in range(U)]
    return inp, table


def test_permute_expression_pair_seq(co, po):
    """The oracle against (1) the lookup argument's own constraints on the result:
    permuted input sorted and a permutation of the input, permuted table a permutation
    of the table, row 0 equal, and every row either equal or repeating the row above;
    and (2) a big-integer restatement of the reference's sequential algorithm (sorted,
    run heads consume one table occurrence, leftovers ascending from the last repeated
    row) on the reference test's patterns (src/cuda/tests.rs:2340-2548)."""
    R = po.R
    rng = random.Random(21)
    for U, n, pattern in ((1, 4, "distinct"), (13, 16, "distinct"), (100, 128, "mixed"), (50, 64, "same"),
                          (120, 128, "small"), (0, 8, "mixed")):
        inp, table = _perm_case(rng, R, U, pattern)
        pad_in = inp + [rng.randrange(R) for _ in range(n - U)]
        pad_tab = table + [rng.randrange(R) for _ in range(n - U)]
        pin, ptab = co.permute_expression_pair(po.fr_array(pad_in), po.fr_array(pad_tab), n, U)
        A, S = po.fr_values(pin), po.fr_values(ptab)
        assert A[U:] == [0] * (n - U) and S[U:] == [0] * (n - U)
        assert A[:U] == sorted(inp) and sorted(S[:U]) == sorted(table)
        for i in range(U):
            assert A[i] == S[i] or (i > 0 and A[i] == A[i - 1])
        if U:
            assert A[0] == S[0]
        # sequential restatement in big integers
        srt = sorted(inp)
        left = {}
        for v in table:
            left[v] = left.get(v, 0) + 1
        exp, rep = [0] * U, []
        for r, v in enumerate(srt):
            if r == 0 or v != srt[r - 1]:
                exp[r] = v
                left[v] -= 1
            else:
                rep.append(r)
        for v in sorted(left):
            for _ in range(left[v]):
                exp[rep.pop()] = v
        assert S[:U] == exp
    with pytest.raises(ValueError):
        co.permute_expression_pair(po.fr_array([1, 2]), po.fr_array([1, 1]), 2, 2)


def _perm_instance(po, rng, n, ncols, blind, cycles_per, break_cell=None):
    """A copy-constrained assignment over ncols columns of n rows: random cycles of 2..4
    cells among the usable rows, every cell of a cycle equal to its first cell, sigma
    values delta^col' omega^row' of the next cell of its cycle (permutation/keygen.rs).
    break_cell = (col, row) changes one cell after the copies."""
    R = po.R
    u = n - blind - 1
    w = po.omega(n.bit_length() - 1)
    delta = pow(7, 1 << 28, R)
    vals = [[rng.randrange(R) for _ in range(n)] for _ in range(ncols)]
    nxt = {(c, r): (c, r) for c in range(ncols) for r in range(n)}
    cells = [(c, r) for c in range(ncols) for r in range(u)]
    rng.shuffle(cells)
    i = 0
    for _ in range(cycles_per):
        ln = rng.randrange(2, 5)
        cyc = cells[i:i + ln]
        i += ln
        for j, (c, r) in enumerate(cyc):
            nxt[(c, r)] = cyc[(j + 1) % len(cyc)]
            vals[c][r] = vals[cyc[0][0]][cyc[0][1]]
    if break_cell is not None:
        c, r = break_cell
        vals[c][r] = (vals[c][r] + 1) % R
    sigma = [[pow(delta, nxt[(c, r)][0], R) * pow(w, nxt[(c, r)][1], R) % R for r in range(n)]
             for c in range(ncols)]
    return vals, sigma, w, delta, u, nxt


def _perm_products(po, vals, sigma, beta, gamma, w, delta, u, chunk_len, rng):
    """The grand products of permutation/prover.rs:96-255 in big integers: set s covers
    columns [s c, s c + c); z_0[0] = 1, z_s[0] = z_{s-1}[u]; z_s[i + 1] = z_s[i] *
    prod_j (v_j + beta delta^j omega^i + gamma) / (v_j + beta sigma_j + gamma) for i < u;
    rows after u random."""
    R = po.R
    n = len(vals[0])
    zs, carry = [], 1
    for s0 in range(0, len(vals), chunk_len):
        z = [0] * n
        z[0] = carry
        for i in range(u):
            num = den = 1
            for j in range(s0, min(s0 + chunk_len, len(vals))):
                num = num * (vals[j][i] + beta * pow(delta, j, R) * pow(w, i, R) + gamma) % R
                den = den * (vals[j][i] + beta * sigma[j][i] + gamma) % R
            z[i + 1] = z[i] * num * pow(den, -1, R) % R
        for i in range(u + 1, n):
            z[i] = rng.randrange(R)
        carry = z[u]
        zs.append(z)
    return zs


def test_permutation_quotient_multiset_vanishes_on_H(co, po):
    """src/cpu/evaluator.rs:63-147 over several sets, evaluated on H itself (rot_scale 1,
    l0 / l_last / l_active as indicator vectors): with grand products built from a valid
    copy-constrained assignment (5 columns, chunk_len 2: three sets, the last partial)
    every row of the quotient is zero -- the first-set start, the last-set closure, the
    cross-set carries l0 (z_i - z_{i-1}(omega^last X)) and the per-set identities with
    delta^j running across sets. A broken copy in the second set leaves the last set's
    product open (z_last(u) != 1), and a wrong carry rotation, a delta restarted per set
    or a dropped set term would each leave rows non-zero."""
    rng = random.Random(2024)
    R = po.R
    n, ncols, blind, chunk = 32, 5, 3, 2
    beta, gamma, y = (rng.randrange(R) for _ in range(3))

    def quotient(vals, sigma, zs, w, delta, u, scale_delta_start=1):
        l0 = [1] + [0] * (n - 1)
        ll = [0] * n
        ll[u] = 1
        la = [1 if i < u else 0 for i in range(n)]
        arr = lambda v: po.fr_array(v)  # noqa: E731
        one = lambda v: po.fr_array([v])[0]  # noqa: E731
        got = co.permutation_quotient(arr([0] * n), [arr(z) for z in zs], [arr(s) for s in sigma],
                                      [arr(v) for v in vals], arr(l0), arr(ll), arr(la), one(beta), one(gamma),
                                      one(y), one(delta), one(beta * scale_delta_start), one(1), one(w), chunk,
                                      -(blind + 1), 1, n)
        return po.fr_values(got)

    vals, sigma, w, delta, u, nxt = _perm_instance(po, random.Random(77), n, ncols, blind, 12)
    # at least one cycle must cross sets for the carries to matter
    assert any(nxt[(c, r)][0] // chunk != c // chunk for c in range(ncols) for r in range(u))
    zs = _perm_products(po, vals, sigma, beta, gamma, w, delta, u, chunk, rng)
    assert len(zs) == 3 and zs[-1][u] == 1
    assert quotient(vals, sigma, zs, w, delta, u) == [0] * n
    # a broken copy in the second set: products rebuilt over the broken values do not close
    cyc_cell = next((c, r) for c in (2, 3) for r in range(u) if nxt[(c, r)] != (c, r))
    bvals, bsig, *_ = _perm_instance(po, random.Random(77), n, ncols, blind, 12, break_cell=cyc_cell)
    assert bsig == sigma and sum(a != b for va, vb in zip(vals, bvals) for a, b in zip(va, vb)) == 1
    bzs = _perm_products(po, bvals, bsig, beta, gamma, w, delta, u, chunk, random.Random(99))
    assert bzs[-1][u] != 1
    assert quotient(bvals, bsig, bzs, w, delta, u)[u] != 0
    # the carry across sets must come from z_{i-1}(omega^last X): a product chain whose second
    # set restarts at 1 breaks row 0
    bad = [list(z) for z in zs]
    ratio = pow(zs[1][0], -1, R)
    bad[1] = [v * ratio % R for v in zs[1]]
    bad[2] = [v * ratio % R for v in zs[2]]
    got = quotient(vals, sigma, bad, w, delta, u)
    assert got[0] != 0
    # delta^j restarted per set (delta_start = beta delta^0 for every set) is not what the
    # oracle computes: the same valid products leave active rows non-zero when built that way
    zs_restart = []
    carry = 1
    for s0 in range(0, ncols, chunk):
        z = [0] * n
        z[0] = carry
        for i in range(u):
            num = den = 1
            for jj, j in enumerate(range(s0, min(s0 + chunk, ncols))):
                num = num * (vals[j][i] + beta * pow(delta, jj, R) * pow(w, i, R) + gamma) % R
                den = den * (vals[j][i] + beta * sigma[j][i] + gamma) % R
            z[i + 1] = z[i] * num * pow(den, -1, R) % R
        for i in range(u + 1, n):
            z[i] = zs[len(zs_restart)][i]
        carry = z[u]
        zs_restart.append(z)
    assert any(v != 0 for v in quotient(vals, sigma, zs_restart, w, delta, u)[:u])
