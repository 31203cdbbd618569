"""Checker: the BN254 optimal ate pairing in Python big integers, written from the
published construction (Miller's algorithm over the loop count 6u + 2 with the two
Frobenius correction lines, then the final exponentiation by (p^12 - 1)/r), for the
KZG verifier's final check e(left, [tau]G2) = e(right, G2)
(src/poly/kzg/msm.rs DualMSM::check; src/poly/kzg/multiopen/shplonk/verifier.rs:131-138).

Representation: Fq12 = Fq[w]/(w^12 - 18 w^6 + 82). With Fq2 = Fq[u]/(u^2 + 1) and the
sextic twist E': y^2 = x^3 + 3/(9 + u), w^6 = 9 + u, so u = w^6 - 9 and
(w^6 - 9)^2 = -1 gives the modulus. A G2 point (x', y') maps to (x' w^2, y' w^3) on
y^2 = x^3 + 3 over Fq12. The Miller loop keeps its running point on the twist (Fq2
arithmetic) and evaluates each line at P in Fq12: for slope m' on the twist the line
through (x', y') is -y_P + m' x_P w + (y' - m' x') w^3. Slow (seconds per pairing
product) but exact; pinned by bilinearity and non-degeneracy in
tests/test_verifier_host.py.
"""
from __future__ import annotations

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
BN_U = 4965661367192848881
ATE_LOOP = 6 * BN_U + 2

# ------------------------------------------------------------------ Fq2


def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_scale(a, s):
    return (a[0] * s % P, a[1] * s % P)


def f2_inv(a):
    t = pow((a[0] * a[0] + a[1] * a[1]) % P, -1, P)
    return (a[0] * t % P, -a[1] * t % P)


def f2_conj(a):
    return (a[0], -a[1] % P)


# ------------------------------------------------------------------ Fq12 (12 coefficients in w)


def f12_one():
    return [1] + [0] * 11


def f12_mul(a, b):
    t = [0] * 23
    for i, ai in enumerate(a):
        if ai:
            for j, bj in enumerate(b):
                if bj:
                    t[i + j] += ai * bj
    for d in range(22, 11, -1):  # w^12 = 18 w^6 - 82
        c = t[d]
        if c:
            t[d - 6] += 18 * c
            t[d - 12] -= 82 * c
    return [v % P for v in t[:12]]


def f12_pow(a, e: int):
    out = f12_one()
    base = a
    while e:
        if e & 1:
            out = f12_mul(out, base)
        base = f12_mul(base, base)
        e >>= 1
    return out


def f12_from_f2_at(c, power: int):
    """The Fq2 element c placed at w^power: c0 + c1 u = (c0 - 9 c1) + c1 w^6."""
    out = [0] * 12
    out[power] = (c[0] - 9 * c[1]) % P
    hi = power + 6
    if hi < 12:
        out[hi] = (out[hi] + c[1]) % P
    else:  # w^(power + 6) = w^(power - 6) w^12 = w^(power - 6) (18 w^6 - 82)
        out[power] = (out[power] + 18 * c[1]) % P
        out[power - 6] = (out[power - 6] - 82 * c[1]) % P
    return out


# ------------------------------------------------------------------ twist arithmetic


def g2_double(t):
    x, y = t
    m = f2_mul(f2_scale(f2_mul(x, x), 3), f2_inv(f2_scale(y, 2)))
    x3 = f2_sub(f2_mul(m, m), f2_scale(x, 2))
    return m, (x3, f2_sub(f2_mul(m, f2_sub(x, x3)), y))


def g2_add_pts(t, q):
    (x1, y1), (x2, y2) = t, q
    m = f2_mul(f2_sub(y2, y1), f2_inv(f2_sub(x2, x1)))
    x3 = f2_sub(f2_sub(f2_mul(m, m), x1), x2)
    return m, (x3, f2_sub(f2_mul(m, f2_sub(x1, x3)), y1))


def line_at(m, t, p):
    """The line of slope m (twist coordinates) through t, evaluated at the G1 point p."""
    xt, yt = t
    xp, yp = p
    out = [0] * 12
    out[0] = -yp % P
    for pw, c in ((1, f2_scale(m, xp)), (3, f2_sub(yt, f2_mul(m, xt)))):
        v = f12_from_f2_at(c, pw)
        out = [(a + b) % P for a, b in zip(out, v)]
    return out


# Frobenius on the twist: pi(x', y') = (conj(x') g2, conj(y') g3) with
# g2 = (9 + u)^((p - 1)/3), g3 = (9 + u)^((p - 1)/2) (the twist's coordinates scale by
# w^2 and w^3, and w^(p - 1) = (w^6)^((p - 1)/6)).
def _f2_pow(a, e):
    out, base = (1, 0), a
    while e:
        if e & 1:
            out = f2_mul(out, base)
        base = f2_mul(base, base)
        e >>= 1
    return out


XI = (9, 1)
FROB_X = _f2_pow(XI, (P - 1) // 3)
FROB_Y = _f2_pow(XI, (P - 1) // 2)


def frobenius(q):
    x, y = q
    return f2_mul(f2_conj(x), FROB_X), f2_mul(f2_conj(y), FROB_Y)


def miller_loop(p, q):
    """f_{6u+2, Q}(P) times the two correction lines (optimal ate), before the final
    exponentiation. p: G1 affine (ints), q: G2 affine over Fq2."""
    if p is None or q is None:
        return f12_one()
    f = f12_one()
    t = q
    for bit in bin(ATE_LOOP)[3:]:
        m, t2 = g2_double(t)
        f = f12_mul(f12_mul(f, f), line_at(m, t, p))
        t = t2
        if bit == "1":
            m, t2 = g2_add_pts(t, q)
            f = f12_mul(f, line_at(m, t, p))
            t = t2
    q1 = frobenius(q)
    q2 = frobenius(q1)
    nq2 = (q2[0], f2_scale(q2[1], P - 1))
    m, t2 = g2_add_pts(t, q1)
    f = f12_mul(f, line_at(m, t, p))
    t = t2
    m, _ = g2_add_pts(t, nq2)
    return f12_mul(f, line_at(m, t, p))


FINAL_EXP = (P ** 12 - 1) // R


def final_exponentiation(f):
    return f12_pow(f, FINAL_EXP)


def pairing(p, q):
    return final_exponentiation(miller_loop(p, q))


def pairing_product_is_one(pairs) -> bool:
    """prod e(p_i, q_i) == 1 (one shared final exponentiation)."""
    f = f12_one()
    for p, q in pairs:
        f = f12_mul(f, miller_loop(p, q))
    return final_exponentiation(f) == f12_one()


def kzg_check(left, s_g2, right, g2) -> bool:
    """DualMSM::check: e(left, s_g2) e(right, -g2) == 1."""
    neg_g2 = (g2[0], f2_scale(g2[1], P - 1))
    return pairing_product_is_one([(left, s_g2), (right, neg_g2)])
