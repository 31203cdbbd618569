"""Checker: a verifier for the proofs of halo2_amd.prover, restating halo2's KZG/SHPLONK
verifier in Python big integers (test infrastructure; nothing in halo2_amd imports it).

  verify_proof            src/plonk/verifier.rs:19-378 (QUERY_INSTANCE = false: the
                          instance values are absorbed as scalars and evaluated at x
                          through the Lagrange basis, verifier.rs:74-202)
  permutation             src/plonk/permutation/verifier.rs:37-257
  lookups                 src/plonk/lookup/verifier.rs:37-211
  vanishing               src/plonk/vanishing/verifier.rs:40-138
  SHPLONK                 src/poly/kzg/multiopen/shplonk/verifier.rs:29-139 with
                          construct_intermediate_sets, shplonk.rs:51-160
  final check             DualMSM: e(left, [tau]G2) = e(right, G2). Two forms are
                          offered: the pairing (bn254_pairing.py) and, since the SRS here
                          comes from a seeded tau, the equivalent group identity
                          tau * left = right in G1.

The transcript is replayed from the proof bytes with the prover's conventions
(Blake2b-512 "Halo2-Transcript", prefixes 0/1/2, 32-byte points: canonical x with
the parity of y in bit 255).
"""
from __future__ import annotations

import hashlib

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
DELTA = pow(7, 1 << 28, R)
G1 = (1, 2)


def omega(k: int) -> int:
    return pow(pow(7, (R - 1) >> 28, R), 1 << (28 - k), R)


class VerifyError(Exception):
    pass


# ------------------------------------------------------------------ G1 (Jacobian)


def _jac_double(p):
    x, y, z = p
    if z == 0 or y == 0:
        return (0, 1, 0)
    a = x * x % P
    b = y * y % P
    c = b * b % P
    d = 2 * ((x + b) * (x + b) - a - c) % P
    e = 3 * a % P
    x3 = (e * e - 2 * d) % P
    return x3, (e * (d - x3) - 8 * c) % P, 2 * y * z % P


def _jac_add(p, q):
    if p[2] == 0:
        return q
    if q[2] == 0:
        return p
    x1, y1, z1 = p
    x2, y2, z2 = q
    z1z1, z2z2 = z1 * z1 % P, z2 * z2 % P
    u1, u2 = x1 * z2z2 % P, x2 * z1z1 % P
    s1, s2 = y1 * z2 * z2z2 % P, y2 * z1 * z1z1 % P
    if u1 == u2:
        return _jac_double(p) if s1 == s2 else (0, 1, 0)
    h = (u2 - u1) % P
    i = 4 * h * h % P
    j = h * i % P
    rr = 2 * (s2 - s1) % P
    v = u1 * i % P
    x3 = (rr * rr - j - 2 * v) % P
    y3 = (rr * (v - x3) - 2 * s1 * j) % P
    z3 = ((z1 + z2) * (z1 + z2) - z1z1 - z2z2) * h % P
    return x3, y3, z3


def to_affine(p):
    if p[2] == 0:
        return None
    zi = pow(p[2], -1, P)
    return p[0] * zi * zi % P, p[1] * zi * zi * zi % P


def from_affine(a):
    return (0, 1, 0) if a is None else (a[0], a[1], 1)


def g1_mul(a, s: int):
    """Affine point times a scalar (double-and-add)."""
    acc, base = (0, 1, 0), from_affine(a)
    s %= R
    while s:
        if s & 1:
            acc = _jac_add(acc, base)
        base = _jac_double(base)
        s >>= 1
    return acc


def msm_eval(terms) -> tuple | None:
    """sum s_i P_i for [(scalar, affine point)]."""
    acc = (0, 1, 0)
    for s, pt in terms:
        if pt is not None:
            acc = _jac_add(acc, g1_mul(pt, s))
    return to_affine(acc)


def on_curve(a) -> bool:
    return a is None or (a[1] * a[1] - a[0] ** 3 - 3) % P == 0


def decompress(data: bytes):
    v = int.from_bytes(data, "little")
    sign, x = v >> 255, v & ((1 << 255) - 1)
    if x >= P:
        raise VerifyError("point x not canonical")
    y2 = (x * x * x + 3) % P
    y = pow(y2, (P + 1) // 4, P)
    if y * y % P != y2:
        raise VerifyError("point not on the curve")
    if y & 1 != sign:
        y = P - y
    return x, y


# ------------------------------------------------------------------ transcript


class TranscriptRead:
    def __init__(self, proof: bytes):
        self.h = hashlib.blake2b(digest_size=64, person=b"Halo2-Transcript")
        self.data = proof
        self.pos = 0

    def _take(self, n: int) -> bytes:
        if self.pos + n > len(self.data):
            raise VerifyError("proof too short")
        out = self.data[self.pos:self.pos + n]
        self.pos += n
        return out

    def common_point(self, pt) -> None:
        if pt is None:
            raise VerifyError("point at infinity in the transcript")
        self.h.update(b"\x01" + pt[0].to_bytes(32, "little") + pt[1].to_bytes(32, "little"))

    def read_point(self):
        pt = decompress(self._take(32))
        self.common_point(pt)
        return pt

    def common_scalar(self, v: int) -> None:
        self.h.update(b"\x02" + (v % R).to_bytes(32, "little"))

    def read_scalar(self) -> int:
        v = int.from_bytes(self._take(32), "little")
        if v >= R:
            raise VerifyError("scalar not canonical")
        self.common_scalar(v)
        return v

    def squeeze(self) -> int:
        self.h.update(b"\x00")
        return int.from_bytes(self.h.copy().digest(), "little") % R


# -------------------------------------------------------------- expressions


def evaluate_expression(e, fixed_evals, advice_evals, instance_evals, cs, challenges=()):
    """Expression::evaluate over the query evaluations (query index = position of
    (column, rotation) in the constraint system's query lists)."""
    kind = e[0]
    if kind == "const":
        return e[1]
    if kind == "fixed":
        return fixed_evals[list(cs.fixed_queries).index((e[1], e[2]))]
    if kind == "advice":
        return advice_evals[list(cs.advice_queries).index((e[1], e[2]))]
    if kind == "instance":
        return instance_evals[list(cs.instance_queries).index((e[1], e[2]))]
    if kind == "challenge":
        return challenges[e[1]]
    rec = lambda a: evaluate_expression(a, fixed_evals, advice_evals, instance_evals, cs, challenges)  # noqa: E731
    if kind == "neg":
        return -rec(e[1]) % R
    if kind == "sum":
        return (rec(e[1]) + rec(e[2])) % R
    if kind == "prod":
        return rec(e[1]) * rec(e[2]) % R
    if kind == "scaled":
        return rec(e[1]) * e[2] % R
    raise ValueError(f"unknown expression {kind}")


# ------------------------------------------------------------------ SHPLONK


def _vanishing(points, z: int) -> int:
    acc = 1
    for p in points:
        acc = acc * (z - p) % R
    return acc


def _lagrange_eval(points, evals, z: int) -> int:
    """The interpolating polynomial of (points, evals) evaluated at z."""
    acc = 0
    for j, pj in enumerate(points):
        num = den = 1
        for m, pm in enumerate(points):
            if m != j:
                num = num * (z - pm) % R
                den = den * (pj - pm) % R
        acc = (acc + evals[j] * num * pow(den, -1, R)) % R
    return acc


def construct_intermediate_sets(queries):
    """shplonk.rs:51-160 over (commitment key, commitment terms, point, eval): commitments
    in first-appearance order with their point sets, then grouped by equal point sets
    (rotation sets in first-appearance order)."""
    per_commitment = []  # [key, terms, {point: eval}]
    super_points = []
    for key, terms, point, ev in queries:
        if point not in super_points:
            super_points.append(point)
        for entry in per_commitment:
            if entry[0] == key:
                entry[2][point] = ev
                break
        else:
            per_commitment.append([key, terms, {point: ev}])
    sets = []  # [points (set), [(terms, evals by point)]]
    for key, terms, evs in per_commitment:
        for s in sets:
            if s[0] == set(evs):
                s[1].append((terms, evs))
                break
        else:
            sets.append([set(evs), [(terms, evs)]])
    return sets, super_points


def shplonk_verify(tr: TranscriptRead, queries, g0):
    """VerifierSHPLONK::verify_proof: returns (left, right) affine points with the
    accepting condition e(left, [tau]G2) = e(right, G2)."""
    sets, super_points = construct_intermediate_sets(queries)
    y = tr.squeeze()
    v = tr.squeeze()
    h1 = tr.read_point()
    u = tr.squeeze()
    h2 = tr.read_point()
    outer = []  # (scalar, point)
    r_outer = 0
    z_0 = z_0_diff_inv = 0
    pv = 1
    for i, (points, commitments) in enumerate(sets):
        pts = sorted(points)
        z_diff = _vanishing([p for p in super_points if p not in points], u)
        if i == 0:
            z_0 = _vanishing(pts, u)
            z_0_diff_inv = pow(z_diff, -1, R)
            z_diff = 1
        else:
            z_diff = z_diff * z_0_diff_inv % R
        py = 1
        r_inner = 0
        inner = []
        for terms, evs in commitments:
            r_inner = (r_inner + py * _lagrange_eval(pts, [evs[p] for p in pts], u)) % R
            inner += [(py * s % R, pt) for s, pt in terms]
            py = py * y % R
        outer += [(s * pv * z_diff % R, pt) for s, pt in inner]
        r_outer = (r_outer + pv * r_inner * z_diff) % R
        pv = pv * v % R
    outer += [(-r_outer % R, g0), (-z_0 % R, h1), (u, h2)]
    return h2, msm_eval(outer)


# ------------------------------------------------------------------ verify_proof


def verify_proof(vk, instance_values, proof: bytes, tau: int | None = None, pairing_check=None, g2=None,
                 s_g2=None) -> bool:
    """verifier.rs:19-378 for one circuit with one instance column. Raises VerifyError on
    a malformed proof; returns whether the final check holds (tau * left == right when
    `tau` is given, or the pairing check `pairing_check(left, s_g2, right, g2)`)."""
    cs, k = vk.cs, vk.k
    n = 1 << k
    w = omega(k)
    w_inv = pow(w, -1, R)
    if len(instance_values) > n - (cs.blinding_factors + 1):
        raise VerifyError("instance too large")
    tr = TranscriptRead(proof)
    tr.common_scalar(vk.transcript_repr)
    for v in instance_values:
        tr.common_scalar(v)
    advice_c = [tr.read_point() for _ in range(cs.num_advice)]
    theta = tr.squeeze()
    nl = len(cs.lookups)
    permuted = [(tr.read_point(), tr.read_point()) for _ in range(nl)]
    beta = tr.squeeze()
    gamma = tr.squeeze()
    perm_c = [tr.read_point() for _ in range(cs.num_permutation_sets)]
    lookup_c = [tr.read_point() for _ in range(nl)]
    random_c = tr.read_point()
    y = tr.squeeze()
    h_c = [tr.read_point() for _ in range(cs.quotient_pieces)]
    x = tr.squeeze()
    xn = pow(x, n, R)
    rot = lambda r: x * pow(w if r >= 0 else w_inv, abs(r), R) % R  # noqa: E731

    def l_i(i: int) -> int:  # EvaluationDomain::l_i_range for one rotation
        wi = pow(w if i >= 0 else w_inv, abs(i), R)
        return wi * (xn - 1) * pow(n * (x - wi) % R, -1, R) % R

    # instance evaluations (QUERY_INSTANCE = false): sum_i v_i l_i(x omega^-rot)
    instance_evals = []
    for col, r in cs.instance_queries:
        vals = instance_values if col == 0 else []
        instance_evals.append(sum(v * l_i(i - r) for i, v in enumerate(vals)) % R)
    advice_evals = [tr.read_scalar() for _ in cs.advice_queries]
    fixed_evals = [tr.read_scalar() for _ in cs.fixed_queries]
    random_eval = tr.read_scalar()
    sigma_evals = [tr.read_scalar() for _ in cs.permutation_columns]
    perm_sets = []
    for s in range(cs.num_permutation_sets):
        ev, nx = tr.read_scalar(), tr.read_scalar()
        last = tr.read_scalar() if s + 1 < cs.num_permutation_sets else None
        perm_sets.append((ev, nx, last))
    lookup_evals = [tuple(tr.read_scalar() for _ in range(5)) for _ in range(nl)]

    # the vanishing argument: expected h(x) from every constraint at x
    bf = cs.blinding_factors
    l_last = l_i(-(bf + 1))
    l_blind = sum(l_i(-i) for i in range(1, bf + 1)) % R
    l_0 = l_i(0)
    active = (1 - l_last - l_blind) % R
    any_eval = {"advice": (advice_evals, cs.advice_queries), "fixed": (fixed_evals, cs.fixed_queries),
                "instance": (instance_evals, cs.instance_queries)}
    col_eval = lambda kind, idx: any_eval[kind][0][list(any_eval[kind][1]).index((idx, 0))]  # noqa: E731
    exprs = [evaluate_expression(g, fixed_evals, advice_evals, instance_evals, cs) for g in cs.gates]
    if perm_sets:
        exprs.append(l_0 * (1 - perm_sets[0][0]) % R)
        zl = perm_sets[-1][0]
        exprs.append((zl * zl - zl) * l_last % R)
        for s in range(1, len(perm_sets)):
            exprs.append((perm_sets[s][0] - perm_sets[s - 1][2]) * l_0 % R)
        c = cs.chunk_len
        for s, (ev, nx, _) in enumerate(perm_sets):
            cols = cs.permutation_columns[s * c:(s + 1) * c]
            left, right = nx, ev
            delta = beta * x * pow(DELTA, s * c, R) % R
            for j, (kind, idx) in enumerate(cols):
                val = col_eval(kind, idx)
                left = left * (val + beta * sigma_evals[s * c + j] + gamma) % R
                right = right * (val + delta + gamma) % R
                delta = delta * DELTA % R
            exprs.append((left - right) * active % R)
    for (inputs, tables), (z, z_next, a, a_prev, s_) in zip(cs.lookups, lookup_evals):
        compress = lambda es: _fold(  # noqa: E731
            [evaluate_expression(e, fixed_evals, advice_evals, instance_evals, cs) for e in es], theta)
        left = z_next * (a + beta) * (s_ + gamma) % R
        right = z * (compress(inputs) + beta) * (compress(tables) + gamma) % R
        exprs += [l_0 * (1 - z) % R, l_last * (z * z - z) % R, (left - right) * active % R,
                  l_0 * (a - s_) % R, (a - s_) * (a - a_prev) * active % R]
    h_eval = 0
    for e in exprs:
        h_eval = (h_eval * y + e) % R
    h_eval = h_eval * pow(xn - 1, -1, R) % R
    h_terms = [(pow(xn, i, R), c) for i, c in enumerate(h_c)]

    # the queries in verifier.rs:317-374 order
    x_next, x_prev, x_last = rot(1), rot(-1), rot(-(bf + 1))
    q = []
    for i, (col, r) in enumerate(cs.advice_queries):
        q.append((("advice", col), [(1, advice_c[col])], rot(r), advice_evals[i]))
    for s, (ev, nx, _) in enumerate(perm_sets):
        q.append((("perm", s), [(1, perm_c[s])], x, ev))
        q.append((("perm", s), [(1, perm_c[s])], x_next, nx))
    for s in reversed(range(len(perm_sets) - 1)):
        q.append((("perm", s), [(1, perm_c[s])], x_last, perm_sets[s][2]))
    for li, ((pin, ptab), zc, (z, z_next, a, a_prev, s_)) in enumerate(zip(permuted, lookup_c, lookup_evals)):
        q += [(("lz", li), [(1, zc)], x, z), (("lin", li), [(1, pin)], x, a), (("ltab", li), [(1, ptab)], x, s_),
              (("lin", li), [(1, pin)], x_prev, a_prev), (("lz", li), [(1, zc)], x_next, z_next)]
    for i, (col, r) in enumerate(cs.fixed_queries):
        q.append((("fixed", col), [(1, vk.fixed_commitments[col])], rot(r), fixed_evals[i]))
    for j, c in enumerate(vk.permutation_commitments):
        q.append((("sigma", j), [(1, c)], x, sigma_evals[j]))
    q.append((("h",), h_terms, x, h_eval))
    q.append((("random",), [(1, random_c)], x, random_eval))

    left, right = shplonk_verify(tr, q, G1)
    if tr.pos != len(proof):
        raise VerifyError("trailing bytes after the proof")
    ok = True
    if tau is not None:
        ok &= to_affine(g1_mul(left, tau)) == right
    if pairing_check is not None:
        ok &= pairing_check(left, s_g2, right, g2)
    return bool(ok)


def _fold(values, theta: int) -> int:
    acc = 0
    for v in values:
        acc = (acc * theta + v) % R
    return acc
