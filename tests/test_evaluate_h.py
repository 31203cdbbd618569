"""evaluate_h on the device (src/plonk/evaluation.rs:1056-1417 of the reference) against
the same part-by-part composition of the CPU oracle (src/cpu/evaluator.rs:150-411:
coset parts, custom-gate rules, permutation and lookup identities, interleave), and the
property that pins the wiring independently of the oracle: for a witness that satisfies
the gates, h = numerator / (X^n - 1) is a polynomial of degree <= 2n - 3, so after the
vanishing division and extended_to_coeff every coefficient from 2n - 2 up is zero."""
import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _gates():
    from halo2_amd.quotient import Advice, Challenge, Constant, Fixed, Instance, Negated, Product, Scaled, Sum

    a, b, c = Advice(0), Advice(1), Advice(2)
    std = Sum(Sum(Sum(Product(Fixed(0), a), Product(Fixed(1), b)), Product(Product(Fixed(2), a), b)),
              Product(Fixed(3), c))
    rot = Sum(Product(Advice(0, 1), Instance(0, -1)), Negated(Product(Scaled(Advice(1, -1), 5), Challenge(0))))
    cst = Sum(Product(Constant(3), Fixed(4, 1)), Negated(c))
    return [std, rot, cst]


def _coset_part(co, coeffs, k, omega, c):
    t = co.distribute_powers_inplace(np.array(coeffs, copy=True), c)
    return co.fft(t, k, omega)


def _oracle_evaluate_h(co, dom, gates_ev, fixed, l0, ll, la, advice, instance, chals, y, beta, gamma, theta,
                       perm, lookups, blinding, cs_degree):
    """The reference's CPU evaluate_h_inner loop restated over the oracle's pieces (one circuit)."""
    from halo2_amd import fields as F

    M = F.mont_limbs
    k, n = dom.k, dom.n
    w = M(dom.omega)
    w_inv = M(dom.omega_inv)
    rules, vp = gates_ev.encode_for_device()
    cs = np.stack([M(v) for v in gates_ev.constants])
    ch = np.stack([M(v) for v in chals]) if chals else np.zeros((0, 4), dtype=np.uint64)
    cur = 1
    parts = []
    for _ in range(dom.extended_len >> k):
        c = M(F.ZETA * cur % F.R_MOD)
        ext = lambda polys: [_coset_part(co, p, k, w, c) for p in polys]  # noqa: E731
        fx, adv, ins = ext(fixed), ext(advice), ext(instance)
        l0e, lle, lae = ext([l0, ll, la])
        values = co.quotient_rules(fx, ins, adv, ch, cs, M(y), rules, vp, n, 1, 0, n)
        if perm is not None:
            zs = ext(perm["products"])
            sig = ext(perm["sigmas"])
            groups = {"fixed": fx, "advice": adv, "instance": ins}
            cols = [groups[kind][i] for kind, i in perm["columns"]]
            values = co.permutation_quotient(values, zs, sig, cols, l0e, lle, lae, M(beta), M(gamma), M(y),
                                             M(F.DELTA), M(beta * F.ZETA % F.R_MOD), M(cur), w, cs_degree - 2,
                                             -(blinding + 1), 1, n)
        for lk in lookups:
            cst, lr, lvp = lk["graph"].encode_for_device_with_runtime_constants(theta, beta, gamma, y)
            table = co.quotient_rules(fx, ins, adv, ch, np.stack([M(v) for v in cst]), M(0), lr, lvp, n, 1, 0, n)
            coeff = lambda lag: co.scale_inplace(co.fft(lag, k, w_inv), M(pow(n, -1, F.R_MOD)))  # noqa: E731
            pin = _coset_part(co, coeff(lk["pin"]), k, w, c)
            ptab = _coset_part(co, coeff(lk["ptab"]), k, w, c)
            prod = _coset_part(co, lk["prod"], k, w, c)
            values = co.quotient_lookups(values, table, prod, pin, ptab, l0e, lle, lae, M(beta), M(gamma), M(y))
        parts.append(values)
        cur = cur * dom.extended_omega % F.R_MOD
    return co.extended_from_lagrange_vec(parts)


@pytest.mark.parametrize("k", [6, 9])
def test_evaluate_h_vs_oracle(lib, co, k):
    from halo2_amd import device as dv, fields as F
    from halo2_amd.domain import EvaluationDomain
    from halo2_amd.evaluation import LookupCommitted, PermutationCommitted, evaluate_h_device
    from halo2_amd.quotient import Advice, Fixed, Product, custom_gates_evaluator, lookup_evaluator

    print(f"evaluate_h k={k}", flush=True)
    n = 1 << k
    dom = EvaluationDomain(5, k)  # quotient degree 4: four parts
    seed = [1000 * k]

    def rnd(cnt):
        out = [rand_fr(co, n, seed[0] + i) for i in range(cnt)]
        seed[0] += cnt
        return out

    fixed, advice, instance = rnd(5), rnd(3), rnd(1)
    l0, ll, la = rnd(3)
    perm = {"products": rnd(3), "sigmas": rnd(5),
            "columns": [("advice", 0), ("advice", 1), ("advice", 2), ("fixed", 0), ("instance", 0)]}
    lk_graph = lookup_evaluator([Advice(0), Product(Advice(1), Fixed(2))], [Fixed(4), Fixed(3)])
    lk_prod, lk_pin, lk_ptab = rnd(3)
    chals = [0xC0FFEE]
    y, beta, gamma, theta = 0x1234567, 0xBEEF01, 0xFACADE, 0x7E7A
    blinding, cs_degree = 3, 4
    gates = custom_gates_evaluator(_gates())
    exp = _oracle_evaluate_h(co, dom, gates, fixed, l0, ll, la, advice, instance, chals, y, beta, gamma, theta,
                             perm, [{"graph": lk_graph, "prod": lk_prod, "pin": lk_pin, "ptab": lk_ptab}],
                             blinding, cs_degree)
    D = dv.to_device
    got = evaluate_h_device(
        dom, gates, [D(x) for x in fixed], D(l0), D(ll), D(la), [D(x) for x in advice], [D(x) for x in instance],
        chals, y, beta, gamma, theta,
        lookups=[LookupCommitted(lk_graph, D(lk_prod), D(lk_pin), D(lk_ptab))],
        permutation=PermutationCommitted([D(x) for x in perm["products"]], perm["columns"],
                                         [D(x) for x in perm["sigmas"]]),
        blinding_factors=blinding, cs_degree=cs_degree)
    assert (dv.to_host(got) == exp).all()
    # the proving key's cached coset parts (filled, then reused) and the permuted
    # columns given in coefficient form give the same values
    key_parts = {}
    pin_c, ptab_c = (dom.lagrange_to_coeff_device_input(D(x)) for x in (lk_pin, lk_ptab))
    for _ in range(2):
        again = evaluate_h_device(
            dom, gates, [D(x) for x in fixed], D(l0), D(ll), D(la), [D(x) for x in advice],
            [D(x) for x in instance], chals, y, beta, gamma, theta,
            lookups=[LookupCommitted(lk_graph, D(lk_prod), D(lk_pin), D(lk_ptab), pin_c, ptab_c)],
            permutation=PermutationCommitted([D(x) for x in perm["products"]], perm["columns"],
                                             [D(x) for x in perm["sigmas"]]),
            blinding_factors=blinding, cs_degree=cs_degree, key_parts=key_parts)
        assert len(key_parts) == 4 and (dv.to_host(again) == exp).all()


def test_evaluate_h_gate_degree(lib, co):
    """A witness satisfying q_a a + q_b b + q_m a b - c = 0 on every row gives h of degree
    <= 2n - 3; one wrong cell makes the high coefficients non-zero."""
    from halo2_amd import device as dv, fields as F
    from halo2_amd.domain import EvaluationDomain
    from halo2_amd.evaluation import evaluate_h_device
    from halo2_amd.quotient import Advice, Fixed, Product, Sum, custom_gates_evaluator

    k = 10
    n = 1 << k
    dom = EvaluationDomain(5, k)
    qa, qb, qm, a, b = (rand_fr(co, n, 7000 + i) for i in range(5))
    qc = np.tile(F.mont_limbs(F.R_MOD - 1), (n, 1))
    mul = lambda x, z: co.field_mul(0, x, z)  # noqa: E731
    add = lambda x, z: co.field_add(0, x, z)  # noqa: E731
    c = add(add(mul(qa, a), mul(qb, b)), mul(mul(qm, a), b))
    gate = Sum(Sum(Sum(Product(Fixed(0), Advice(0)), Product(Fixed(1), Advice(1))),
                   Product(Product(Fixed(2), Advice(0)), Advice(1))), Product(Fixed(3), Advice(2)))
    ev = custom_gates_evaluator([gate])
    zero = dv.to_device(np.zeros((n, 4), dtype=np.uint64))

    def high_coeffs(c_col):
        cols = dom.lagrange_to_coeff_many_device([qa, qb, qm, qc, a, b, c_col])
        num = evaluate_h_device(dom, ev, cols[:4], zero, zero, zero, cols[4:], [], [], 5, 0, 0, 0)
        h = dom.extended_to_coeff_device(dom.divide_by_vanishing_poly_device(num))
        return dv.to_host(h)[2 * n - 2:]

    assert not high_coeffs(c).any()
    bad = c.copy()
    bad[17] = add(bad[17:18], np.tile(F.mont_limbs(1), (1, 1)))[0]
    assert high_coeffs(bad).any()
