"""Custom-gate quotient (the `_halo2_quotient*` rule interpreter).

CPU: the expression encoder (mirror of GraphEvaluator, src/plonk/evaluation.rs:362-700)
and the oracle interpreter are pinned together against a direct big-integer
evaluation of the gate expressions folded with Horner in y (the semantics of
src/cpu/evaluator.rs:150-411).
GPU: all three entry points bit-exact against the oracle, with row sub-ranges,
the per-row Horner seed, wrapped rotations and more rules than fit in LDS.
"""
import random

import numpy as np
import pytest

from conftest import rand_fr


def _circuit(n_extra_gates=0, only_standard=False):
    from halo2_amd.quotient import Advice, Challenge, Constant, Fixed, Instance, Negated, Product, Scaled, Sum

    # the standard gate q_a a + q_b b + q_m a b + q_c c (fixed 0..3, advice 0..2)
    a, b, c = Advice(0), Advice(1), Advice(2)
    std = Sum(Sum(Sum(Product(Fixed(0), a), Product(Fixed(1), b)), Product(Product(Fixed(2), a), b)),
              Product(Fixed(3), c))
    # rotations (next / previous / far), subtraction, negation, doubling, squaring, challenge, instance
    g2 = Sum(Product(Advice(0, 1), Advice(0, 1)), Negated(Product(Scaled(Advice(1, -1), 5), Challenge(0))))
    g3 = Sum(Product(Constant(2), Instance(0, 0)), Negated(Sum(Advice(2, 3), Constant(7))))
    g4 = Product(Sum(Fixed(4, -2), Negated(Advice(0))), Sum(Challenge(1), Constant(1)))
    polys = [std] if only_standard else [std, g2, g3, g4]
    rnd = random.Random(5)
    for i in range(n_extra_gates):
        x = Sum(Product(Advice(i % 3, rnd.randint(-3, 3)), Fixed(i % 5)), Scaled(Instance(0, rnd.randint(-2, 2)),
                                                                                  rnd.randrange(1, 1000)))
        polys.append(Product(x, Sum(Advice((i + 1) % 3, 0), Challenge(i % 2))))
    return polys


def _direct(polys, fixed, instance, advice, challenges, y, nq, rot_scale, rows, prev=None):
    """Big-integer evaluation of the expressions, Horner-folded in y, row by row."""
    from halo2_amd import fields as F

    R = F.R_MOD
    vals = {}

    def col(arrs, c, rot, r):
        key = (id(arrs), c)
        if key not in vals:
            vals[key] = [F.value_of(x) for x in arrs[c]]
        return vals[key][(r + rot * rot_scale) % nq]

    def ev(e, r):
        k = e[0]
        if k == "const":
            return e[1]
        if k == "fixed":
            return col(fixed, e[1], e[2], r)
        if k == "advice":
            return col(advice, e[1], e[2], r)
        if k == "instance":
            return col(instance, e[1], e[2], r)
        if k == "challenge":
            return challenges[e[1]]
        if k == "neg":
            return -ev(e[1], r) % R
        if k == "sum":
            return (ev(e[1], r) + ev(e[2], r)) % R
        if k == "prod":
            return ev(e[1], r) * ev(e[2], r) % R
        if k == "scaled":
            return ev(e[1], r) * e[2] % R
        raise ValueError(k)

    out = []
    for i, r in enumerate(rows):
        v = prev[i] if prev is not None else 0
        for p in polys:
            v = (v * y + ev(p, r)) % R
        out.append(v)
    return out


def _inputs(co, nq, seed):
    fixed = [rand_fr(co, nq, seed + i) for i in range(5)]
    instance = [rand_fr(co, nq, seed + 10)]
    advice = [rand_fr(co, nq, seed + 20 + i) for i in range(3)]
    return fixed, instance, advice


def test_encoder_and_oracle_vs_direct(co):
    from halo2_amd import fields as F
    from halo2_amd.quotient import custom_gates_evaluator

    nq, rot_scale = 64, 4
    polys = _circuit(2)
    ev = custom_gates_evaluator(polys)
    rules, vp = ev.encode_for_device()
    # the encoder emits one Store per distinct query and de-duplicates repeats
    assert len(vp) == len(polys)
    fixed, instance, advice = _inputs(co, nq, 300)
    challenges = [123456789, 987654321]
    y = 0xABCDEF
    ch = np.stack([F.mont_limbs(c) for c in challenges])
    cs = np.stack([F.mont_limbs(c) for c in ev.constants])
    prev_vals = [random.Random(1).randrange(F.R_MOD) for _ in range(nq)]
    prev = np.stack([F.mont_limbs(v) for v in prev_vals])
    for (r0, r1, pv) in ((0, nq, None), (5, 40, prev[5:40])):
        got = co.quotient_rules(fixed, instance, advice, ch, cs, F.mont_limbs(y), rules, vp, nq, rot_scale, r0, r1,
                                prev=pv)
        exp = _direct(polys, fixed, instance, advice, challenges, y, nq, rot_scale, range(r0, r1),
                      prev=prev_vals[r0:r1] if pv is not None else None)
        assert [F.value_of(x) for x in got] == exp


def _gpu_case(lib, co, nq, rot_scale, n_extra, seed, r0=0, r1=None, with_prev=False, only_standard=False):
    from halo2_amd import device as dv, fields as F
    from halo2_amd.quotient import custom_gates_evaluator, quotient_device, quotient_host

    r1 = nq if r1 is None else r1
    polys = _circuit(n_extra, only_standard)
    ev = custom_gates_evaluator(polys)
    rules, vp = ev.encode_for_device()
    fixed, instance, advice = _inputs(co, nq, seed)
    challenges = [seed * 7 + 1, seed * 11 + 3]
    y = seed * 1000003 % F.R_MOD
    ch = np.stack([F.mont_limbs(c) for c in challenges])
    cs = np.stack([F.mont_limbs(c) for c in ev.constants])
    prev = rand_fr(co, r1 - r0, seed + 99) if with_prev else None
    exp = co.quotient_rules(fixed, instance, advice, ch, cs, F.mont_limbs(y), rules, vp, nq, rot_scale, r0, r1,
                            prev=prev)
    d = [[dv.to_device(x) for x in group] for group in (fixed, instance, advice)]
    got = quotient_device(ev, *d, challenges, y, nq, rot_scale, r0, r1,
                          prev=dv.to_device(prev) if with_prev else None)
    assert (dv.to_host(got) == exp).all(), "device columns, device out"
    if not with_prev:
        assert (quotient_host(ev, fixed, instance, advice, challenges, y, nq, rot_scale, r0, r1) == exp).all(), \
            "host columns"
        assert (quotient_host(ev, *d, challenges, y, nq, rot_scale, r0, r1, device_columns=True) == exp).all(), \
            "device columns, host out"
    return len(rules)


@pytest.mark.gpu
def test_quotient_rules_small(lib, co):
    nrules = _gpu_case(lib, co, 1 << 10, 4, 0, 11, only_standard=True)
    assert nrules <= 16  # the LDS-intermediates path
    _gpu_case(lib, co, 1 << 10, 4, 0, 12)


@pytest.mark.gpu
def test_quotient_rules_many_rules_global_path(lib, co):
    nrules = _gpu_case(lib, co, 1 << 12, 4, 12, 21)
    assert nrules > 16


@pytest.mark.gpu
def test_quotient_rules_subrange_and_prev(lib, co):
    _gpu_case(lib, co, 1 << 12, 4, 3, 31, r0=1000, r1=3000, with_prev=True)
    _gpu_case(lib, co, 1 << 12, 4, 3, 32, r0=4000, r1=4096)


@pytest.mark.gpu
def test_quotient_rules_large_part(lib, co):
    """One extended part at k = 18 (rot_scale 4 as for a 4-part extended domain)."""
    print("quotient rules nq=2^18", flush=True)
    _gpu_case(lib, co, 1 << 18, 4, 2, 41, with_prev=True)
