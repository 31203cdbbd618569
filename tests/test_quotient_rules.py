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


# ---------------------------------------------------------------- raw rule programs
# Hand-built CalcRules (ValueSource = src | idx << 4 | |rot| << 24 | sign << 39) that
# reach the program compiler's corners: many intermediates live at once (slots in LDS
# with one row per thread, or in the global buffer), rules nothing reads, Horner
# steps over constants only, and no value parts at all.

def _vs(src, idx=0, rot=0):
    return src | (idx << 4) | (abs(rot) << 24) | ((1 << 39) if rot < 0 else 0)


def _rule(a, b, c1, c2, d):
    return [a | (c1 << 40) | (d << 44), b | (c2 << 40)]


FIX, INS, ADV, INT, CON, CHA, DUM = range(7)


def _raw_call(lib, fixed, instance, advice, challenges, constants, y, rules, vp, nq, rot_scale, r0, r1, prev=None):
    import ctypes

    from halo2_amd import _lib, device as dv, fields as F

    rules = np.ascontiguousarray(np.array(rules, dtype=np.uint64).reshape(-1))
    nr = len(rules) // 2
    nvp = len(vp)
    vpa = np.ascontiguousarray(np.array(vp if nvp else [0], dtype=np.uint64))
    ch = np.ascontiguousarray(challenges if len(challenges) else np.zeros((1, 4), dtype=np.uint64))
    cs = np.ascontiguousarray(constants if len(constants) else np.zeros((1, 4), dtype=np.uint64))
    expr = np.stack([F.mont_limbs(v) for v in (0, 1, F.R_MOD - 1, 2, y)])
    d = [[dv.to_device(x) for x in g] for g in (fixed, instance, advice)]
    ptrs = [(ctypes.c_void_p * max(1, len(g)))(*[dv.ptr(c) for c in g]) for g in d]
    out = dv.empty_u64(4 * (r1 - r0))
    ws = int(lib._halo2_quotient_workspace_size(rules.ctypes.data if nr else None, nr, nvp, nq, len(fixed),
                                                len(instance), len(advice), len(challenges), len(constants),
                                                rot_scale, r0, r1))
    scratch = dv.empty_bytes(ws)
    d_prev = dv.to_device(prev) if prev is not None else None
    err = lib._halo2_quotient_device_columns_device_out(
        ptrs[0], len(fixed), ptrs[1], len(instance), ptrs[2], len(advice), _lib.obj(ch.ctypes.data), len(challenges),
        _lib.obj(cs.ctypes.data), len(constants), _lib.obj(expr.ctypes.data), rules.ctypes.data, nr,
        vpa.ctypes.data, nvp, dv.ptr(out), nq, dv.ptr(d_prev) if d_prev is not None else None, rot_scale, r0, r1,
        dv.ptr(scratch), ws, dv.stream())
    _lib.check(err, "_halo2_quotient_device_columns_device_out")
    return dv.to_host(out)


def _wide_program(width):
    """`width` products of rotated columns all computed first, then summed in a chain:
    `width` intermediates are live together before the chain starts. Two rules that
    nothing reads sit in the middle."""
    rules = []
    for i in range(width):
        rules += _rule(_vs(FIX, i % 5, (i % 7) - 3), _vs(ADV, i % 3, (i % 5) - 2), 1, 0, 1)
    rules += _rule(_vs(ADV, 0, 1), _vs(CHA, 1), 1, 1, 0)  # dead
    rules += _rule(_vs(INT, width), _vs(DUM), 3, 0, 0)  # dead (reads the dead one)
    acc = 0
    for i in range(1, width):
        rules += _rule(_vs(INT, acc), _vs(INT, i), 1, 2 if i % 4 == 3 else 1, 0)
        acc = len(rules) // 2 - 1
    rules += _rule(_vs(INT, acc), _vs(CON, 1), 1, 0, 1)
    last = len(rules) // 2 - 1
    rules += _rule(_vs(INT, last), _vs(DUM), 2, 0, 0)  # negation
    vp = [_vs(CON, 0), _vs(INT, last), _vs(INT, len(rules) // 2 - 1), _vs(INT, 0), _vs(CON, 1)]
    return rules, vp


@pytest.mark.gpu
@pytest.mark.parametrize("width,nq,r0,r1,with_prev", [
    (3, 1000, 0, 1000, False),  # two rows per thread; 1000 is not a multiple of the 128-row tile
    (12, 4096, 77, 3001, True),  # one row per thread, slots in LDS
    (40, 4096, 0, 4096, True),  # more live slots than LDS holds: global buffer
    (40, 1 << 17, 5, (1 << 17) - 3, False),  # global buffer, grid-stride over many tiles
])
def test_quotient_rules_live_slots(lib, co, width, nq, r0, r1, with_prev):
    from halo2_amd import fields as F

    print(f"raw program width={width} nq={nq} rows=[{r0},{r1})", flush=True)
    fixed, instance, advice = _inputs(co, nq, 500 + width)
    ch = rand_fr(co, 2, 7)
    cs = np.stack([F.mont_limbs(v) for v in (5, 0xC0FFEE)])
    y = 0x5A5A5A5A + width
    rules, vp = _wide_program(width)
    prev = rand_fr(co, r1 - r0, 9) if with_prev else None
    exp = co.quotient_rules(fixed, instance, advice, ch, cs, F.mont_limbs(y), rules, vp, nq, 4, r0, r1, prev=prev)
    got = _raw_call(lib, fixed, instance, advice, ch, cs, y, rules, vp, nq, 4, r0, r1, prev=prev)
    assert (got == exp).all()


@pytest.mark.gpu
def test_quotient_rules_degenerate_programs(lib, co):
    """No value parts (the seed or zero comes back), constants-only Horner, no rules."""
    from halo2_amd import fields as F

    nq = 300
    fixed, instance, advice = _inputs(co, nq, 700)
    ch = rand_fr(co, 1, 3)
    cs = np.stack([F.mont_limbs(v) for v in (3, 4)])
    prev = rand_fr(co, nq, 4)
    rules = _rule(_vs(FIX, 0), _vs(ADV, 1, -1), 1, 0, 1)
    for rl, vp, pv in ((rules, [], None), (rules, [], prev), ([], [_vs(CON, 1), _vs(CON, 0)], None),
                       ([], [_vs(CON, 1)], prev)):
        exp = co.quotient_rules(fixed, instance, advice, ch, cs, F.mont_limbs(77), rl, vp, nq, 2, 0, nq, prev=pv)
        got = _raw_call(lib, fixed, instance, advice, ch, cs, 77, rl, vp, nq, 2, 0, nq, prev=pv)
        assert (got == exp).all()


@pytest.mark.gpu
def test_quotient_rules_rejects_bad_operands(lib, co):
    """An operand naming a missing column, constant, challenge or a not-yet-computed
    intermediate fails on the host before any launch."""
    from halo2_amd import _lib, fields as F

    nq = 64
    fixed, instance, advice = _inputs(co, nq, 800)
    ch = rand_fr(co, 2, 3)
    cs = np.stack([F.mont_limbs(v) for v in (3, 4)])
    ok = _rule(_vs(FIX, 0), _vs(ADV, 0), 1, 0, 1)
    bad = [
        (ok + _rule(_vs(FIX, 5), _vs(ADV, 0), 1, 1, 0), [_vs(INT, 1)]),  # fixed 5 of 5
        (ok + _rule(_vs(INS, 1), _vs(ADV, 0), 1, 1, 0), [_vs(INT, 1)]),  # instance 1 of 1
        (ok + _rule(_vs(ADV, 0), _vs(ADV, 3), 1, 2, 0), [_vs(INT, 1)]),  # advice 3 of 3
        (ok + _rule(_vs(INT, 1), _vs(ADV, 0), 1, 1, 0), [_vs(INT, 1)]),  # reads itself
        (ok + _rule(_vs(CON, 2), _vs(ADV, 0), 1, 1, 0), [_vs(INT, 1)]),  # constant 2 of 2
        (ok + _rule(_vs(CHA, 2), _vs(ADV, 0), 1, 1, 0), [_vs(INT, 1)]),  # challenge 2 of 2
        (ok, [_vs(INT, 1)]),  # value part past the last rule
        (ok, [_vs(CON, 2)]),  # value part constant 2 of 2
    ]
    for rules, vp in bad:
        with pytest.raises(_lib.Halo2GpuError):
            _raw_call(lib, fixed, instance, advice, ch, cs, 5, rules, vp, nq, 1, 0, nq)
    # an operand the rule does not use is not checked (negation ignores b)
    _raw_call(lib, fixed, instance, advice, ch, cs, 5, ok + _rule(_vs(INT, 0), _vs(FIX, 99), 2, 0, 0),
              [_vs(INT, 1)], nq, 1, 0, nq)
