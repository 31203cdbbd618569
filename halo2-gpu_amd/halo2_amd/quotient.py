"""Custom-gate quotient evaluation over the C ABI.

`GraphEvaluator` mirrors src/plonk/evaluation.rs:362-700 of the reference: it
compiles gate expressions into de-duplicated calculations over value sources
(constants, intermediates, fixed/advice/instance queries with rotations,
challenges) and encodes them into the 128-bit CalcRule / 64-bit value-part rule
words the `_halo2_quotient*` entry points interpret (evaluation.rs:107-350).

Expressions are small tuples, built with the helpers below:
    Constant(v), Fixed(col, rot), Advice(col, rot), Instance(col, rot), Challenge(i),
    Negated(e), Sum(a, b), Product(a, b), Scaled(e, k)
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import R_MOD, mont_limbs

# ---------------------------------------------------------------- expressions


def Constant(v):
    return ("const", v % R_MOD)


def Fixed(col, rot=0):
    return ("fixed", col, rot)


def Advice(col, rot=0):
    return ("advice", col, rot)


def Instance(col, rot=0):
    return ("instance", col, rot)


def Challenge(i):
    return ("challenge", i)


def Negated(e):
    return ("neg", e)


def Sum(a, b):
    return ("sum", a, b)


def Product(a, b):
    return ("prod", a, b)


def Scaled(e, k):
    return ("scaled", e, k % R_MOD)


# ValueSource variants in declaration order (the derived PartialOrd orders by variant, then fields)
_VS_ORDER = {"Constant": 0, "Intermediate": 1, "Fixed": 2, "Advice": 3, "Instance": 4, "Challenge": 5,
             "Beta": 6, "Gamma": 7, "Theta": 8, "Y": 9, "PreviousValue": 10}


def _vs(kind, *fields):
    return (_VS_ORDER[kind], kind) + tuple(fields)


def _vs_key(v):
    return (v[0],) + tuple(v[2:])


class GraphEvaluator:
    """evaluation.rs:362-700 (constants start as [0, 1, 2])."""

    def __init__(self):
        self.constants = [0, 1, 2]
        self.rotations: list[int] = []
        self.calculations: list[tuple] = []  # (calculation, target)
        self.num_intermediates = 0

    def add_rotation(self, rot: int) -> int:
        if rot in self.rotations:
            return self.rotations.index(rot)
        self.rotations.append(rot)
        return len(self.rotations) - 1

    def add_constant(self, c: int):
        c %= R_MOD
        if c in self.constants:
            return _vs("Constant", self.constants.index(c))
        self.constants.append(c)
        return _vs("Constant", len(self.constants) - 1)

    def add_calculation(self, calc):
        for existing, target in self.calculations:
            if existing == calc:
                return _vs("Intermediate", target)
        target = self.num_intermediates
        self.calculations.append((calc, target))
        self.num_intermediates += 1
        return _vs("Intermediate", target)

    def add_expression(self, e):
        C0, C1, C2 = _vs("Constant", 0), _vs("Constant", 1), _vs("Constant", 2)
        k = e[0]
        if k == "const":
            return self.add_constant(e[1])
        if k in ("fixed", "advice", "instance"):
            kind = {"fixed": "Fixed", "advice": "Advice", "instance": "Instance"}[k]
            return self.add_calculation(("Store", _vs(kind, e[1], self.add_rotation(e[2]))))
        if k == "challenge":
            return self.add_calculation(("Store", _vs("Challenge", e[1])))
        if k == "neg":
            if e[1][0] == "const":
                return self.add_constant(-e[1][1])
            a = self.add_expression(e[1])
            return a if a == C0 else self.add_calculation(("Negate", a))
        if k == "sum":
            a_e, b_e = e[1], e[2]
            if b_e[0] == "neg":
                a = self.add_expression(a_e)
                b = self.add_expression(b_e[1])
                if a == C0:
                    return self.add_calculation(("Negate", b))
                if b == C0:
                    return a
                return self.add_calculation(("Sub", a, b))
            a = self.add_expression(a_e)
            b = self.add_expression(b_e)
            if a == C0:
                return b
            if b == C0:
                return a
            if _vs_key(a) <= _vs_key(b):
                return self.add_calculation(("Add", a, b))
            return self.add_calculation(("Add", b, a))
        if k == "prod":
            a = self.add_expression(e[1])
            b = self.add_expression(e[2])
            if a == C0 or b == C0:
                return C0
            if a == C1:
                return b
            if b == C1:
                return a
            if a == C2:
                return self.add_calculation(("Double", b))
            if b == C2:
                return self.add_calculation(("Double", a))
            if a == b:
                return self.add_calculation(("Square", a))
            if _vs_key(a) <= _vs_key(b):
                return self.add_calculation(("Mul", a, b))
            return self.add_calculation(("Mul", b, a))
        if k == "scaled":
            f = e[2]
            if f == 0:
                return C0
            if f == 1:
                return self.add_expression(e[1])
            cst = self.add_constant(f)
            a = self.add_expression(e[1])
            return self.add_calculation(("Mul", a, cst))
        raise ValueError(f"unknown expression {e!r}")

    # ------------------------------------------------------------- encoding

    def _encode_vs(self, v) -> int:
        kind = v[1]
        if kind in ("Fixed", "Instance", "Advice"):
            src = {"Fixed": 0, "Instance": 1, "Advice": 2}[kind]
            idx, rot = v[2], self.rotations[v[3]]
        elif kind == "Intermediate":
            src, idx, rot = 3, v[2], 0
        elif kind == "Constant":
            src, idx, rot = 4, v[2], 0
        elif kind == "Challenge":
            src, idx, rot = 5, v[2], 0
        else:
            raise ValueError(f"{kind} is not encodable")
        sign = 1 if rot < 0 else 0
        return src | (idx << 4) | (abs(rot) << 24) | (sign << 39)

    def _encode_calc(self, calc) -> tuple[int, int]:
        dummy = 6
        op = calc[0]
        enc = self._encode_vs
        if op == "Add":
            a, b, c1, c2, d = enc(calc[1]), enc(calc[2]), 1, 1, 0
        elif op == "Sub":
            a, b, c1, c2, d = enc(calc[1]), enc(calc[2]), 1, 2, 0
        elif op == "Mul":
            a, b, c1, c2, d = enc(calc[1]), enc(calc[2]), 1, 0, 1
        elif op == "Negate":
            a, b, c1, c2, d = enc(calc[1]), dummy, 2, 0, 0
        elif op == "Store":
            a, b, c1, c2, d = enc(calc[1]), dummy, 1, 0, 0
        elif op == "Double":
            a, b, c1, c2, d = enc(calc[1]), dummy, 3, 0, 0
        elif op == "Square":
            a, b, c1, c2, d = enc(calc[1]), enc(calc[1]), 1, 0, 1
        else:
            raise ValueError(op)
        return a | (c1 << 40) | (d << 44), b | (c2 << 40)

    def encode_for_device(self) -> tuple[np.ndarray, np.ndarray]:
        """(rules: (n, 2) uint64, value_part_rules: (m,) uint64); the last calculation must be
        the terminal Horner (evaluation.rs:641-650)."""
        *body, (horner, _) = self.calculations
        assert horner[0] == "Horner"
        rules = np.array([self._encode_calc(c) for c, _ in body], dtype=np.uint64).reshape(-1, 2)
        vp = np.array([self._encode_vs(p) for p in horner[2]], dtype=np.uint64)
        return rules, vp


def custom_gates_evaluator(gate_polys) -> GraphEvaluator:
    """Evaluator::new_inner (evaluation.rs:395-410): all gate polynomials, then
    Horner(PreviousValue, parts, Y)."""
    g = GraphEvaluator()
    parts = [g.add_expression(p) for p in gate_polys]
    g.add_calculation(("Horner", _vs("PreviousValue"), tuple(parts), _vs("Y")))
    return g


# ---------------------------------------------------------------- launchers


def _fr_host(values) -> np.ndarray:
    if len(values) == 0:
        return np.zeros((1, 4), dtype=np.uint64)
    return np.ascontiguousarray(np.stack([mont_limbs(v) for v in values]))


def quotient_device(evaluator: GraphEvaluator, fixed, instance, advice, challenges, y: int, num_quotient: int,
                    rot_scale: int, row_start: int = 0, row_end: int | None = None,
                    prev: torch.Tensor | None = None) -> torch.Tensor:
    """`_halo2_quotient_device_columns_device_out` (evaluation.rs:916-1044 `quotient_device`):
    device columns in, device quotient values for rows [row_start, row_end) out."""
    lib = _lib.load()
    row_end = num_quotient if row_end is None else row_end
    rules, vp = evaluator.encode_for_device()
    rules = np.ascontiguousarray(rules)
    vp = np.ascontiguousarray(vp) if len(vp) else np.zeros(1, dtype=np.uint64)
    consts = _fr_host(evaluator.constants)
    chals = _fr_host(challenges)
    expr = _fr_host([0, 1, R_MOD - 1, 2, y])
    ptrs = [(ctypes.c_void_p * max(1, len(cols)))(*[dv.ptr(c) for c in cols]) for cols in (fixed, instance, advice)]
    out = dv.empty_u64(4 * (row_end - row_start))
    nr, nvp = len(rules), len(evaluator.calculations[-1][0][2])
    ws = int(lib._halo2_quotient_workspace_size(rules.ctypes.data, nr, nvp, num_quotient, len(fixed),
                                                len(instance), len(advice), len(challenges),
                                                len(evaluator.constants), rot_scale, row_start, row_end))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_quotient_device_columns_device_out(
        ptrs[0], len(fixed), ptrs[1], len(instance), ptrs[2], len(advice),
        _lib.obj(chals.ctypes.data), len(challenges), _lib.obj(consts.ctypes.data), len(evaluator.constants),
        _lib.obj(expr.ctypes.data), rules.ctypes.data, nr, vp.ctypes.data, nvp, dv.ptr(out), num_quotient,
        dv.ptr(prev) if prev is not None else None, rot_scale, row_start, row_end, dv.ptr(scratch), ws,
        dv.stream()), "_halo2_quotient_device_columns_device_out")
    return out


def quotient_host(evaluator: GraphEvaluator, fixed, instance, advice, challenges, y: int, num_quotient: int,
                  rot_scale: int, row_start: int = 0, row_end: int | None = None,
                  device_columns: bool = False) -> np.ndarray:
    """`_halo2_quotient` (host columns) or `_halo2_quotient_device_columns` (device columns), host output."""
    lib = _lib.load()
    row_end = num_quotient if row_end is None else row_end
    rules, vp = evaluator.encode_for_device()
    rules = np.ascontiguousarray(rules)
    nvp = len(vp)
    vp = np.ascontiguousarray(vp) if nvp else np.zeros(1, dtype=np.uint64)
    consts = _fr_host(evaluator.constants)
    chals = _fr_host(challenges)
    expr = _fr_host([0, 1, R_MOD - 1, 2, y])
    out = np.zeros((row_end - row_start, 4), dtype=np.uint64)
    ws = int(lib._halo2_quotient_workspace_size(rules.ctypes.data, len(rules), nvp, num_quotient, len(fixed),
                                                len(instance), len(advice), len(challenges),
                                                len(evaluator.constants), rot_scale, row_start, row_end))
    scratch = dv.empty_bytes(ws)
    common = (_lib.obj(chals.ctypes.data), len(challenges), _lib.obj(consts.ctypes.data), len(evaluator.constants),
              _lib.obj(expr.ctypes.data), rules.ctypes.data, len(rules), vp.ctypes.data, nvp,
              _lib.obj(out.ctypes.data), num_quotient, rot_scale, row_start, row_end, dv.ptr(scratch), ws,
              dv.stream())
    if device_columns:
        ptrs = [(ctypes.c_void_p * max(1, len(c)))(*[dv.ptr(x) for x in c]) for c in (fixed, instance, advice)]
        _lib.check(lib._halo2_quotient_device_columns(ptrs[0], len(fixed), ptrs[1], len(instance), ptrs[2],
                                                      len(advice), *common), "_halo2_quotient_device_columns")
    else:
        hosts = [[np.ascontiguousarray(x, dtype=np.uint64) for x in c] for c in (fixed, instance, advice)]
        objs = [_lib.obj_array([x.ctypes.data for x in h]) for h in hosts]
        _lib.check(lib._halo2_quotient(objs[0], len(fixed), objs[1], len(instance), objs[2], len(advice), *common),
                   "_halo2_quotient")
    return out
