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

    def encode_for_device_with_runtime_constants(self, theta: int, beta: int, gamma: int, y: int):
        """evaluation.rs:695-815: (constants + [theta, beta, gamma, y], rules, value-part rules).
        Beta/Gamma/Theta/Y sources become Constant slots, internal Horners are unrolled into
        Mul/Add chains (intermediates renumbered), a non-Horner terminal becomes the single
        part of Horner(Constant(0), [term], Constant(0))."""
        constants = list(self.constants)
        slot = {}
        for name, v in (("Theta", theta), ("Beta", beta), ("Gamma", gamma), ("Y", y)):
            slot[name] = len(constants)
            constants.append(v % R_MOD)
        remap: list[int] = []
        new: list[tuple] = []

        def rw(v):
            if v[1] in slot:
                return _vs("Constant", slot[v[1]])
            if v[1] == "Intermediate":
                return _vs("Intermediate", remap[v[2]])
            return v

        *internal, (terminal, _) = self.calculations
        for calc, _target in internal:
            if calc[0] == "Horner":
                start, parts, factor = rw(calc[1]), calc[2], rw(calc[3])
                if not parts:
                    new.append(("Store", start))
                    remap.append(len(new) - 1)
                    continue
                acc = start
                for part in parts:
                    new.append(("Mul", acc, factor))
                    new.append(("Add", _vs("Intermediate", len(new) - 1), rw(part)))
                    acc = _vs("Intermediate", len(new) - 1)
                remap.append(len(new) - 1)
            else:
                new.append((calc[0],) + tuple(rw(a) for a in calc[1:]))
                remap.append(len(new) - 1)
        if terminal[0] == "Horner":
            parts = tuple(rw(p) for p in terminal[2])
        else:
            new.append((terminal[0],) + tuple(rw(a) for a in terminal[1:]))
            parts = (_vs("Intermediate", len(new) - 1),)
        rules = np.array([self._encode_calc(c) for c in new], dtype=np.uint64).reshape(-1, 2)
        vp = np.array([self._encode_vs(p) for p in parts], dtype=np.uint64)
        return constants, rules, vp

    def encode_for_device(self) -> tuple[np.ndarray, np.ndarray]:
        """(rules: (n, 2) uint64, value_part_rules: (m,) uint64); the last calculation must be
        the terminal Horner (evaluation.rs:641-650)."""
        *body, (horner, _) = self.calculations
        assert horner[0] == "Horner"
        rules = np.array([self._encode_calc(c) for c, _ in body], dtype=np.uint64).reshape(-1, 2)
        vp = np.array([self._encode_vs(p) for p in horner[2]], dtype=np.uint64)
        return rules, vp


def lookup_evaluator(input_expressions, table_expressions) -> GraphEvaluator:
    """Evaluator::new_inner, lookup part (evaluation.rs:417-446): theta-compressed input and
    table, then (input + beta) * (table + gamma)."""
    g = GraphEvaluator()

    def lc(exprs):
        parts = tuple(g.add_expression(e) for e in exprs)
        return g.add_calculation(("Horner", _vs("Constant", 0), parts, _vs("Theta")))

    ci = lc(input_expressions)
    ct = lc(table_expressions)
    right = g.add_calculation(("Add", ct, _vs("Gamma")))
    left = g.add_calculation(("Add", ci, _vs("Beta")))
    g.add_calculation(("Mul", left, right))
    return g


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


# ------------------------------------------------------ column pool / compress


class InsufficientGpuMemory(_lib.Halo2GpuError):
    """HaloGpuError::InsufficientGpuMemory (src/cuda/error.rs): the all-at-once budget check failed."""

    def __init__(self, context: str, magnitude: int, free_bytes: int):
        super().__init__(-2, f"{context}: {magnitude} columns do not fit in {free_bytes} free bytes")
        self.context = context
        self.magnitude = magnitude
        self.free_bytes = free_bytes


RESERVED_BYTES = 256 << 20


def query_device_free_bytes_for_chunking() -> int:
    """Free device memory minus a 256 MiB reservation (src/cuda/utils.rs:173-198)."""
    free, _ = torch.cuda.mem_get_info()
    return max(0, int(free) - RESERVED_BYTES)


class ColumnPool:
    """Per-prove device column set (src/cuda/funcs/column_pool.rs:21-283): fixed / advice /
    instance columns resident on the device with the pointer arrays the
    `_halo2_quotient_device_columns*` entry points read. Columns are either uploaded
    (`try_init`) or borrowed from caller-owned device tensors (`try_init_device`); the
    memory check is all-at-once, before any upload, and a failed check leaves the pool
    uninitialised so the caller can take its host arm."""

    def __init__(self, n: int):
        self.n = int(n)
        self._owned: list[torch.Tensor] = []
        self._ptrs = {"fixed": [], "advice": [], "instance": []}
        self._arrays = {}
        self.initialized = False

    def is_initialized(self) -> bool:
        return self.initialized

    @staticmethod
    def estimate_resident_bytes(n: int, n_fixed: int, n_advice: int, n_instance: int) -> int:
        return 32 * int(n) * (n_fixed + n_advice + n_instance)

    def is_gpu_memory_enough(self, n_fixed: int, n_advice: int, n_instance: int) -> bool:
        total = self.estimate_resident_bytes(self.n, n_fixed, n_advice, n_instance)
        return total < query_device_free_bytes_for_chunking()

    def _check(self, context: str, nf: int, na: int, ni: int) -> None:
        if not self.is_gpu_memory_enough(nf, na, ni):
            raise InsufficientGpuMemory(context, nf + na + ni, query_device_free_bytes_for_chunking())

    def _upload(self, kind: str, cols) -> None:
        for c in cols:
            c = np.ascontiguousarray(c, dtype=np.uint64)
            assert c.shape == (self.n, 4), f"{kind} column shape {c.shape} != ({self.n}, 4)"
            d = dv.to_device(c)
            self._owned.append(d)
            self._ptrs[kind].append(dv.ptr(d))

    def _borrow(self, kind: str, cols) -> None:
        for d in cols:
            assert d.numel() * 8 >= 32 * self.n, f"{kind} device column shorter than {self.n} rows"
            self._ptrs[kind].append(dv.ptr(d))

    def _finish(self) -> None:
        for kind, p in self._ptrs.items():
            self._arrays[kind] = (ctypes.c_void_p * max(1, len(p)))(*p)
        self.initialized = True

    def try_init(self, fixed_values, advice_values, instance_values) -> None:
        """Upload host columns ((n, 4) uint64 Montgomery words each). Idempotent."""
        if self.initialized:
            return
        self._check("ColumnPool::try_init", len(fixed_values), len(advice_values), len(instance_values))
        self._upload("fixed", fixed_values)
        self._upload("advice", advice_values)
        self._upload("instance", instance_values)
        self._finish()

    def try_init_device(self, pk_fixed_lagrange_mirror, fixed_values, advice_values_device,
                        instance_values_device) -> None:
        """Borrow device columns; fixed columns come from the proving key's device mirror
        when it is present (and matches the host column count), else are uploaded."""
        if self.initialized:
            return
        mirror = pk_fixed_lagrange_mirror
        borrowed = mirror is not None and len(mirror) > 0 and (len(fixed_values) == 0
                                                                or len(mirror) == len(fixed_values))
        nf = 0 if borrowed else len(fixed_values)
        self._check("ColumnPool::try_init_device", nf, 0, 0)
        if borrowed:
            self._borrow("fixed", mirror)
        else:
            self._upload("fixed", fixed_values)
        self._borrow("advice", advice_values_device)
        self._borrow("instance", instance_values_device)
        self._finish()

    def fixed_ptrs(self):
        return self._arrays["fixed"]

    def advice_ptrs(self):
        return self._arrays["advice"]

    def instance_ptrs(self):
        return self._arrays["instance"]

    def num_fixed(self) -> int:
        return len(self._ptrs["fixed"])

    def num_advice(self) -> int:
        return len(self._ptrs["advice"])

    def num_instance(self) -> int:
        return len(self._ptrs["instance"])


def for_compress(expressions) -> GraphEvaluator:
    """GraphEvaluator::for_compress (evaluation.rs:640-652): every expression, then
    Horner(Constant(0), parts, Y). The kernel starts the fold at zero and folds with
    expr_constants[4], where the caller places theta."""
    g = GraphEvaluator()
    parts = [g.add_expression(e) for e in expressions]
    g.add_calculation(("Horner", _vs("Constant", 0), tuple(parts), _vs("Y")))
    return g


def _pool_chunks(lib, rules, nvp: int, size: int, pool: ColumnPool, n_chal: int, n_const: int,
                 rot_scale: int):
    batch = int(lib._halo2_evaluate_h_max_rows(rules.ctypes.data, len(rules), nvp, size, pool.num_fixed(),
                                               pool.num_instance(), pool.num_advice(), n_chal, n_const, rot_scale,
                                               0, size, query_device_free_bytes_for_chunking()))
    if batch <= 0:
        raise _lib.Halo2GpuError(-2, "_halo2_evaluate_h_max_rows: no row fits in free device memory")
    return [(s, min(size, s + batch)) for s in range(0, size, batch)]


def _launch_pool(lib, graph: GraphEvaluator, expr_constants, size: int, rot_scale: int, pool: ColumnPool,
                 challenges, out_host: np.ndarray | None, out_dev: torch.Tensor | None, encoded=None) -> None:
    assert pool.is_initialized(), "ColumnPool is not initialised"
    if encoded is None:
        rules, vp = graph.encode_for_device()
        constants = graph.constants
    else:
        constants, rules, vp = encoded
    rules = np.ascontiguousarray(rules)
    nvp = len(vp)
    assert nvp > 0
    vp = np.ascontiguousarray(vp)
    consts = _fr_host(constants)
    chals = _fr_host(challenges)
    expr = _fr_host(expr_constants)
    heads = (pool.fixed_ptrs(), pool.num_fixed(), pool.instance_ptrs(), pool.num_instance(), pool.advice_ptrs(),
             pool.num_advice(), _lib.obj(chals.ctypes.data), len(challenges), _lib.obj(consts.ctypes.data),
             len(constants), _lib.obj(expr.ctypes.data), rules.ctypes.data, len(rules), vp.ctypes.data, nvp)
    for r0, r1 in _pool_chunks(lib, rules, nvp, size, pool, len(challenges), len(constants), rot_scale):
        ws = int(lib._halo2_quotient_workspace_size(rules.ctypes.data, len(rules), nvp, size, pool.num_fixed(),
                                                    pool.num_instance(), pool.num_advice(), len(challenges),
                                                    len(constants), rot_scale, r0, r1))
        scratch = dv.empty_bytes(ws)
        if out_host is not None:
            chunk = out_host[r0:r1]
            _lib.check(lib._halo2_quotient_device_columns(*heads, _lib.obj(chunk.ctypes.data), size, rot_scale, r0,
                                                          r1, dv.ptr(scratch), ws, dv.stream()),
                       "_halo2_quotient_device_columns")
        else:
            _lib.check(lib._halo2_quotient_device_columns_device_out(
                *heads, dv.ptr(out_dev) + 32 * r0, size, None, rot_scale, r0, r1, dv.ptr(scratch), ws,
                dv.stream()), "_halo2_quotient_device_columns_device_out")


def compress_expressions_device(expressions, theta: int, size: int, rot_scale: int, pool: ColumnPool,
                                challenges) -> np.ndarray:
    """evaluation.rs:1582-1715: sum over expressions folded with theta
    (acc = acc * theta + expr), evaluated on the device from the pool's columns, host output."""
    lib = _lib.load()
    graph = for_compress(expressions)
    out = np.zeros((size, 4), dtype=np.uint64)
    _launch_pool(lib, graph, [0, 1, R_MOD - 1, 2, theta], size, rot_scale, pool, challenges, out, None)
    return out


def compress_expressions_in_place_device(graph: GraphEvaluator, expr_constants, size: int, rot_scale: int,
                                         pool: ColumnPool, challenges, d_out: torch.Tensor) -> None:
    """evaluation.rs:1722-1830: a pre-built graph, caller-supplied expression constants
    (slot 4 = the fold factor), device output in `d_out` (size rows)."""
    lib = _lib.load()
    assert d_out.numel() * 8 >= 32 * size
    _launch_pool(lib, graph, expr_constants, size, rot_scale, pool, challenges, None, d_out)


def compress_expressions_with_runtime_constants_device(graph: GraphEvaluator, theta: int, beta: int, gamma: int,
                                                       y: int, size: int, rot_scale: int, pool: ColumnPool,
                                                       challenges, d_out: torch.Tensor) -> None:
    """evaluation.rs:1856-1990: a graph that reads Beta/Gamma/Theta/Y (the lookup evaluator),
    flattened with those values as trailing constants; expression slot 4 is 0, so each row's
    value is the terminal term itself. Device output in `d_out` (size rows)."""
    lib = _lib.load()
    assert d_out.numel() * 8 >= 32 * size
    enc = graph.encode_for_device_with_runtime_constants(theta, beta, gamma, y)
    _launch_pool(lib, graph, [0, 1, R_MOD - 1, 2, 0], size, rot_scale, pool, challenges, None, d_out, encoded=enc)
