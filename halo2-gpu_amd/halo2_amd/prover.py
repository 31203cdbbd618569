"""A proving pipeline over this library's device primitives, in the protocol order of
halo2's create_proof for KZG/SHPLONK (src/plonk/prover.rs:141-629 of the reference):
the verifying-key digest and the instance values absorbed, advice commitments, lookup
permutation (commit_permuted), permutation and lookup grand products, the vanishing
argument's random polynomial, h(X) over the extended domain (evaluate_h) and its
pieces, the evaluations at x and the SHPLONK multi-open.

The circuit is the standard-gate circuit of the k = 20 proving benchmark:
  advice  a, b, c = q_a a + q_b b + q_m a b, a lookup input d, a free column e;
  fixed   q_a, q_b, q_m, q_c = -1 and a range table (16-bit by default);
  instance one column holding NUM_PUBLIC public values (rows 0..NUM_PUBLIC-1);
  equality over a, b, e and the instance column (4 columns, chunk_len = degree - 2 = 2,
  so the permutation argument has two sets): random copy cycles of 2..8 cells over the
  usable rows of a, b, e, and every public value copied into 1..3 advice cells;
  the lookup d in table.
Rows from n - blinding - 1 on carry random blinders, as halo2 does.

Protocol choices that follow the reference's GPU prover rather than upstream halo2:
the vanishing argument's random polynomial is the constant 1 (its commitment is g[0]
and its evaluation 1, vanishing/prover.rs:68-95 and :236-240), and instances are not
opened (QUERY_INSTANCE = false for SHPLONK, shplonk/prover.rs:90): their values are
absorbed as scalars and the verifier evaluates them itself.

Transcript: Blake2b-512 with personalisation "Halo2-Transcript", prefix 0 before a
squeeze, 1 before a point (its affine x, y as 32-byte little-endian canonical
integers), 2 before a scalar; challenges are the 64-byte state digest reduced mod r.
Points are written to the proof as 32 bytes: canonical x with the parity of y in the
top bit. The verifying-key digest is a Blake2b-512 ("Halo2-Verify-Key") of this
library's own text form of the key (halo2 hashes Rust's Debug form of the pinned key,
which cannot be reproduced here), so proof bytes are not byte-identical to halo2's;
the tests hold a verifier of halo2's verify_proof + SHPLONK verifier that accepts
these proofs.
"""
from __future__ import annotations

import ctypes
import hashlib
import threading
import time
from dataclasses import dataclass, field, replace

import numpy as np
import torch

from . import _lib
from . import device as dv
from . import msm, ntt, polyops
from .domain import EvaluationDomain
from .evaluation import LookupCommitted, PermutationCommitted, evaluate_h_device
from .fields import DELTA, P_MOD, R_MOD, mont_limbs, value_of
from .lookup import lookup_product_device, permute_expression_pair_device
from .quotient import Advice, Fixed, Product, Sum, custom_gates_evaluator, lookup_evaluator
from .shplonk import ProverQuery, create_proof_device

CS_DEGREE = 4  # max(gate 3, permutation 3, lookup 4)
BLINDING_FACTORS = 5  # max(3, queries per advice column = 1) + 2
NUM_PUBLIC = 8
NUM_ADVICE, NUM_FIXED, NUM_INSTANCE = 5, 5, 1
A, B, C, D, E = range(5)  # advice columns
Q_A, Q_B, Q_M, Q_C, TABLE = range(5)  # fixed columns


# ------------------------------------------------------------- constraint system


def standard_gate():
    a, b, c = Advice(A), Advice(B), Advice(C)
    return Sum(Sum(Sum(Product(Fixed(Q_A), a), Product(Fixed(Q_B), b)), Product(Product(Fixed(Q_M), a), b)),
               Product(Fixed(Q_C), c))


@dataclass(frozen=True)
class ConstraintSystem:
    """The circuit's shape as halo2's ConstraintSystem records it (plonk/circuit.rs):
    the queries in the order configure() creates them (the equality columns first, then
    the gate's, then the lookup's), the permutation columns in enable_equality order,
    the gate polynomials and the lookups (input expressions, table expressions)."""

    advice_queries: tuple = ((A, 0), (B, 0), (E, 0), (C, 0), (D, 0))
    fixed_queries: tuple = ((Q_A, 0), (Q_B, 0), (Q_M, 0), (Q_C, 0), (TABLE, 0))
    instance_queries: tuple = ((0, 0),)
    permutation_columns: tuple = (("advice", A), ("advice", B), ("advice", E), ("instance", 0))
    gates: tuple = (standard_gate(),)
    lookups: tuple = (((Advice(D),), (Fixed(TABLE),)),)
    degree: int = CS_DEGREE
    blinding_factors: int = BLINDING_FACTORS
    num_fixed: int = NUM_FIXED
    num_advice: int = NUM_ADVICE
    num_instance: int = NUM_INSTANCE

    @property
    def chunk_len(self) -> int:
        return self.degree - 2

    @property
    def num_permutation_sets(self) -> int:
        return -(-len(self.permutation_columns) // self.chunk_len)

    @property
    def quotient_pieces(self) -> int:
        return self.degree - 1


STANDARD_CS = ConstraintSystem()


# ------------------------------------------------------------------ transcript


def affine_of(out96) -> tuple | None:
    """The MSM's 96-byte projective-out point (x, y, 1) / (0, 1, 0), Montgomery limbs,
    as canonical affine integers (None for the identity)."""
    w = np.ascontiguousarray(np.asarray(out96, dtype=np.uint64).reshape(12))
    if not w[8:12].any():
        return None
    return value_of(w[0:4], P_MOD), value_of(w[4:8], P_MOD)


def compress(pt: tuple) -> bytes:
    """32-byte point encoding: canonical x little-endian, parity of y in the top bit."""
    x, y = pt
    return (x | ((y & 1) << 255)).to_bytes(32, "little")


class Transcript:
    """Blake2b transcript (see the module doc): common_* absorb, write_* absorb and
    append to the proof, squeeze returns a challenge."""

    def __init__(self):
        self.h = hashlib.blake2b(digest_size=64, person=b"Halo2-Transcript")
        self.proof = bytearray()

    def common_point(self, pt: tuple | None) -> None:
        if pt is None:
            raise ValueError("cannot write the point at infinity to the transcript")
        self.h.update(b"\x01" + pt[0].to_bytes(32, "little") + pt[1].to_bytes(32, "little"))

    def write_point(self, pt: tuple | None) -> None:
        self.common_point(pt)
        self.proof += compress(pt)

    def common_scalar(self, v: int) -> None:
        self.h.update(b"\x02" + (v % R_MOD).to_bytes(32, "little"))

    def write_scalar(self, v: int) -> None:
        self.common_scalar(v)
        self.proof += (v % R_MOD).to_bytes(32, "little")

    def squeeze(self) -> int:
        self.h.update(b"\x00")
        return int.from_bytes(self.h.copy().digest(), "little") % R_MOD


# ---------------------------------------------------------------- the circuit


def _mont_column(values: np.ndarray) -> torch.Tensor:
    """Small canonical integers (< 2^64) -> a device column in Montgomery form: the raw
    word times R^2 through one Montgomery product per element."""
    n = len(values)
    raw = np.zeros((n, 4), dtype=np.uint64)
    raw[:, 0] = values.astype(np.uint64)
    d_raw = dv.to_device(raw)
    d_r2 = dv.to_device(np.tile(mont_limbs(pow(2, 256, R_MOD)), (n, 1)))
    polyops.poly_elementwise_multiply_device(d_raw, d_raw, d_r2)
    return d_raw


def _set_rows(col: torch.Tensor, start: int, values: torch.Tensor) -> None:
    col.view(-1, 4)[start:].copy_(values.view(-1, 4)[: col.numel() // 4 - start])


@dataclass
class Circuit:
    """One circuit of the family: fixed columns and copy structure (the proving key's
    part) plus the current witness (advice, instance and the public values)."""

    k: int
    fixed: list  # device Lagrange columns: q_a, q_b, q_m, q_c, table
    advice: list  # device Lagrange columns a, b, c, d, e
    instance: list  # one device Lagrange column
    instance_values: list  # the public values (canonical ints), rows 0..NUM_PUBLIC-1
    cycles: list  # copy cycles, each a list of (permutation column, row); first cell = source
    usable: int
    table_size: int
    copy_src: torch.Tensor | None = None  # flat cell indices col * n + row (device)
    copy_dst: torch.Tensor | None = None
    cs: ConstraintSystem = STANDARD_CS


def synthesize_standard(k: int, seed: int, table_bits: int = 16) -> Circuit:
    """The standard-gate circuit with a satisfying random witness (see the module doc)."""
    n = 1 << k
    u = n - BLINDING_FACTORS - 1
    assert u > NUM_PUBLIC
    rng = np.random.default_rng(seed)
    q_a, q_b, q_m = (msm.random_fr(n, seed * 16 + i) for i in range(3))
    q_c = dv.to_device(np.tile(mont_limbs(R_MOD - 1), (n, 1)))
    zeros = torch.zeros(4 * n, dtype=torch.int64, device=dv.device())
    for q in (q_a, q_b, q_m, q_c):
        _set_rows(q, u, zeros)  # fixed columns are zero outside the usable rows
    table_size = min(1 << table_bits, u)
    tv = np.zeros(n, dtype=np.uint64)
    tv[:table_size] = np.arange(table_size, dtype=np.uint64)
    table = _mont_column(tv)
    # copy cycles over the permutation columns (0 = a, 1 = b, 2 = e, 3 = instance):
    # each public value starts a cycle into 1..3 advice cells, then random cycles of
    # 2..8 advice cells
    cells = rng.permutation(3 * u)[: max(2 * NUM_PUBLIC + 2, (3 * u) // 8)]
    cycles, i = [], 0
    for j in range(NUM_PUBLIC):
        ln = int(rng.integers(1, 4))
        cycles.append([(3, j)] + [(int(c) // u, int(c) % u) for c in cells[i: i + ln]])
        i += ln
    while i + 2 <= len(cells):
        ln = int(rng.integers(2, 9))
        cyc = [(int(c) // u, int(c) % u) for c in cells[i: i + ln]]
        if len(cyc) >= 2:
            cycles.append(cyc)
        i += ln
    src = np.fromiter((cyc[0][0] * n + cyc[0][1] for cyc in cycles for _ in cyc[1:]), dtype=np.int64)
    dst = np.fromiter((c * n + r for cyc in cycles for c, r in cyc[1:]), dtype=np.int64)
    circuit = Circuit(k, [q_a, q_b, q_m, q_c, table], [], [], [], cycles, u, table_size,
                      torch.from_numpy(src).to(dv.device()), torch.from_numpy(dst).to(dv.device()))
    return fresh_witness(circuit, seed)


def fresh_witness(circuit: Circuit, seed: int) -> Circuit:
    """The circuit with a new satisfying witness (fixed columns, copy cycles and table
    unchanged, so the proving key stays valid), generated on the device: a, b, e and the
    public values random, every copied cell set to its cycle's first cell,
    c = q_a a + q_b b + q_m a b, d drawn from the table, random blinders on the last
    rows of every advice column."""
    k, u = circuit.k, circuit.usable
    n = 1 << k
    q_a, q_b, q_m, _, table = circuit.fixed
    inst = torch.zeros(4 * n, dtype=torch.int64, device=dv.device())
    _set_rows(inst[: 4 * NUM_PUBLIC], 0, msm.random_fr(NUM_PUBLIC, seed * 16 + 9))
    flat = torch.cat([msm.random_fr(n, seed * 16 + 5), msm.random_fr(n, seed * 16 + 6),
                      msm.random_fr(n, seed * 16 + 7), inst]).view(-1, 4)
    flat[circuit.copy_dst] = flat[circuit.copy_src]  # sources are first cells, never destinations
    a, b, e, inst = (flat[j * n:(j + 1) * n].reshape(-1) for j in range(4))
    a, b, e, inst = a.clone(), b.clone(), e.clone(), inst.clone()
    c = dv.empty_u64(4 * n)
    t = dv.empty_u64(4 * n)
    polyops.poly_elementwise_multiply_device(c, q_a, a)
    polyops.poly_elementwise_multiply_device(t, q_b, b)
    polyops.poly_multiply_add_device(c, t, 1)
    polyops.poly_elementwise_multiply_device(t, q_m, a)
    polyops.poly_elementwise_multiply_device(t, t, b)
    polyops.poly_multiply_add_device(c, t, 1)  # zero selectors leave c free on the blinding rows
    gen = torch.Generator(device=dv.device())
    gen.manual_seed(seed * 7919 + 13)
    idx = torch.randint(0, circuit.table_size, (n,), generator=gen, device=dv.device())
    d = table.view(-1, 4)[idx].reshape(-1).clone()
    for j, col in enumerate((a, b, c, d, e)):  # blinders
        _set_rows(col, u, msm.random_fr(n - u, seed * 16 + 10 + j))
    public = [value_of(r) for r in dv.to_host(inst[: 4 * NUM_PUBLIC])]
    return replace(circuit, advice=[a, b, c, d, e], instance=[inst], instance_values=public)


# ---------------------------------------------------------------- proving key


@dataclass
class VerifyingKey:
    """keygen_vk: the domain size, the constraint system, the commitments of the fixed
    columns and of the permutation's sigma columns (affine integer pairs) and the
    digest absorbed first into every transcript."""

    k: int
    cs: ConstraintSystem
    fixed_commitments: list
    permutation_commitments: list
    transcript_repr: int = 0

    def pinned(self) -> str:
        """The text form the digest hashes: every part of the key that fixes the
        statement (domain, queries, gates, lookups, permutation, commitments)."""
        cs = self.cs
        hexpt = lambda p: "identity" if p is None else f"({p[0]:064x},{p[1]:064x})"  # noqa: E731
        return (f"k={self.k};degree={cs.degree};blinding={cs.blinding_factors};"
                f"fixed={cs.num_fixed};advice={cs.num_advice};instance={cs.num_instance};"
                f"advice_queries={list(cs.advice_queries)};fixed_queries={list(cs.fixed_queries)};"
                f"instance_queries={list(cs.instance_queries)};permutation={list(cs.permutation_columns)};"
                f"gates={list(cs.gates)};lookups={list(cs.lookups)};"
                f"fixed_commitments=[{','.join(hexpt(p) for p in self.fixed_commitments)}];"
                f"permutation_commitments=[{','.join(hexpt(p) for p in self.permutation_commitments)}]")

    def compute_transcript_repr(self) -> int:
        s = self.pinned().encode()
        h = hashlib.blake2b(digest_size=64, person=b"Halo2-Verify-Key")
        h.update(len(s).to_bytes(8, "little"))
        h.update(s)
        return int.from_bytes(h.digest(), "little") % R_MOD


@dataclass
class ProvingKey:
    """keygen_pk for the circuit's fixed part (src/plonk/keygen.rs in halo2-axiom): the
    domain, the fixed columns (coefficients, and their Lagrange values in
    `fixed_values`), l0 / l_last / l_active and the permutation's sigma columns
    (Lagrange and coefficients), all resident on the device, plus the verifying key.
    pk_io writes and reads it as bytes."""

    domain: EvaluationDomain
    fixed_polys: list
    l0: torch.Tensor
    l_last: torch.Tensor
    l_active: torch.Tensor
    sigma_values: list  # Lagrange
    sigma_polys: list  # coefficients
    vk: VerifyingKey | None = None
    gates: object = None
    lookup: object = None
    key_parts: dict = field(default_factory=dict)  # coset values of the fixed part, per extended part
    timings: dict = field(default_factory=dict)
    fixed_values: list = field(default_factory=list)  # Lagrange


def keygen(circuit: Circuit, params: msm.ParamsKZG | None = None) -> ProvingKey:
    """The proving key; with `params`, also the verifying key (commitments of the fixed
    and sigma columns against params' Lagrange basis) that create_proof needs."""
    k, n, u, cs = circuit.k, 1 << circuit.k, circuit.usable, circuit.cs
    dom = EvaluationDomain(cs.degree, k)
    fixed_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.fixed)
    lv = np.zeros((3, n), dtype=np.uint64)
    lv[0, 0] = 1
    lv[1, u] = 1
    lv[2, :u] = 1
    l0, l_last, l_active = dom.lagrange_to_coeff_many_device_inputs([_mont_column(r) for r in lv])
    # sigma: cell (col, row) -> the next cell of its cycle; every other cell maps to itself
    m = len(cs.permutation_columns)
    mapping = np.zeros((m, n, 2), dtype=np.uint32)
    mapping[:, :, 0] = np.arange(m, dtype=np.uint32)[:, None]
    mapping[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
    for cyc in circuit.cycles:
        for i, (c, r) in enumerate(cyc):
            mapping[c, r] = cyc[(i + 1) % len(cyc)]
    flat = mapping.reshape(-1, 2)
    sigma_values = [polyops.generate_omegadelta_device(flat, dom.omega, DELTA, k, j * n, (j + 1) * n - 1, m)
                    for j in range(m)]
    sigma_polys = dom.lagrange_to_coeff_many_device_inputs(sigma_values)
    pk = ProvingKey(dom, fixed_polys, l0, l_last, l_active, sigma_values, sigma_polys,
                    fixed_values=list(circuit.fixed))
    pk.gates = custom_gates_evaluator(list(cs.gates))
    (inputs, tables), = cs.lookups
    pk.lookup = lookup_evaluator(list(inputs), list(tables))
    if params is not None:
        pk.vk = keygen_vk(params, circuit, pk)
    return pk


def keygen_vk(params: msm.ParamsKZG, circuit: Circuit, pk: ProvingKey) -> VerifyingKey:
    pts = params.commit_lagrange_many_device(list(circuit.fixed) + list(pk.sigma_values))
    nf = len(circuit.fixed)
    vk = VerifyingKey(circuit.k, circuit.cs, [affine_of(p) for p in pts[:nf]], [affine_of(p) for p in pts[nf:]])
    vk.transcript_repr = vk.compute_transcript_repr()
    return vk


# ---------------------------------------------------------------------- prove


def _scalar_dev(v: int) -> torch.Tensor:
    return dv.to_device(mont_limbs(v % R_MOD).reshape(1, 4))


def _scan_into_z(ratio: torch.Tensor, d_prefix: torch.Tensor, n: int, seed: int) -> torch.Tensor:
    """z[0] = prefix, z[i] = prefix * prod_{j < i} ratio[j] for i < n - blinding, the last
    BLINDING_FACTORS rows random (permutation/prover.rs:179-240, lookup/prover.rs:
    294-461). `d_prefix` is a device element (the previous set's z on its last usable
    row, or one)."""
    lib = _lib.load()
    z = dv.empty_u64(4 * n)
    zv, rv = z.view(-1, 4), ratio.view(-1, 4)
    zv[0].copy_(torch.from_numpy(mont_limbs(1).view(np.int64)).to(z.device, non_blocking=True))
    zv[1:].copy_(rv[: n - 1])
    _lib.check(lib._halo2_grand_product_device_inputs(ctypes.byref(_lib.obj(dv.ptr(z))),
                                                      ctypes.byref(_lib.obj(dv.ptr(d_prefix))), n, 0, dv.stream()),
               "_halo2_grand_product_device_inputs")
    _set_rows(z, n - BLINDING_FACTORS, msm.random_fr(BLINDING_FACTORS, seed))
    return z


_side = threading.local()


def _side_stream() -> torch.cuda.Stream:
    """A second stream per host thread for transforms that overlap a commitment MSM
    (its sort, bucket and reduction kernels leave the device partly idle)."""
    if getattr(_side, "stream", None) is None:
        _side.stream = torch.cuda.Stream()
    return _side.stream


def _closes(z: torch.Tensor, u: int) -> bool:
    return bool((z.view(-1, 4)[u].cpu().numpy().view(np.uint64) == mont_limbs(1)).all())


def rotate(dom: EvaluationDomain, x: int, rot: int) -> int:
    """domain.rotate_omega: x * omega^rot."""
    w = dom.omega if rot >= 0 else dom.omega_inv
    return x * pow(w, abs(rot), R_MOD) % R_MOD


def create_proof(params: msm.ParamsKZG, pk: ProvingKey, circuit: Circuit, seed: int = 1, phase_hook=None) -> dict:
    """One proof; returns the proof bytes, the public values, the per-phase times (ms,
    synchronised on this thread's stream) and the self-checks. `phase_hook(name)`, when
    given, runs after each phase has finished on the device (a serving loop uses it to
    let smaller proofs run first)."""
    lib = _lib.load()
    cs = circuit.cs
    dom, n, u = pk.domain, 1 << circuit.k, circuit.usable
    assert pk.vk is not None, "create_proof needs the verifying key: keygen(circuit, params)"
    tr = Transcript()
    times, checks = {}, {}
    t0 = time.perf_counter()

    def phase(name):
        nonlocal t0
        torch.cuda.current_stream().synchronize()  # this proof's stream only: proofs may run side by side
        t = time.perf_counter()
        times[name] = round((t - t0) * 1e3, 3)
        if phase_hook is not None:
            phase_hook(name)
        t0 = time.perf_counter()

    # 0. the key's digest and the instance values (prover.rs:165; witness.rs:237-245)
    tr.common_scalar(pk.vk.transcript_repr)
    for v in circuit.instance_values:
        tr.common_scalar(v)

    main, side = torch.cuda.current_stream(), _side_stream()
    # 1. advice (witness.rs:300-345): commitments in Lagrange form; the coefficient forms
    #    are computed on the side stream meanwhile
    side.wait_stream(main)
    with torch.cuda.stream(side):
        advice_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.advice)
        instance_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.instance)
    for pt in params.commit_lagrange_many_device(circuit.advice):
        tr.write_point(affine_of(pt))
    main.wait_stream(side)
    phase("advice")

    # 2. lookup commit_permuted (lookup/prover.rs:552-627): theta-compression of one
    #    column is the column itself; permuted pair, blinding rows, commitments
    theta = tr.squeeze()
    inp, tab = circuit.advice[D], (pk.fixed_values or circuit.fixed)[TABLE]
    p_in, p_tab = dv.empty_u64(4 * n), dv.empty_u64(4 * n)
    permute_expression_pair_device(inp, tab, p_in, p_tab, n, u)
    _set_rows(p_in, u, msm.random_fr(n - u, seed * 7 + 1))
    _set_rows(p_tab, u, msm.random_fr(n - u, seed * 7 + 2))
    for pt in params.commit_lagrange_many_device([p_in, p_tab]):
        tr.write_point(affine_of(pt))
    phase("lookup_permuted")

    # 3. grand products (permutation/prover.rs:69-273, lookup/prover.rs:294-461): one
    #    permutation product per set of chunk_len columns, each starting from the previous
    #    set's value on the last usable row, delta^j running across the sets
    beta, gamma = tr.squeeze(), tr.squeeze()
    d_beta, d_gamma, d_delta, d_omega = (_scalar_dev(v) for v in (beta, gamma, DELTA, dom.omega))
    ws = int(lib._halo2_permutation_product_device_inputs_workspace_size(n))
    groups = {"advice": circuit.advice, "instance": circuit.instance}
    perm_values = [groups[kind][idx] for kind, idx in cs.permutation_columns]

    def perm_ratio(c0: int) -> torch.Tensor:
        cols = perm_values[c0:c0 + cs.chunk_len]
        sig = pk.sigma_values[c0:c0 + cs.chunk_len]
        ratio = dv.empty_u64(4 * n)
        polyops.poly_fill_one_device(ratio)  # the permutation product multiplies into it
        scratch = dv.empty_bytes(ws)
        # delta^c0: the set's first column power, advanced by delta once per column
        d_deltaomega = _scalar_dev(pow(DELTA, c0, R_MOD))
        _lib.check(lib._halo2_permutation_product_device_inputs(
            ctypes.byref(_lib.obj(dv.ptr(ratio))), _lib.obj_array([dv.ptr(x) for x in sig]),
            _lib.obj_array([dv.ptr(x) for x in cols]), dv.ptr(d_beta), dv.ptr(d_gamma), dv.ptr(d_delta),
            dv.ptr(d_omega), dv.ptr(d_deltaomega), n, 0, len(cols), dv.ptr(scratch), ws, dv.stream()),
            "_halo2_permutation_product_device_inputs")
        return ratio

    # the sets' ratios (products, one batch inversion each: latency-bound) and the lookup
    # product are independent; only the permutation scans chain, through the prefix. The
    # first set runs here, the other sets on one side stream, the lookup on another.
    starts = list(range(0, len(perm_values), cs.chunk_len))
    side_a, side_b = dv.side_streams(2)
    side_a.wait_stream(main)
    side_b.wait_stream(main)
    with torch.cuda.stream(side_a):
        later = [perm_ratio(c0) for c0 in starts[1:]]
    with torch.cuda.stream(side_b):
        lratio = lookup_product_device(p_in, p_tab, inp, tab, beta, gamma)
        z_lookup = _scan_into_z(lratio, _scalar_dev(1), n, seed * 7 + 4)
    z_perm, prefix = [], _scalar_dev(1)
    for s, c0 in enumerate(starts):
        if s == 1:
            main.wait_stream(side_a)
        ratio = perm_ratio(c0) if s == 0 else later[s - 1]
        z = _scan_into_z(ratio, prefix, n, seed * 7 + 10 + s)
        z_perm.append(z)
        prefix = z.view(-1, 4)[u]
    for t in later:  # made on side_a, scanned here
        t.record_stream(main)
    main.wait_stream(side_b)
    z_lookup.record_stream(main)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        polys = dom.lagrange_to_coeff_many_device_inputs(z_perm + [z_lookup, p_in, p_tab])
    z_perm_polys, (z_lookup_poly, p_in_poly, p_tab_poly) = polys[:len(z_perm)], polys[len(z_perm):]
    pts = params.commit_lagrange_many_device(z_perm + [z_lookup])
    # read back after the commitments (which wait for the stream anyway), not before them
    checks["permutation_product_closes"] = _closes(z_perm[-1], u)
    checks["lookup_product_closes"] = _closes(z_lookup, u)
    for pt in pts:
        tr.write_point(affine_of(pt))
    main.wait_stream(side)
    phase("grand_products")

    # 4. the vanishing argument's random polynomial: the constant 1, committed as g[0]
    #    (vanishing/prover.rs:68-95)
    random_poly = _scalar_dev(1).reshape(-1)
    tr.write_point(affine_of(params.g0_out12()))
    y = tr.squeeze()

    # 5. h(X) (evaluation.rs:1056-1417), divided by X^n - 1, in pieces of n
    num = evaluate_h_device(
        dom, pk.gates, pk.fixed_polys, pk.l0, pk.l_last, pk.l_active, advice_polys, instance_polys, [], y, beta,
        gamma, theta, lookups=[LookupCommitted(pk.lookup, z_lookup_poly, p_in, p_tab, p_in_poly, p_tab_poly)],
        permutation=PermutationCommitted(z_perm_polys, list(cs.permutation_columns), pk.sigma_polys),
        blinding_factors=cs.blinding_factors, cs_degree=cs.degree, key_parts=pk.key_parts)
    dom.divide_by_vanishing_poly_device(num)
    ntt.fft_normal_device(ntt.IFFT, dom.extended_k, num, num, dom.extended_omega_inv, dom.extended_ifft_divisor)
    ntt.distribute_powers_zeta_device(num, [dom.g_coset_inv, dom.g_coset])
    top = num.view(-1, 4)[cs.quotient_pieces * n:]
    nonzero_top = top.any()  # read back after the commitment below
    pieces = [num[4 * i * n: 4 * (i + 1) * n] for i in range(cs.quotient_pieces)]
    for pt in params.commit_many_device(pieces):
        tr.write_point(affine_of(pt))
    checks["h_is_a_polynomial"] = not bool(nonzero_top.item())
    phase("quotient")

    # 6. x, h = sum x^(n i) h_i (vanishing/prover.rs:181-230), the queries in halo2's open
    #    order (prover.rs:574-616): advice, permutation products (x, omega x per set, then
    #    omega^last x for every set but the last), lookups (product x, permuted input x,
    #    permuted table x, permuted input omega^-1 x, product omega x), fixed, sigma, h,
    #    the random polynomial. All are evaluated in one batch; the transcript gets them
    #    in its own order (prover.rs:398-520): advice, fixed, random, sigma, products,
    #    lookups; h's value is not written (the verifier derives it).
    x = tr.squeeze()
    xn = pow(x, n, R_MOD)
    h_poly = pieces[0].clone()
    for i in range(1, len(pieces)):
        polyops.poly_multiply_add_device(h_poly, pieces[i], pow(xn, i, R_MOD))
    x_next, x_prev = rotate(dom, x, 1), rotate(dom, x, -1)
    x_last = rotate(dom, x, -(cs.blinding_factors + 1))
    adv = [(advice_polys[c], rotate(dom, x, r)) for c, r in cs.advice_queries]
    perm = [q for z in z_perm_polys for q in ((z, x), (z, x_next))] + [(z, x_last) for z in z_perm_polys[:-1]][::-1]
    lkp = [(z_lookup_poly, x), (p_in_poly, x), (p_tab_poly, x), (p_in_poly, x_prev), (z_lookup_poly, x_next)]
    fix = [(pk.fixed_polys[c], rotate(dom, x, r)) for c, r in cs.fixed_queries]
    sig = [(p, x) for p in pk.sigma_polys]
    van = [(h_poly, x), (random_poly, x)]
    queries = adv + perm + lkp + fix + sig + van
    values = [value_of(e) for e in polyops.batch_eval_polynomial_d2h([p for p, _ in queries],
                                                                      [pt for _, pt in queries])]
    ev = {(id(p), pt): v for (p, pt), v in zip(queries, values)}
    written = adv + fix + [(random_poly, x)] + sig
    for z in z_perm_polys[:-1]:
        written += [(z, x), (z, x_next), (z, x_last)]
    written += [(z_perm_polys[-1], x), (z_perm_polys[-1], x_next)]
    written += [(z_lookup_poly, x), (z_lookup_poly, x_next), (p_in_poly, x), (p_in_poly, x_prev), (p_tab_poly, x)]
    for p, pt in written:
        tr.write_scalar(ev[(id(p), pt)])
    phase("evaluations")

    # 7. SHPLONK multi-open (shplonk/prover.rs:97-441): y and v squeezed, h written,
    #    u squeezed from the transcript that holds h, h_final written
    sy, sv = tr.squeeze(), tr.squeeze()
    ids: dict = {}
    pq = [ProverQuery(pt, p, ids.setdefault(id(p), len(ids))) for p, pt in queries]

    def u_after_h(h):
        tr.write_point(affine_of(h))
        return tr.squeeze()

    out = create_proof_device(params, pq, sy, sv, u_after_h, evals=values)
    tr.write_point(affine_of(out["h_final"]))
    phase("opening")
    return {"proof": bytes(tr.proof), "instance": list(circuit.instance_values), "times_ms": times,
            "checks": checks}
