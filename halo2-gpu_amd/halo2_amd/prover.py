"""A proving pipeline over this library's device primitives, in the phase order of
halo2's create_proof (src/plonk/prover.rs:141-629 of the reference): advice
commitments, lookup permutation (commit_permuted), permutation and lookup grand
products, the vanishing argument's random polynomial, h(X) over the extended
domain (evaluate_h), its pieces, the evaluations at x and the SHPLONK opening.

The circuit is the standard-gate circuit of the k = 20 proving benchmark:
5 advice columns (a, b, c = q_a a + q_b b + q_m a b, a lookup input d, a free e),
fixed selectors q_a, q_b, q_m, q_c = -1 and a 16-bit table, one instance column,
copy constraints in random cycles of 2..8 cells over a and b, and the lookup
d in table. Rows from n - blinding - 1 on carry random blinders, as halo2 does.

Not the reference's byte format: the transcript is a Blake2b absorb/squeeze in
halo2's style (prefixes 0 = challenge, 1 = point, 2 = scalar), not pinned to the
reference's bytes, and no verifier runs. What the pipeline proves about itself:
the numerator of h is divisible by X^n - 1 (the vanishing division leaves every
coefficient from (cs_degree - 1) n up zero), which holds only if the witness
satisfies every gate, copy constraint and lookup and the grand products close.
"""
from __future__ import annotations

import ctypes
import hashlib
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import device as dv
from . import msm, ntt, polyops
from .domain import EvaluationDomain
from .evaluation import LookupCommitted, PermutationCommitted, evaluate_h_device
from .fields import DELTA, P_MOD, R_MOD, mont_limbs
from .lookup import lookup_product_device, permute_expression_pair_device
from .quotient import Advice, Fixed, Product, Sum, custom_gates_evaluator, lookup_evaluator
from .shplonk import ProverQuery, create_proof_device

CS_DEGREE = 4  # max(gate 3, permutation 3, lookup 4)
BLINDING_FACTORS = 5


# ------------------------------------------------------------------ transcript


class Transcript:
    """Blake2b transcript in halo2's style (transcript/blake2b: personal "Halo2-Transcript",
    64-byte state, prefix 1 before points, 2 before scalars, 0 before each squeeze; a
    challenge is the 64-byte digest reduced mod r)."""

    def __init__(self):
        self.h = hashlib.blake2b(digest_size=64, person=b"Halo2-Transcript")
        self.proof = bytearray()

    def write_point(self, out96: np.ndarray) -> None:
        x = int.from_bytes(np.ascontiguousarray(out96[:4]).tobytes(), "little") * pow(1 << 256, -1, P_MOD) % P_MOD
        y = int.from_bytes(np.ascontiguousarray(out96[4:8]).tobytes(), "little") * pow(1 << 256, -1, P_MOD) % P_MOD
        if not np.asarray(out96[8:12]).any():
            x = y = 0  # the identity
        data = x.to_bytes(32, "little") + y.to_bytes(32, "little")
        self.h.update(b"\x01" + data)
        self.proof += data

    def write_scalar(self, v: int) -> None:
        data = (v % R_MOD).to_bytes(32, "little")
        self.h.update(b"\x02" + data)
        self.proof += data

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
    k: int
    fixed: list  # device Lagrange columns: q_a, q_b, q_m, q_c, table
    advice: list  # a, b, c, d, e
    instance: list
    cycles: list  # copy cycles, each a list of (column 0 = a / 1 = b, row)
    usable: int


def synthesize_standard(k: int, seed: int, table_bits: int = 16) -> Circuit:
    """The standard-gate circuit with a satisfying random witness (see the module doc)."""
    n = 1 << k
    u = n - BLINDING_FACTORS - 1
    rng = np.random.default_rng(seed)
    zero_tail = lambda col: _set_rows(col, u, torch.zeros(4 * n, dtype=torch.int64, device=dv.device()))  # noqa
    q_a, q_b, q_m = (msm.random_fr(n, seed * 16 + i) for i in range(3))
    q_c = dv.to_device(np.tile(mont_limbs(R_MOD - 1), (n, 1)))
    for q in (q_a, q_b, q_m, q_c):
        zero_tail(q)  # fixed columns are zero outside the usable rows
    table_size = min(1 << table_bits, u)
    tv = np.zeros(n, dtype=np.uint64)
    tv[:table_size] = np.arange(table_size, dtype=np.uint64)
    table = _mont_column(tv)
    # random copy cycles of 2..8 cells over a and b (column 0 = a, 1 = b)
    cells = rng.permutation(2 * u)[: max(2, (2 * u) // 8)]
    cycles, i = [], 0
    while i + 2 <= len(cells):
        ln = int(rng.integers(2, 9))
        cyc = [(int(c) // u, int(c) % u) for c in cells[i: i + ln]]
        if len(cyc) >= 2:
            cycles.append(cyc)
        i += ln
    instance = [torch.zeros(4 * n, dtype=torch.int64, device=dv.device())]
    circuit = Circuit(k, [q_a, q_b, q_m, q_c, table], [], instance, cycles, u)
    circuit.advice = fresh_witness(circuit, seed, table_size)
    return circuit


def fresh_witness(circuit: Circuit, seed: int, table_size: int | None = None) -> list:
    """A new satisfying assignment of the advice columns for the same circuit (fixed
    columns, copy cycles and table unchanged, so the proving key stays valid): a, b
    random with every cell of a cycle equal to its first cell, c = q_a a + q_b b + q_m a b,
    d drawn from the table, e free, random blinders on the last rows."""
    k, u = circuit.k, circuit.usable
    n = 1 << k
    if table_size is None:
        table_size = min(1 << 16, u)
    rng = np.random.default_rng(seed + 0x5EED)
    q_a, q_b, q_m = circuit.fixed[:3]
    a = msm.random_fr(n, seed * 16 + 5)
    b = msm.random_fr(n, seed * 16 + 6)
    flat = np.concatenate([dv.to_host(a), dv.to_host(b)])  # cell (col, row) at col * n + row
    src = np.fromiter((cyc[0][0] * n + cyc[0][1] for cyc in circuit.cycles for _ in cyc[1:]), dtype=np.int64)
    dst = np.fromiter((c * n + r for cyc in circuit.cycles for c, r in cyc[1:]), dtype=np.int64)
    flat[dst] = flat[src]  # sources are first cells, never destinations
    a, b = dv.to_device(flat[:n]), dv.to_device(flat[n:])
    # c = q_a a + q_b b + q_m a b (q_c = -1); zero selectors leave c free on the blinding rows
    c = dv.empty_u64(4 * n)
    t = dv.empty_u64(4 * n)
    polyops.poly_elementwise_multiply_device(c, q_a, a)
    polyops.poly_elementwise_multiply_device(t, q_b, b)
    polyops.poly_multiply_add_device(c, t, 1)
    polyops.poly_elementwise_multiply_device(t, q_m, a)
    polyops.poly_elementwise_multiply_device(t, t, b)
    polyops.poly_multiply_add_device(c, t, 1)
    d = _mont_column(rng.integers(0, table_size, n, dtype=np.uint64))
    e = msm.random_fr(n, seed * 16 + 7)
    for j, col in enumerate((a, b, c, d)):  # blinders
        _set_rows(col, u, msm.random_fr(n - u, seed * 16 + 8 + j))
    return [a, b, c, d, e]


# ---------------------------------------------------------------- proving key


def standard_gate():
    a, b, c = Advice(0), Advice(1), Advice(2)
    return Sum(Sum(Sum(Product(Fixed(0), a), Product(Fixed(1), b)), Product(Product(Fixed(2), a), b)),
               Product(Fixed(3), c))


@dataclass
class ProvingKey:
    """keygen_pk for the circuit's fixed part (src/plonk/keygen.rs in halo2-axiom): the
    domain, fixed polynomials, l0 / l_last / l_active and the permutation's sigma
    polynomials, all resident on the device."""

    domain: EvaluationDomain
    fixed_polys: list
    l0: torch.Tensor
    l_last: torch.Tensor
    l_active: torch.Tensor
    sigma_values: list  # Lagrange
    sigma_polys: list  # coefficients
    gates: object = None
    lookup: object = None
    key_parts: dict = field(default_factory=dict)  # coset values of the fixed part, per extended part
    timings: dict = field(default_factory=dict)


def keygen(circuit: Circuit) -> ProvingKey:
    k, n, u = circuit.k, 1 << circuit.k, circuit.usable
    dom = EvaluationDomain(CS_DEGREE, k)
    fixed_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.fixed)
    lv = np.zeros((3, n), dtype=np.uint64)
    lv[0, 0] = 1
    lv[1, u] = 1
    lv[2, :u] = 1
    l0, l_last, l_active = dom.lagrange_to_coeff_many_device_inputs([_mont_column(r) for r in lv])
    # sigma: cell (col, row) -> the next cell of its cycle; every other cell maps to itself
    mapping = np.zeros((2, n, 2), dtype=np.uint32)
    mapping[:, :, 0] = np.arange(2, dtype=np.uint32)[:, None]
    mapping[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
    for cyc in circuit.cycles:
        for i, (c, r) in enumerate(cyc):
            nc, nr = cyc[(i + 1) % len(cyc)]
            mapping[c, r] = (nc, nr)
    flat = mapping.reshape(-1, 2)
    sigma_values = [polyops.generate_omegadelta_device(flat, dom.omega, DELTA, k, j * n, (j + 1) * n - 1, 2)
                    for j in range(2)]
    sigma_polys = dom.lagrange_to_coeff_many_device_inputs(sigma_values)
    pk = ProvingKey(dom, fixed_polys, l0, l_last, l_active, sigma_values, sigma_polys)
    pk.gates = custom_gates_evaluator([standard_gate()])
    pk.lookup = lookup_evaluator([Advice(3)], [Fixed(4)])
    return pk


# ---------------------------------------------------------------------- prove


def _grand_product_z(ratio: torch.Tensor, n: int, seed: int) -> torch.Tensor:
    """z[0] = 1, z[i + 1] = z[i] ratio[i], the last BLINDING_FACTORS rows random
    (permutation/prover.rs and lookup/prover.rs)."""
    lib = _lib.load()
    z = dv.empty_u64(4 * n)
    zv, rv = z.view(-1, 4), ratio.view(-1, 4)
    zv[0].copy_(torch.from_numpy(mont_limbs(1).view(np.int64)).to(z.device))
    zv[1:].copy_(rv[: n - 1])
    one = dv.to_device(mont_limbs(1).reshape(1, 4))
    _lib.check(lib._halo2_grand_product_device_inputs(ctypes.byref(_lib.obj(dv.ptr(z))),
                                                      ctypes.byref(_lib.obj(dv.ptr(one))), n, 0, dv.stream()),
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


def _check_closing(z: torch.Tensor, u: int) -> bool:
    return bool((z.view(-1, 4)[u].cpu().numpy().view(np.uint64) == mont_limbs(1)).all())


def create_proof(params: msm.ParamsKZG, pk: ProvingKey, circuit: Circuit, seed: int = 1) -> dict:
    """One proof; returns the transcript bytes, the per-phase times (ms, device-synchronised)
    and the self-checks."""
    lib = _lib.load()
    dom, n, u = pk.domain, 1 << circuit.k, circuit.usable
    tr = Transcript()
    times, checks = {}, {}
    t0 = time.perf_counter()

    def phase(name):
        nonlocal t0
        torch.cuda.current_stream().synchronize()  # this proof's stream only: proofs may run side by side
        t = time.perf_counter()
        times[name] = round((t - t0) * 1e3, 3)
        t0 = t

    def commit_lagrange_many(cols):
        return params.commit_lagrange_many_device(cols)

    main, side = torch.cuda.current_stream(), _side_stream()
    # 1. advice (prover.rs:282-466): commitments in Lagrange form; the coefficient forms
    #    are computed on the side stream meanwhile
    side.wait_stream(main)
    with torch.cuda.stream(side):
        advice_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.advice)
        instance_polys = dom.lagrange_to_coeff_many_device_inputs(circuit.instance)
    for pt in commit_lagrange_many(circuit.advice):
        tr.write_point(pt)
    main.wait_stream(side)
    phase("advice")

    # 2. lookup commit_permuted (lookup/prover.rs:552-627): theta-compression of one
    #    column is the column itself; permuted pair, blinding rows, commitments
    theta = tr.squeeze()
    inp, tab = circuit.advice[3], circuit.fixed[4]
    p_in, p_tab = dv.empty_u64(4 * n), dv.empty_u64(4 * n)
    permute_expression_pair_device(inp, tab, p_in, p_tab, n, u)
    _set_rows(p_in, u, msm.random_fr(n - u, seed * 7 + 1))
    _set_rows(p_tab, u, msm.random_fr(n - u, seed * 7 + 2))
    for pt in commit_lagrange_many([p_in, p_tab]):
        tr.write_point(pt)
    phase("lookup_permuted")

    # 3. grand products (permutation/prover.rs:69-273, lookup/prover.rs:294-461)
    beta, gamma = tr.squeeze(), tr.squeeze()
    ratio = dv.empty_u64(4 * n)
    polyops.poly_fill_one_device(ratio)  # the permutation product multiplies into it
    sc = [dv.to_device(mont_limbs(v).reshape(1, 4)) for v in (beta, gamma, DELTA, dom.omega, 1)]
    ws = int(lib._halo2_permutation_product_device_inputs_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib._halo2_permutation_product_device_inputs(
        ctypes.byref(_lib.obj(dv.ptr(ratio))), _lib.obj_array([dv.ptr(x) for x in pk.sigma_values]),
        _lib.obj_array([dv.ptr(x) for x in circuit.advice[:2]]), *[dv.ptr(x) for x in sc], n, 0, 2,
        dv.ptr(scratch), ws, dv.stream()), "_halo2_permutation_product_device_inputs")
    z_perm = _grand_product_z(ratio, n, seed * 7 + 3)
    lratio = lookup_product_device(p_in, p_tab, inp, tab, beta, gamma)
    z_lookup = _grand_product_z(lratio, n, seed * 7 + 4)
    # 4. the vanishing argument's random polynomial (vanishing/prover.rs:68-120) depends
    #    on no challenge: drawn as n uniform Lagrange values (a uniform polynomial of
    #    degree < n, as halo2's uniform coefficients), committed in the same batch as the
    #    grand products against g_lagrange; its transcript point still follows theirs
    random_values = msm.random_fr(n, seed * 7 + 5)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        z_perm_poly, z_lookup_poly, p_in_poly, p_tab_poly, random_poly = dom.lagrange_to_coeff_many_device_inputs(
            [z_perm, z_lookup, p_in, p_tab, random_values])
    z_perm_pt, z_lookup_pt, random_pt = commit_lagrange_many([z_perm, z_lookup, random_values])
    # read back after the commitment (which waits for the stream anyway), not before it
    checks["permutation_product_closes"] = _check_closing(z_perm, u)
    checks["lookup_product_closes"] = _check_closing(z_lookup, u)
    tr.write_point(z_perm_pt)
    tr.write_point(z_lookup_pt)
    main.wait_stream(side)
    phase("grand_products")

    tr.write_point(random_pt)
    y = tr.squeeze()
    phase("vanishing_commit")

    # 5. h(X) (evaluation.rs:1056-1417), divided by X^n - 1, in pieces of n
    num = evaluate_h_device(
        dom, pk.gates, pk.fixed_polys, pk.l0, pk.l_last, pk.l_active, advice_polys, instance_polys, [], y, beta,
        gamma, theta, lookups=[LookupCommitted(pk.lookup, z_lookup_poly, p_in, p_tab, p_in_poly, p_tab_poly)],
        permutation=PermutationCommitted([z_perm_poly], [("advice", 0), ("advice", 1)], pk.sigma_polys),
        blinding_factors=BLINDING_FACTORS, cs_degree=CS_DEGREE, key_parts=pk.key_parts)
    dom.divide_by_vanishing_poly_device(num)
    ntt.fft_normal_device(ntt.IFFT, dom.extended_k, num, num, dom.extended_omega_inv, dom.extended_ifft_divisor)
    ntt.distribute_powers_zeta_device(num, [dom.g_coset_inv, dom.g_coset])
    top = num.view(-1, 4)[(CS_DEGREE - 1) * n:]
    nonzero_top = top.any()  # read back after the commitment below
    pieces = [num[4 * i * n: 4 * (i + 1) * n] for i in range(CS_DEGREE - 1)]
    for pt in params.commit_many_device(pieces):
        tr.write_point(pt)
    checks["h_is_a_polynomial"] = not bool(nonzero_top.item())
    phase("quotient")

    # 6. evaluations at x (prover.rs:513-592), in halo2's order: instance, advice, fixed,
    #    the vanishing random polynomial, sigma, permutation products at x and omega x,
    #    lookup product at x and omega x, permuted input at x and omega^-1 x, permuted
    #    table at x. h is opened as h_0 + x^n h_1 + x^2n h_2 (vanishing/prover.rs:
    #    Constructed::evaluate); its value is not written, the verifier derives it.
    x = tr.squeeze()
    xw, xwi = x * dom.omega % R_MOD, x * dom.omega_inv % R_MOD
    queries = [(p, x) for p in instance_polys + advice_polys + pk.fixed_polys + [random_poly] + pk.sigma_polys]
    queries += [(z_perm_poly, x), (z_perm_poly, xw), (z_lookup_poly, x), (z_lookup_poly, xw), (p_in_poly, x),
                (p_in_poly, xwi), (p_tab_poly, x)]
    evals = polyops.batch_eval_polynomial_d2h([p for p, _ in queries], [pt for _, pt in queries])
    for e in evals:
        tr.write_scalar(int.from_bytes(np.ascontiguousarray(e).tobytes(), "little") * pow(1 << 256, -1, R_MOD) % R_MOD)
    xn = pow(x, n, R_MOD)
    h_poly = pieces[0].clone()
    for i in range(1, len(pieces)):
        polyops.poly_multiply_add_device(h_poly, pieces[i], pow(xn, i, R_MOD))
    queries.append((h_poly, x))
    phase("evaluations")

    # 7. SHPLONK multi-open (shplonk/prover.rs:97-441): y and v squeezed first, h written,
    #    then u squeezed from the transcript that holds h, then h_final written
    sy, sv = tr.squeeze(), tr.squeeze()
    ids = {}
    pq = [ProverQuery(pt, p, ids.setdefault(id(p), len(ids))) for p, pt in queries]

    def u_after_h(h):
        tr.write_point(h)
        return tr.squeeze()

    out = create_proof_device(params, pq, sy, sv, u_after_h)
    tr.write_point(out["h_final"])
    phase("opening")
    return {"proof": bytes(tr.proof), "times_ms": times, "checks": checks}
