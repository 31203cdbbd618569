"""SHPLONK multi-opening on the device, mirroring the GPU arm of the reference's
ProverSHPLONK::create_proof (src/poly/kzg/multiopen/shplonk/prover.rs:97-441) and
construct_intermediate_sets (src/poly/kzg/multiopen/shplonk.rs:51-160).

The transcript is not part of this path: the challenges y, v, u that the
reference squeezes from it are arguments here, and the two commitments it
writes (h, then h_final) are returned. Device polynomials are int64 torch
tensors of 4*n words (coefficient form); points and challenges are Python ints.
The small host folds (Lagrange interpolation over at most a few points, the
vanishing products) run on Python integers, as the reference runs them on the
CPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import device as dv
from . import polyops
from .fields import R_MOD, mont_limbs, value_of


@dataclass
class ProverQuery:
    """src/poly/query.rs ProverQuery: open `poly` (device coefficients) at `point`;
    `commitment` identifies the committed polynomial (queries on one polynomial share it)."""

    point: int
    poly: torch.Tensor
    commitment: object


@dataclass
class _Commitment:
    key: object
    poly: torch.Tensor
    evals: list[int]
    low_degree_equivalent: list[int] | None = None


@dataclass
class RotationSet:
    points: list[int]
    commitments: list[_Commitment]


def powers(x: int, count: int) -> list[int]:
    out, cur = [], 1
    for _ in range(count):
        out.append(cur)
        cur = cur * x % R_MOD
    return out


def evaluate_vanishing_polynomial(roots, z: int) -> int:
    """prod (z - root) (halo2-axiom arithmetic::evaluate_vanishing_polynomial)."""
    acc = 1
    for r in roots:
        acc = acc * (z - r) % R_MOD
    return acc


def lagrange_interpolate(points: list[int], evals: list[int]) -> list[int]:
    """Coefficients of the unique polynomial of degree < len(points) through (points, evals)
    (halo2-axiom arithmetic::lagrange_interpolate)."""
    m = len(points)
    out = [0] * m
    for j in range(m):
        basis = [1]
        denom = 1
        for k in range(m):
            if k == j:
                continue
            basis = [((basis[i - 1] if i > 0 else 0) - points[k] * (basis[i] if i < len(basis) else 0)) % R_MOD
                     for i in range(len(basis) + 1)]
            denom = denom * (points[j] - points[k]) % R_MOD
        scale = evals[j] * pow(denom, -1, R_MOD) % R_MOD
        for i, c in enumerate(basis):
            out[i] = (out[i] + c * scale) % R_MOD
    return out


def eval_host(coeffs: list[int], x: int) -> int:
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R_MOD
    return acc


def construct_intermediate_sets(queries: list[ProverQuery],
                                evals: list[int] | None = None) -> tuple[list[RotationSet], list[int]]:
    """shplonk.rs:51-160: evaluate every query once (batched on the device, unless the
    caller passes the values, query by query), group the points per commitment, then
    the commitments per point set. Commitment keys must be hashable. Points keep their
    first appearance order (the reference iterates hash sets; the results do not depend
    on the order)."""
    if evals is None:
        evals = [value_of(e) for e in
                 polyops.batch_eval_polynomial_d2h([q.poly for q in queries], [q.point for q in queries])]
    assert len(evals) == len(queries)
    eval_of = {(q.commitment, q.point % R_MOD): e % R_MOD for q, e in zip(queries, evals)}
    super_points: list[int] = []
    per_commitment: dict = {}  # key -> (poly, points), first-appearance order
    for q in queries:
        p = q.point % R_MOD
        if p not in super_points:
            super_points.append(p)
        _, pts = per_commitment.setdefault(q.commitment, (q.poly, []))
        if p not in pts:
            pts.append(p)
    sets: list[RotationSet] = []
    for key, (poly, pts) in per_commitment.items():
        for rs in sets:
            if set(rs.points) == set(pts):
                rs.commitments.append(_Commitment(key, poly, [eval_of[(key, p)] for p in rs.points]))
                break
        else:
            sets.append(RotationSet(list(pts), [_Commitment(key, poly, [eval_of[(key, p)] for p in pts])]))
    return sets, super_points


def _dev_table(vals: list[int]) -> torch.Tensor:
    return dv.to_device(np.stack([mont_limbs(v % R_MOD) for v in vals]))


class TableBuilder:
    """Host scalars gathered into one device table (one staged upload): add() returns a
    (start, count) slot, view() the slot's words on the device."""

    def __init__(self):
        self.vals: list[int] = []

    def add(self, vals: list[int]) -> tuple[int, int]:
        slot = (len(self.vals), len(vals))
        self.vals += [v % R_MOD for v in vals]
        return slot

    def upload(self) -> torch.Tensor:
        return _dev_table(self.vals or [0])

    @staticmethod
    def view(dev: torch.Tensor, slot: tuple[int, int]) -> torch.Tensor:
        return dev[4 * slot[0]:4 * (slot[0] + slot[1])]


def _zeros(n: int) -> torch.Tensor:
    return torch.zeros(4 * n, dtype=torch.int64, device=dv.device())


def create_proof_device(params, queries: list[ProverQuery], y: int, v: int, u,
                        evals: list[int] | None = None) -> dict:
    """prover.rs:97-441 with the transcript's challenges given: returns the two commitments
    ("h", "h_final", 96-byte projective points as (12,) u64) and the device polynomials
    h_x (n coefficients) and h_x_final (n - 1 coefficients). `u` is the challenge or, as
    in the protocol (u is squeezed after h is written, prover.rs:388-392), a function
    that takes h's commitment and returns u. `evals`, when given, are the queries'
    values (query by query), so they are not evaluated again."""
    n = params.n
    rotation_sets, super_points = construct_intermediate_sets(queries, evals)
    for rs in rotation_sets:
        for c in rs.commitments:
            c.low_degree_equivalent = lagrange_interpolate(rs.points, c.evals)
    max_c = max([len(rs.commitments) for rs in rotation_sets] + [1])

    # every scalar the first half needs, in one upload (each upload is a copy launch and
    # host work between the short launches of the sets): y and v powers, and per set
    # r(X)'s coefficients and the set's points
    table = TableBuilder()
    t_y = table.add(powers(y, max_c))
    t_v = table.add(powers(v, max(1, len(rotation_sets))))
    t_rx, t_pt = [], []
    for rs in rotation_sets:
        r_x = [0] * max(len(c.low_degree_equivalent) for c in rs.commitments)
        for c, py in zip(rs.commitments, powers(y, len(rs.commitments))):
            for i, coef in enumerate(c.low_degree_equivalent):
                r_x[i] = (r_x[i] + coef * py) % R_MOD
        t_rx.append(table.add(r_x))
        t_pt.append(table.add(rs.points))
    dev = table.upload()
    d_y_powers, d_v_powers = table.view(dev, t_y), table.view(dev, t_v)

    quotient_polys, p_xs = [], []
    # the rotation sets are independent until h is combined: each set's folds and
    # divisions (latency-bound chains of small launches) run on its own stream
    main = torch.cuda.current_stream()
    streams = [main] + dv.side_streams(min(len(rotation_sets), 3) - 1)
    for st in streams[1:]:
        st.wait_stream(main)
    for si, rs in enumerate(rotation_sets):  # quotient_contribution, prover.rs:183-268
        st = streams[si % len(streams)]
        with torch.cuda.stream(st):
            d_p_x = _zeros(n)
            polyops.poly_multiply_add_many_device(d_p_x, [c.poly for c in rs.commitments], d_y_powers)
            d_n_x = dv.empty_u64(4 * n)
            polyops.poly_sub_short_out_of_place_device(d_n_x, d_p_x, table.view(dev, t_rx[si]))
            d_q = d_n_x
            points = table.view(dev, t_pt[si])
            for i in range(len(rs.points)):
                d_root = points[4 * i:4 * i + 4]
                if i + 1 == len(rs.points):
                    d_q = polyops.kate_division_device_padded_with_d_root(d_q, d_root, n)
                else:
                    d_q = polyops.kate_division_device_with_d_root(d_q, d_root)
            quotient_polys.append(d_q)
            p_xs.append(d_p_x)

    for st in streams[1:]:
        main.wait_stream(st)
    for t in quotient_polys + p_xs:  # possibly made on a side stream, used on this one from here
        t.record_stream(main)
    d_h_x = _zeros(n)
    polyops.poly_multiply_add_many_device(d_h_x, quotient_polys, d_v_powers)
    h = params.commit_device(d_h_x)
    if callable(u):
        u = u(h) % R_MOD

    # the second half's scalars (they depend on u), again in one upload
    z_diffs, r_is = [], []
    for rs in rotation_sets:  # linearisation_contribution, prover.rs:331-386
        diffs = [p for p in super_points if p not in rs.points]
        z_diffs.append(evaluate_vanishing_polynomial(diffs, u))
        r_i = 0
        for c, py in zip(rs.commitments, powers(y, len(rs.commitments))):
            r_i = (r_i + eval_host(c.low_degree_equivalent, u) * py) % R_MOD
        r_is.append(r_i)
    zt_eval = evaluate_vanishing_polynomial(super_points, u)
    z0_inv = pow(z_diffs[0], -1, R_MOD)
    table2 = TableBuilder()
    t_r = table2.add(r_is)
    t_z = table2.add([z - 1 for z in z_diffs])
    t_last = table2.add([-zt_eval, z0_inv - 1, u])
    dev2 = table2.upload()
    d_r, d_z, d_last = table2.view(dev2, t_r), table2.view(dev2, t_z), table2.view(dev2, t_last)
    lin = []
    for si, d_l_x in enumerate(p_xs):
        polyops.poly_sub_scalar_at_zero_device(d_l_x, d_r[4 * si:4 * si + 4])
        polyops.poly_scale_device_with_d_s_minus_one(d_l_x, d_z[4 * si:4 * si + 4])
        lin.append(d_l_x)
    d_l_x = _zeros(n)
    polyops.poly_multiply_add_many_device(d_l_x, lin, d_v_powers)
    polyops.poly_multiply_add_device_with_d_scalar(d_l_x, d_h_x, d_last[0:4])
    d_h_final = polyops.kate_division_device_with_d_root(d_l_x, d_last[8:12])
    polyops.poly_scale_device_with_d_s_minus_one(d_h_final, d_last[4:8])
    h_final = params.commit_device(d_h_final)
    return {"h": h, "h_final": h_final, "h_x": d_h_x, "h_x_final": d_h_final, "l_x": d_l_x}
