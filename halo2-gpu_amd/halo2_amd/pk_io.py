"""Proving- and verifying-key bytes: ProvingKey::write / read / to_bytes and
VerifyingKey::write / read (src/plonk.rs:440-466 of the reference, which delegates them
to halo2-axiom's canonical keys; the round trip CPU keygen -> bytes -> GPU proof is the
reference's tests/cross_prover_pk_equivalence.rs:252-330).

Layout, in halo2's serialization order and conventions (lengths and counts as u32
big-endian, field elements and points by SerdeFormat as in srs.py):

  verifying key   k; the number of fixed commitments and each commitment; the
                  permutation's sigma commitments (their count comes from the
                  constraint system's equality columns). The constraint system has no
                  selector columns, so no packed selector bits follow.
  proving key     the verifying key; l0, l_last, l_active_row; the fixed columns'
                  Lagrange values, then their coefficients; the permutation's sigma
                  columns' Lagrange values, then their coefficients. A polynomial is its
                  length and its elements; a list of polynomials is its count and each.

Field elements: RawBytes / RawBytesUnchecked write the 32-byte Montgomery word the
device holds (four little-endian 64-bit limbs); Processed writes the canonical value
little-endian. RawBytes and Processed reject a value >= r on read. Points: as srs.py
(Processed: 32 bytes, canonical x with the parity of y in the top bit, all zeros for
the identity; raw: x and y as Montgomery words, 64 zero bytes for the identity).

Differences from halo2-axiom's own bytes, which cannot be checked here (the crate is
not part of the reference tree): l0, l_last and l_active_row are held and written in
coefficient form (the form the prover consumes), the extended-domain cosets of the
fixed and sigma columns are not written (they are recomputed per extended part on
first use), and the circuit's shape is passed to `read_pk` as the ConstraintSystem
instead of being rebuilt by a Circuit type's configure().
"""
from __future__ import annotations

import io

import numpy as np
import torch

from . import device as dv
from .domain import EvaluationDomain
from .fields import P_MOD, R_MOD, limbs
from .polyops import poly_multiply_add_device
from .quotient import custom_gates_evaluator, lookup_evaluator
from .srs import SerdeFormat, _read_exact

_R_LIMBS = limbs(R_MOD).astype(np.uint64)
_MONT_R = (1 << 256) % R_MOD  # R mod r: canonical word -> Montgomery word
_MONT_R_INV = pow(1 << 256, -1, R_MOD)  # R^-1 mod r: Montgomery word -> canonical word


def _u32(v: int) -> bytes:
    return int(v).to_bytes(4, "big")


def _read_u32(reader) -> int:
    return int.from_bytes(_read_exact(reader, 4), "big")


# ---------------------------------------------------------------------- points


def point_to_bytes(pt: tuple | None, fmt: SerdeFormat) -> bytes:
    """One G1 point (canonical affine integers, None = identity)."""
    if fmt == SerdeFormat.Processed:
        if pt is None:
            return bytes(32)
        x, y = pt
        return (x | ((y & 1) << 255)).to_bytes(32, "little")
    if pt is None:
        return bytes(64)
    return b"".join((v * (1 << 256) % P_MOD).to_bytes(32, "little") for v in pt)


def _on_curve(x: int, y: int) -> bool:
    return (y * y - x * x * x - 3) % P_MOD == 0


def point_from_bytes(b: bytes, fmt: SerdeFormat) -> tuple | None:
    if fmt == SerdeFormat.Processed:
        v = int.from_bytes(b, "little")
        sign, x = v >> 255, v & ((1 << 255) - 1)
        if x == 0 and sign == 0:
            return None
        if x >= P_MOD:
            raise ValueError("invalid point encoding")
        rhs = (x * x * x + 3) % P_MOD
        y = pow(rhs, (P_MOD + 1) // 4, P_MOD)  # p = 3 mod 4
        if y * y % P_MOD != rhs:
            raise ValueError("invalid point encoding")
        if (y & 1) != sign:
            y = P_MOD - y
        return x, y
    xw, yw = int.from_bytes(b[:32], "little"), int.from_bytes(b[32:], "little")
    if xw == 0 and yw == 0:
        return None
    if fmt == SerdeFormat.RawBytes and (xw >= P_MOD or yw >= P_MOD):
        raise ValueError("invalid point encoding")
    inv = pow(1 << 256, -1, P_MOD)
    x, y = xw * inv % P_MOD, yw * inv % P_MOD
    if fmt == SerdeFormat.RawBytes and not _on_curve(x, y):
        raise ValueError("invalid point encoding")
    return x, y


def _point_size(fmt: SerdeFormat) -> int:
    return 32 if fmt == SerdeFormat.Processed else 64


# ----------------------------------------------------------------- polynomials


def _below_r(w: np.ndarray) -> np.ndarray:
    """Rows of 4 little-endian u64 limbs: value < r, per row."""
    lt = np.zeros(len(w), dtype=bool)
    eq = np.ones(len(w), dtype=bool)
    for i in (3, 2, 1, 0):
        lt |= eq & (w[:, i] < _R_LIMBS[i])
        eq &= w[:, i] == _R_LIMBS[i]
    return lt


def _scaled(d: torch.Tensor, factor: int) -> torch.Tensor:
    """Every element times a constant on the device (a Montgomery product)."""
    out = torch.zeros_like(d)
    poly_multiply_add_device(out, d, factor)
    return out


def poly_to_bytes(d: torch.Tensor, fmt: SerdeFormat) -> bytes:
    n = d.numel() // 4
    words = _scaled(d, _MONT_R_INV) if fmt == SerdeFormat.Processed else d
    return _u32(n) + dv.to_host(words).astype("<u8").tobytes()


def poly_from_bytes(reader, fmt: SerdeFormat, n_expected: int | None = None) -> torch.Tensor:
    n = _read_u32(reader)
    if n_expected is not None and n != n_expected:
        raise ValueError(f"polynomial of {n} elements where {n_expected} are expected")
    w = np.frombuffer(_read_exact(reader, 32 * n), dtype="<u8").reshape(n, 4).astype(np.uint64)
    if fmt != SerdeFormat.RawBytesUnchecked and n and not _below_r(w).all():
        raise ValueError("field element not below the modulus")
    d = dv.to_device(np.ascontiguousarray(w))
    return _scaled(d, _MONT_R) if fmt == SerdeFormat.Processed else d


def _polys_to_bytes(polys: list, fmt: SerdeFormat) -> bytes:
    return _u32(len(polys)) + b"".join(poly_to_bytes(p, fmt) for p in polys)


def _polys_from_bytes(reader, fmt: SerdeFormat, n: int, count: int | None = None) -> list:
    m = _read_u32(reader)
    if count is not None and m != count:
        raise ValueError(f"{m} polynomials where {count} are expected")
    return [poly_from_bytes(reader, fmt, n) for _ in range(m)]


# ------------------------------------------------------------ verifying key


def write_vk(vk, writer, fmt: SerdeFormat) -> None:
    writer.write(_u32(vk.k))
    writer.write(_u32(len(vk.fixed_commitments)))
    for pt in vk.fixed_commitments:
        writer.write(point_to_bytes(pt, fmt))
    for pt in vk.permutation_commitments:
        writer.write(point_to_bytes(pt, fmt))


def read_vk(reader, fmt: SerdeFormat, cs):
    """The verifying key of a circuit with constraint system `cs`; its digest is
    recomputed from what was read."""
    from .prover import VerifyingKey

    k = _read_u32(reader)
    nf = _read_u32(reader)
    if nf != cs.num_fixed:
        raise ValueError(f"{nf} fixed commitments where the constraint system has {cs.num_fixed} fixed columns")
    size = _point_size(fmt)
    fixed = [point_from_bytes(_read_exact(reader, size), fmt) for _ in range(nf)]
    perm = [point_from_bytes(_read_exact(reader, size), fmt) for _ in cs.permutation_columns]
    vk = VerifyingKey(k, cs, fixed, perm)
    vk.transcript_repr = vk.compute_transcript_repr()
    return vk


def vk_to_bytes(vk, fmt: SerdeFormat) -> bytes:
    buf = io.BytesIO()
    write_vk(vk, buf, fmt)
    return buf.getvalue()


# -------------------------------------------------------------- proving key


def write_pk(pk, writer, fmt: SerdeFormat) -> None:
    if pk.vk is None:
        raise ValueError("the proving key has no verifying key (keygen with params)")
    if len(pk.fixed_values) != len(pk.fixed_polys):
        raise ValueError("the proving key holds no fixed column values")
    write_vk(pk.vk, writer, fmt)
    for p in (pk.l0, pk.l_last, pk.l_active):
        writer.write(poly_to_bytes(p, fmt))
    writer.write(_polys_to_bytes(pk.fixed_values, fmt))
    writer.write(_polys_to_bytes(pk.fixed_polys, fmt))
    writer.write(_polys_to_bytes(pk.sigma_values, fmt))
    writer.write(_polys_to_bytes(pk.sigma_polys, fmt))


def pk_to_bytes(pk, fmt: SerdeFormat) -> bytes:
    buf = io.BytesIO()
    write_pk(pk, buf, fmt)
    return buf.getvalue()


def read_pk(reader, fmt: SerdeFormat, cs):
    """A proving key, device-resident, for a circuit with constraint system `cs`
    (halo2-axiom's ProvingKey::read::<_, ConcreteCircuit>)."""
    from .prover import ProvingKey

    vk = read_vk(reader, fmt, cs)
    n = 1 << vk.k
    l0, l_last, l_active = (poly_from_bytes(reader, fmt, n) for _ in range(3))
    fixed_values = _polys_from_bytes(reader, fmt, n, cs.num_fixed)
    fixed_polys = _polys_from_bytes(reader, fmt, n, cs.num_fixed)
    m = len(cs.permutation_columns)
    sigma_values = _polys_from_bytes(reader, fmt, n, m)
    sigma_polys = _polys_from_bytes(reader, fmt, n, m)
    pk = ProvingKey(EvaluationDomain(cs.degree, vk.k), fixed_polys, l0, l_last, l_active, sigma_values, sigma_polys,
                    vk=vk, fixed_values=fixed_values)
    pk.gates = custom_gates_evaluator(list(cs.gates))
    (inputs, tables), = cs.lookups
    pk.lookup = lookup_evaluator(list(inputs), list(tables))
    return pk


def pk_from_bytes(data: bytes, fmt: SerdeFormat, cs):
    reader = io.BytesIO(data)
    pk = read_pk(reader, fmt, cs)
    if reader.read(1):
        raise ValueError("trailing bytes after the proving key")
    return pk
