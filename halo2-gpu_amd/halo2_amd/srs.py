"""SRS files: ParamsKZG::write_custom / read_custom and write / read
(src/poly/kzg/commitment.rs:317-414 and :510-518 of the reference).

File layout (commitment.rs:317-333): k as u32 little-endian, then the n = 2^k
points of g, the n points of g_lagrange, then g2 and s_g2. Point encodings by
SerdeFormat (the halo2curves-axiom 0.7 enum the reference re-exports):

* RawBytes / RawBytesUnchecked: G1 = x, y as 4 little-endian 64-bit
  Montgomery limbs each (64 bytes, identity (0, 0)), the in-memory and device
  layout of the bases; G2 = x.c0, x.c1, y.c0, y.c1 the same way (128 bytes).
  RawBytes rejects a coordinate >= p or a point off its curve; the unchecked
  form reads as is. `write`/`read` use RawBytesUnchecked (commitment.rs:510-518).
* Processed: G1 = canonical x little-endian with the parity of canonical y in
  bit 7 of the last byte (32 bytes); G2 = canonical x.c0 || x.c1 with the parity
  of canonical y.c0 in the top bit (64 bytes); all-zero bytes are the identity.

The G1 halves are encoded, validated and decompressed on the device
(h2amd_g1_*); the two G2 points are handled here with Python integers.
"""
from __future__ import annotations

import ctypes
import enum
import io

import numpy as np
import torch

from . import _lib
from . import device as dv
from .fields import P_MOD, R_MOD

# ------------------------------------------------------------------ Fq2 (u^2 = -1)


def _fq2_add(a, b):
    return ((a[0] + b[0]) % P_MOD, (a[1] + b[1]) % P_MOD)


def _fq2_sub(a, b):
    return ((a[0] - b[0]) % P_MOD, (a[1] - b[1]) % P_MOD)


def _fq2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P_MOD, (a[0] * b[1] + a[1] * b[0]) % P_MOD)


def _fq2_inv(a):
    d = pow((a[0] * a[0] + a[1] * a[1]) % P_MOD, -1, P_MOD)
    return (a[0] * d % P_MOD, -a[1] * d % P_MOD)


def _fq_sqrt(a: int) -> int | None:
    r = pow(a, (P_MOD + 1) // 4, P_MOD)  # p = 3 mod 4
    return r if r * r % P_MOD == a % P_MOD else None


def fq2_sqrt(a) -> tuple[int, int] | None:
    """A square root in Fq2 = Fq[u]/(u^2 + 1), or None: with the norm's root t,
    x0^2 = (a0 + t) / 2 (or (a0 - t) / 2) and x1 = a1 / (2 x0)."""
    a0, a1 = a[0] % P_MOD, a[1] % P_MOD
    if a1 == 0:
        r = _fq_sqrt(a0)
        if r is not None:
            return (r, 0)
        r = _fq_sqrt(-a0 % P_MOD)  # sqrt(a0) = u * sqrt(-a0)
        return None if r is None else (0, r)
    t = _fq_sqrt((a0 * a0 + a1 * a1) % P_MOD)
    if t is None:
        return None
    half = pow(2, -1, P_MOD)
    for d in ((a0 + t) * half % P_MOD, (a0 - t) * half % P_MOD):
        x0 = _fq_sqrt(d)
        if x0:
            x1 = a1 * pow(2 * x0, -1, P_MOD) % P_MOD
            if _fq2_mul((x0, x1), (x0, x1)) == (a0, a1):
                return (x0, x1)
    return None


# ------------------------------------------------------------ G2 (the twist)

# y^2 = x^3 + 3 / (9 + u)
G2_B = _fq2_mul((3, 0), _fq2_inv((9, 1)))
# the standard BN254 G2 generator (EIP-197), coordinates as (c0, c1)
G2_GENERATOR = (
    (0x1800DEEF121F1E76426A00665E5C4479674322D4F75EDADD46DEBD5CD992F6ED,
     0x198E9393920D483A7260BFB731FB5D25F1AA493335A9E71297E485B7AEF312C2),
    (0x12C85EA5DB8C6DEB4AAB71808DCB408FE3D1E7690C43D37B4CE6CC0166FA7DAA,
     0x090689D0585FF075EC9E99AD690C3395BC4B313370B38EF355ACDADCD122975B),
)


def g2_on_curve(p) -> bool:
    if p is None:
        return True
    x, y = p
    return _fq2_mul(y, y) == _fq2_add(_fq2_mul(_fq2_mul(x, x), x), G2_B)


def g2_add(p, q):
    """Affine addition (None = identity)."""
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if _fq2_add(y1, y2) == (0, 0):
            return None
        lam = _fq2_mul(_fq2_mul((3, 0), _fq2_mul(x1, x1)), _fq2_inv(_fq2_add(y1, y1)))
    else:
        lam = _fq2_mul(_fq2_sub(y2, y1), _fq2_inv(_fq2_sub(x2, x1)))
    x3 = _fq2_sub(_fq2_sub(_fq2_mul(lam, lam), x1), x2)
    return (x3, _fq2_sub(_fq2_mul(lam, _fq2_sub(x1, x3)), y1))


def g2_mul(p, s: int):
    acc, base = None, p
    s %= R_MOD
    while s:
        if s & 1:
            acc = g2_add(acc, base)
        base = g2_add(base, base)
        s >>= 1
    return acc


def _fq_raw(v: int) -> bytes:
    return (v * (1 << 256) % P_MOD).to_bytes(32, "little")


def _fq_from_raw(b: bytes) -> int:
    return int.from_bytes(b, "little") * pow(1 << 256, -1, P_MOD) % P_MOD


def g2_to_bytes(p, fmt: "SerdeFormat") -> bytes:
    if fmt == SerdeFormat.Processed:
        if p is None:
            return bytes(64)
        (x0, x1), (y0, _) = p
        out = bytearray(x0.to_bytes(32, "little") + x1.to_bytes(32, "little"))
        out[63] |= (y0 & 1) << 7
        return bytes(out)
    if p is None:
        return bytes(128)
    (x0, x1), (y0, y1) = p
    return b"".join(_fq_raw(v) for v in (x0, x1, y0, y1))


def g2_from_bytes(b: bytes, fmt: "SerdeFormat"):
    if fmt == SerdeFormat.Processed:
        sign = b[63] >> 7
        raw = bytearray(b)
        raw[63] &= 0x7F
        x0, x1 = int.from_bytes(raw[:32], "little"), int.from_bytes(raw[32:], "little")
        if x0 >= P_MOD or x1 >= P_MOD:
            raise ValueError("invalid point encoding")
        if x0 == 0 and x1 == 0 and not sign:
            return None
        x = (x0, x1)
        y = fq2_sqrt(_fq2_add(_fq2_mul(_fq2_mul(x, x), x), G2_B))
        if y is None:
            raise ValueError("invalid point encoding")
        if (y[0] & 1) != sign:
            y = ((-y[0]) % P_MOD, (-y[1]) % P_MOD)
        return (x, y)
    words = [int.from_bytes(b[32 * i: 32 * i + 32], "little") for i in range(4)]
    if fmt == SerdeFormat.RawBytes and any(w >= P_MOD for w in words):
        raise ValueError("invalid point encoding")
    if not any(words):
        return None
    x0, x1, y0, y1 = (w * pow(1 << 256, -1, P_MOD) % P_MOD for w in words)
    p = ((x0, x1), (y0, y1))
    if fmt == SerdeFormat.RawBytes and not g2_on_curve(p):
        raise ValueError("invalid point encoding")
    return p


# ----------------------------------------------------------------- formats


class SerdeFormat(enum.IntEnum):
    """halo2curves SerdeFormat (re-exported by the reference, lib.rs:40)."""

    Processed = 0
    RawBytes = 1
    RawBytesUnchecked = 2


def _g1_flags_call(fn, *args) -> tuple[int, int]:
    lib = _lib.load()
    ws = int(lib.h2amd_g1_flags_workspace_size())
    scratch = dv.empty_bytes(ws)
    bad, first = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _lib.check(getattr(lib, fn)(*args, ctypes.byref(bad), ctypes.byref(first), dv.ptr(scratch), ws, dv.stream()),
               fn)
    return bad.value, first.value


def g1_validate_device(d_points: torch.Tensor) -> tuple[int, int]:
    """(number of invalid points, index of the first) of device affine points: a
    coordinate >= p or a point off y^2 = x^3 + 3 (the identity (0, 0) is valid)."""
    n = d_points.numel() // 8
    return _g1_flags_call("h2amd_g1_validate", dv.ptr(d_points), n)


def g1_compress_device(d_points: torch.Tensor) -> torch.Tensor:
    """Processed encoding of device affine points: (n * 32) bytes as a uint8 tensor."""
    n = d_points.numel() // 8
    out = torch.empty(32 * max(1, n), dtype=torch.uint8, device=d_points.device)
    _lib.check(_lib.load().h2amd_g1_compress(dv.ptr(out), dv.ptr(d_points), n, dv.stream()), "h2amd_g1_compress")
    return out[: 32 * n]


def g1_decompress_device(d_bytes: torch.Tensor) -> tuple[torch.Tensor, int, int]:
    """Affine points from Processed bytes on the device; (points, invalid count, first
    invalid index). Invalid encodings decode to the identity."""
    n = d_bytes.numel() // 32
    pts = dv.empty_u64(8 * max(1, n))
    bad, first = _g1_flags_call("h2amd_g1_decompress", dv.ptr(pts), dv.ptr(d_bytes), n)
    return pts[: 8 * n], bad, first


def _g1_bytes(d_points: torch.Tensor, fmt: SerdeFormat) -> bytes:
    if fmt == SerdeFormat.Processed:
        return g1_compress_device(d_points).cpu().numpy().tobytes()
    return dv.to_host(d_points, cols=8).tobytes()


def _g1_read(data: bytes, n: int, fmt: SerdeFormat) -> torch.Tensor:
    if fmt == SerdeFormat.Processed:
        raw = torch.from_numpy(np.frombuffer(data, dtype=np.uint8).copy()).to(dv.device())
        pts, bad, _ = g1_decompress_device(raw)
        if bad:
            raise ValueError("invalid point encoding")
        return pts
    pts = dv.to_device(np.frombuffer(data, dtype=np.uint64).reshape(n, 8).copy())
    if fmt == SerdeFormat.RawBytes:
        bad, _ = g1_validate_device(pts)
        if bad:
            raise ValueError("invalid point encoding")
    return pts


def write_custom(params, writer: io.RawIOBase, fmt: SerdeFormat) -> None:
    """commitment.rs:317-333: k, g, g_lagrange, g2, s_g2."""
    writer.write(int(params.k).to_bytes(4, "little"))
    writer.write(_g1_bytes(params.g_device, fmt))
    writer.write(_g1_bytes(params.g_lagrange_device, fmt))
    writer.write(g2_to_bytes(params.g2, fmt))
    writer.write(g2_to_bytes(params.s_g2, fmt))


def _read_exact(reader, size: int) -> bytes:
    data = reader.read(size)
    if len(data) != size:
        raise EOFError("failed to fill whole buffer")
    return data


def read_custom(reader, fmt: SerdeFormat, fixed_base: bool = True):
    """commitment.rs:335-414: ParamsKZG with g and g_lagrange resident on the device."""
    from .msm import ParamsKZG

    k = int.from_bytes(_read_exact(reader, 4), "little")
    n = 1 << k
    g1_size = 32 if fmt == SerdeFormat.Processed else 64
    g = _g1_read(_read_exact(reader, g1_size * n), n, fmt)
    g_lagrange = _g1_read(_read_exact(reader, g1_size * n), n, fmt)
    g2_size = 64 if fmt == SerdeFormat.Processed else 128
    g2 = g2_from_bytes(_read_exact(reader, g2_size), fmt)
    s_g2 = g2_from_bytes(_read_exact(reader, g2_size), fmt)
    return ParamsKZG.from_parts(k, g, g_lagrange, g2, s_g2, fixed_base=fixed_base)


def write(params, writer) -> None:
    """Params::write (commitment.rs:510-513): RawBytesUnchecked."""
    write_custom(params, writer, SerdeFormat.RawBytesUnchecked)


def read(reader, fixed_base: bool = True):
    """Params::read (commitment.rs:515-518): RawBytesUnchecked."""
    return read_custom(reader, SerdeFormat.RawBytesUnchecked, fixed_base)
