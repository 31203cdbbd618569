"""Proving- and verifying-key bytes (halo2_amd.pk_io; ProvingKey::write / read of
src/plonk.rs:440-466, the round trip of tests/cross_prover_pk_equivalence.rs:252-330).

Host tests: the verifying key and the point encodings round trip in every SerdeFormat,
the layout is the documented one (u32 big-endian k and count, then the points), and
malformed bytes are refused. GPU tests: a k = 10 proving key written in each format
reads back to the same device words, writes back to the same bytes, and proves the same
proof bytes as the key it came from (and the proof verifies); RawBytes and Processed
refuse an element >= r, RawBytesUnchecked takes it. Byte identity with halo2-axiom's own
key files is not checkable here (the crate is not in the reference tree)."""
import io
import random

import pytest

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47


def _g1_mul(k: int):
    """k * (1, 2) on y^2 = x^3 + 3 (affine, None = identity)."""
    def add(a, b):
        if a is None:
            return b
        if b is None:
            return a
        if a[0] == b[0]:
            if (a[1] + b[1]) % P == 0:
                return None
            m = 3 * a[0] * a[0] * pow(2 * a[1], -1, P) % P
        else:
            m = (b[1] - a[1]) * pow(b[0] - a[0], -1, P) % P
        x = (m * m - a[0] - b[0]) % P
        return x, (m * (a[0] - x) - a[1]) % P
    acc, base = None, (1, 2)
    while k:
        if k & 1:
            acc = add(acc, base)
        base = add(base, base)
        k >>= 1
    return acc


def _vk():
    from halo2_amd import prover

    rng = random.Random(7)
    cs = prover.STANDARD_CS
    pts = [_g1_mul(rng.randrange(1, 1 << 64)) for _ in range(cs.num_fixed + len(cs.permutation_columns))]
    pts[1] = None  # an identity commitment (an all-zero fixed column)
    vk = prover.VerifyingKey(12, cs, pts[:cs.num_fixed], pts[cs.num_fixed:])
    vk.transcript_repr = vk.compute_transcript_repr()
    return vk


def test_point_encodings_round_trip():
    from halo2_amd import pk_io
    from halo2_amd.srs import SerdeFormat

    rng = random.Random(3)
    for pt in [None, (1, 2), _g1_mul(5)] + [_g1_mul(rng.randrange(1, 1 << 80)) for _ in range(6)]:
        for fmt in SerdeFormat:
            b = pk_io.point_to_bytes(pt, fmt)
            assert len(b) == (32 if fmt == SerdeFormat.Processed else 64)
            assert pk_io.point_from_bytes(b, fmt) == pt
    # the generator's encodings, written out: canonical x = 1 with y = 2 even; raw = x R, y R mod p
    assert pk_io.point_to_bytes((1, 2), SerdeFormat.Processed) == (1).to_bytes(32, "little")
    assert pk_io.point_to_bytes((1, P - 2), SerdeFormat.Processed) == (1 | 1 << 255).to_bytes(32, "little")
    assert pk_io.point_to_bytes((1, 2), SerdeFormat.RawBytes)[:32] == ((1 << 256) % P).to_bytes(32, "little")
    bad_x = P.to_bytes(32, "little")  # x = p
    with pytest.raises(ValueError):
        pk_io.point_from_bytes(bad_x, SerdeFormat.Processed)
    # x = 4: 4^3 + 3 = 67 is not a square mod p, so no point has x = 4
    assert pow(67, (P - 1) // 2, P) == P - 1
    with pytest.raises(ValueError):
        pk_io.point_from_bytes((4).to_bytes(32, "little"), SerdeFormat.Processed)
    off = ((1 << 256) % P).to_bytes(32, "little") + ((3 << 256) % P).to_bytes(32, "little")  # (1, 3)
    with pytest.raises(ValueError):
        pk_io.point_from_bytes(off, SerdeFormat.RawBytes)
    assert pk_io.point_from_bytes(off, SerdeFormat.RawBytesUnchecked) == (1, 3)


def test_vk_bytes_layout_and_round_trip():
    from halo2_amd import pk_io, prover
    from halo2_amd.srs import SerdeFormat

    vk = _vk()
    cs = prover.STANDARD_CS
    for fmt in SerdeFormat:
        b = pk_io.vk_to_bytes(vk, fmt)
        size = 32 if fmt == SerdeFormat.Processed else 64
        assert b[:4] == (12).to_bytes(4, "big") and b[4:8] == cs.num_fixed.to_bytes(4, "big")
        assert len(b) == 8 + size * (cs.num_fixed + len(cs.permutation_columns))
        assert b[8 + size: 8 + 2 * size] == bytes(size)  # the identity commitment
        back = pk_io.read_vk(io.BytesIO(b), fmt, cs)
        assert back.k == vk.k and back.fixed_commitments == vk.fixed_commitments
        assert back.permutation_commitments == vk.permutation_commitments
        assert back.transcript_repr == vk.transcript_repr
        with pytest.raises(EOFError):
            pk_io.read_vk(io.BytesIO(b[:-1]), fmt, cs)
    wrong = bytearray(pk_io.vk_to_bytes(vk, SerdeFormat.RawBytes))
    wrong[7] = 4  # 4 fixed commitments against the constraint system's 5
    with pytest.raises(ValueError):
        pk_io.read_vk(io.BytesIO(bytes(wrong)), SerdeFormat.RawBytes, cs)


@pytest.mark.gpu
def test_pk_bytes_round_trip_and_same_proof(lib):
    import torch

    import plonk_verifier as pv
    from halo2_amd import fields as F, msm, pk_io, prover
    from halo2_amd.srs import SerdeFormat

    circuit = prover.synthesize_standard(10, 21, table_bits=8)
    params = msm.ParamsKZG(10, 0x5EED_0F_7A0 % F.R_MOD, fixed_base=False)
    pk = prover.keygen(circuit, params)
    ref = prover.create_proof(params, pk, circuit, seed=5)
    cs = prover.STANDARD_CS
    n = 1 << 10
    sizes = {}
    for fmt in SerdeFormat:
        print("format", fmt.name, flush=True)
        b = pk_io.pk_to_bytes(pk, fmt)
        sizes[fmt] = len(b)
        back = pk_io.pk_from_bytes(b, fmt, cs)
        for a, c in zip([pk.l0, pk.l_last, pk.l_active] + pk.fixed_values + pk.fixed_polys + pk.sigma_values
                        + pk.sigma_polys, [back.l0, back.l_last, back.l_active] + back.fixed_values
                        + back.fixed_polys + back.sigma_values + back.sigma_polys):
            assert torch.equal(a.view(-1)[: 4 * n], c.view(-1)[: 4 * n])
        assert pk_io.pk_to_bytes(back, fmt) == b
        out = prover.create_proof(params, back, circuit, seed=5)
        assert out["proof"] == ref["proof"], fmt.name
        assert pv.verify_proof(back.vk, out["instance"], out["proof"], tau=params.tau)
    vk_raw = 8 + 64 * (cs.num_fixed + len(cs.permutation_columns))
    polys = 3 + 2 * cs.num_fixed + 2 * len(cs.permutation_columns)
    assert sizes[SerdeFormat.RawBytes] == vk_raw + polys * (4 + 32 * n) + 4 * 4
    # an element >= r (the first of l0): refused by RawBytes and Processed, taken by
    # RawBytesUnchecked
    for fmt in SerdeFormat:
        bad = bytearray(pk_io.pk_to_bytes(pk, fmt))
        at = len(pk_io.vk_to_bytes(pk.vk, fmt)) + 4
        bad[at: at + 32] = (F.R_MOD + 1).to_bytes(32, "little")
        if fmt == SerdeFormat.RawBytesUnchecked:
            pk_io.pk_from_bytes(bytes(bad), fmt, cs)
        else:
            with pytest.raises(ValueError):
                pk_io.pk_from_bytes(bytes(bad), fmt, cs)
    with pytest.raises(ValueError):
        pk_io.pk_from_bytes(pk_io.pk_to_bytes(pk, SerdeFormat.RawBytes) + b"\0", SerdeFormat.RawBytes, cs)
