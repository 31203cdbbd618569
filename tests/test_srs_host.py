"""SRS formats, host side (src/poly/kzg/commitment.rs:317-414): the G2 generator and
s_g2 = tau G2 arithmetic, Fq2 square roots and the G2 point encodings, pinned by the
curve equation, the group order and known constants (EIP-197's generator)."""
import random

from halo2_amd import srs
from halo2_amd.fields import P_MOD, R_MOD


def test_g2_generator_and_order():
    x, y = srs.G2_GENERATOR
    # EIP-197 lists the generator in decimal
    assert x == (10857046999023057135944570762232829481370756359578518086990519993285655852781,
                 11559732032986387107991004021392285783925812861821192530917403151452391805634)
    assert y == (8495653923123431417604973247489272438418190587263600148770280649306958101930,
                 4082367875863433681332203403145435568316851327593401208105741076214120093531)
    assert srs.g2_on_curve(srs.G2_GENERATOR)
    # order r: (r - 1) G2 + G2 = O
    assert srs.g2_add(srs.g2_mul(srs.G2_GENERATOR, R_MOD - 1), srs.G2_GENERATOR) is None


def test_g2_scalar_mul_is_a_homomorphism():
    rng = random.Random(3)
    a, b = rng.randrange(R_MOD), rng.randrange(R_MOD)
    g = srs.G2_GENERATOR
    ab = srs.g2_mul(g, a * b % R_MOD)
    assert ab == srs.g2_mul(srs.g2_mul(g, b), a)
    assert srs.g2_add(srs.g2_mul(g, a), srs.g2_mul(g, b)) == srs.g2_mul(g, (a + b) % R_MOD)
    assert srs.g2_on_curve(ab)


def test_fq2_sqrt():
    rng = random.Random(5)
    found_non_square = False
    for _ in range(40):
        a = (rng.randrange(P_MOD), rng.randrange(P_MOD))
        r = srs.fq2_sqrt(srs._fq2_mul(a, a))
        assert r is not None and srs._fq2_mul(r, r) == srs._fq2_mul(a, a)
        b = (rng.randrange(P_MOD), rng.randrange(P_MOD))
        rb = srs.fq2_sqrt(b)
        if rb is None:
            found_non_square = True
        else:
            assert srs._fq2_mul(rb, rb) == b
    assert found_non_square
    for c0 in (0, 1, 4, P_MOD - 1):  # purely real inputs, squares and non-squares of Fq
        r = srs.fq2_sqrt((c0, 0))
        assert r is not None and srs._fq2_mul(r, r) == (c0, 0)


def test_g2_encodings_round_trip_and_checks():
    import pytest

    p = srs.g2_mul(srs.G2_GENERATOR, 0xC0FFEE)
    for fmt in srs.SerdeFormat:
        b = srs.g2_to_bytes(p, fmt)
        assert len(b) == (64 if fmt == srs.SerdeFormat.Processed else 128)
        assert srs.g2_from_bytes(b, fmt) == p
        assert srs.g2_from_bytes(srs.g2_to_bytes(None, fmt), fmt) is None
    # raw words are Montgomery limbs: x.c0 * 2^256 mod p little-endian first
    raw = srs.g2_to_bytes(p, srs.SerdeFormat.RawBytes)
    assert int.from_bytes(raw[:32], "little") == p[0][0] * (1 << 256) % P_MOD
    # an off-curve point: RawBytes rejects it, RawBytesUnchecked reads it as is
    (x, y) = p
    bad = srs.g2_to_bytes((x, ((y[0] + 1) % P_MOD, y[1])), srs.SerdeFormat.RawBytesUnchecked)
    with pytest.raises(ValueError):
        srs.g2_from_bytes(bad, srs.SerdeFormat.RawBytes)
    assert srs.g2_from_bytes(bad, srs.SerdeFormat.RawBytesUnchecked)[1][0] == (y[0] + 1) % P_MOD
    # a coordinate word >= p
    big = bytearray(raw)
    big[0:32] = (P_MOD + 5).to_bytes(32, "little")
    with pytest.raises(ValueError):
        srs.g2_from_bytes(bytes(big), srs.SerdeFormat.RawBytes)
