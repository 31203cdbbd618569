"""CPU pins of the checker verifier's building blocks (tests/plonk_verifier.py,
tests/bn254_pairing.py): the pairing is bilinear, non-degenerate and lands in the
order-r subgroup; the KZG check accepts a genuine opening identity and rejects a wrong
one; points decompress to the curve with the encoded parity."""
import random

import pytest


def _g2():
    import sys
    import os

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "halo2-gpu_amd"))
    from halo2_amd.srs import G2_GENERATOR, g2_mul

    return G2_GENERATOR, g2_mul


def test_pairing_bilinear_nondegenerate():
    import bn254_pairing as bp
    import plonk_verifier as pv

    g2, g2_mul = _g2()
    e = bp.pairing(pv.G1, g2)
    assert e != bp.f12_one()
    assert bp.f12_pow(e, bp.R) == bp.f12_one()
    a, b = 0x1234567, 0x89ABCDEF
    pa = pv.to_affine(pv.g1_mul(pv.G1, a))
    qb = g2_mul(g2, b)
    assert bp.pairing(pa, qb) == bp.f12_pow(e, a * b % bp.R)
    # e(aP, Q) e(-aP, Q) = 1 and e(P, bQ) = e(bP, Q)
    neg = (pa[0], (-pa[1]) % pv.P)
    assert bp.pairing_product_is_one([(pa, g2), (neg, g2)])
    assert bp.pairing(pv.G1, qb) == bp.pairing(pv.to_affine(pv.g1_mul(pv.G1, b)), g2)


def test_kzg_check_accepts_and_rejects():
    """e(W, tau G2) = e(tau W, G2) holds; a wrong right-hand point fails."""
    import bn254_pairing as bp
    import plonk_verifier as pv

    g2, g2_mul = _g2()
    tau = 0x5EED_1234_5678
    w = pv.to_affine(pv.g1_mul(pv.G1, 777))
    right = pv.to_affine(pv.g1_mul(w, tau))
    s_g2 = g2_mul(g2, tau)
    assert bp.kzg_check(w, s_g2, right, g2)
    assert not bp.kzg_check(w, s_g2, pv.to_affine(pv.g1_mul(w, tau + 1)), g2)


def test_point_decompression_round_trip():
    import plonk_verifier as pv

    rng = random.Random(5)
    for _ in range(20):
        pt = pv.to_affine(pv.g1_mul(pv.G1, rng.randrange(1, pv.R)))
        enc = (pt[0] | ((pt[1] & 1) << 255)).to_bytes(32, "little")
        assert pv.decompress(enc) == pt
    with pytest.raises(pv.VerifyError):
        pv.decompress((pv.P + 1).to_bytes(32, "little"))


def test_lagrange_and_vanishing_helpers():
    import plonk_verifier as pv

    rng = random.Random(6)
    pts = [rng.randrange(pv.R) for _ in range(3)]
    coeffs = [rng.randrange(pv.R) for _ in range(3)]
    f = lambda z: sum(c * pow(z, i, pv.R) for i, c in enumerate(coeffs)) % pv.R  # noqa: E731
    z = rng.randrange(pv.R)
    assert pv._lagrange_eval(pts, [f(p) for p in pts], z) == f(z)
    assert pv._vanishing(pts, pts[1]) == 0
