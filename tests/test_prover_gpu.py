"""The device proving pipeline (halo2_amd.prover, create_proof's protocol order,
src/plonk/prover.rs:141-629) checked by a verifier that restates halo2's verify_proof
and SHPLONK verifier (tests/plonk_verifier.py), with the final KZG check done both as
the pairing e(left, [tau]G2) = e(right, G2) and as the group identity tau * left = right
(the SRS comes from a seeded tau). The circuit's permutation spans two sets (a, b | e,
instance), so the cross-set carry is part of what the verifier checks.

Covered: the k = 20 proof of the benchmark circuit (16-bit range table) is accepted;
one flipped byte, wrong public inputs or a proof under another key are rejected; a
witness broken in a copied cell of the second permutation set, in a gate or in the
lookup breaks exactly the matching prover self-checks and the verifier rejects the
proof; same seed, same bytes; fixed-base and plain MSMs give the same bytes; a fresh
witness proves under the same key."""
import pytest

pytestmark = pytest.mark.gpu

TAU = 0x5EED_0F_7A0


def _setup(k, seed, table_bits=8, fixed_base=False, mutate=None):
    from halo2_amd import fields as F, msm, prover

    circuit = prover.synthesize_standard(k, seed, table_bits=table_bits)
    if mutate:
        mutate(circuit)
    params = msm.ParamsKZG(k, TAU % F.R_MOD, fixed_base=fixed_base)
    pk = prover.keygen(circuit, params)
    return params, pk, circuit


def _verify(params, pk, public, proof, pairing=True):
    import bn254_pairing as bp
    import plonk_verifier as pv

    ok_tau = pv.verify_proof(pk.vk, public, proof, tau=params.tau)
    if not pairing:
        return ok_tau
    ok_pair = pv.verify_proof(pk.vk, public, proof, pairing_check=bp.kzg_check, g2=params.g2, s_g2=params.s_g2)
    assert ok_tau == ok_pair
    return ok_pair


def _rejects(params, pk, public, proof):
    import plonk_verifier as pv

    try:
        return not pv.verify_proof(pk.vk, public, proof, tau=params.tau)
    except pv.VerifyError:
        return True


def test_proof_verifies_and_self_checks(lib):
    from halo2_amd import prover

    params, pk, circuit = _setup(10, 3)
    out = prover.create_proof(params, pk, circuit, seed=3)
    print("k=10 proof", len(out["proof"]), "bytes", flush=True)
    assert out["checks"] == {"permutation_product_closes": True, "lookup_product_closes": True,
                             "h_is_a_polynomial": True}
    assert set(out["times_ms"]) >= {"advice", "quotient", "opening"}
    assert len(out["instance"]) == prover.NUM_PUBLIC
    assert _verify(params, pk, out["instance"], out["proof"])
    assert prover.create_proof(params, pk, circuit, seed=3)["proof"] == out["proof"]  # deterministic
    # the fixed-base tables give the same commitments, hence the same transcript
    fparams, fpk, fcircuit = _setup(10, 3, fixed_base=True)
    assert prover.create_proof(fparams, fpk, fcircuit, seed=3)["proof"] == out["proof"]


def test_verifier_rejects_tampering(lib):
    from halo2_amd import prover

    params, pk, circuit = _setup(10, 8)
    out = prover.create_proof(params, pk, circuit, seed=8)
    proof, public = out["proof"], out["instance"]
    assert _verify(params, pk, public, proof, pairing=False)
    for pos in (5, 32 * 7 + 3, len(proof) - 40, len(proof) - 1):  # points and scalars
        bad = bytearray(proof)
        bad[pos] ^= 0x04
        assert _rejects(params, pk, public, bytes(bad)), pos
    assert _rejects(params, pk, public, proof[:-32])
    wrong = list(public)
    wrong[3] = (wrong[3] + 1) % prover.R_MOD
    assert _rejects(params, pk, wrong, proof)
    # the same proof against another circuit's key
    _, pk2, _ = _setup(10, 9)
    assert _rejects(params, pk2, public, proof)


def test_bad_witness_is_caught(lib):
    from halo2_amd import device as dv, fields as F, prover

    def break_copy(c):  # a copied cell of the second permutation set (column e)
        cyc = next(cy for cy in c.cycles if any(col == 2 for col, _ in cy[1:]))
        col, row = next((col, row) for col, row in cyc[1:] if col == 2)
        c.advice[prover.E].view(-1, 4)[row].add_(1)

    def break_gate(c):
        c.advice[prover.C].view(-1, 4)[5, 0] ^= 1  # c on an active row

    def break_lookup(c):
        big = dv.to_device(F.mont_limbs(12345678).reshape(1, 4))  # outside the 8-bit table
        c.advice[prover.D].view(-1, 4)[7].copy_(big.view(-1, 4)[0])

    for name, fn, expect in (("copy", break_copy, {"permutation_product_closes": False, "lookup_product_closes": True}),
                             ("gate", break_gate, {"permutation_product_closes": True, "lookup_product_closes": True}),
                             ("lookup", break_lookup,
                              {"permutation_product_closes": True, "lookup_product_closes": False})):
        print("broken", name, flush=True)
        params, pk, circuit = _setup(10, 4, mutate=fn)
        out = prover.create_proof(params, pk, circuit, seed=4)
        for key, val in expect.items():
            assert out["checks"][key] == val, (name, key)
        assert not out["checks"]["h_is_a_polynomial"], name
        assert _rejects(params, pk, out["instance"], out["proof"]), name


def test_fresh_witness_same_key(lib):
    """A fresh witness (new public inputs) proves under the same proving key; the proof
    differs and verifies against its own public inputs only."""
    from halo2_amd import prover

    params, pk, circuit = _setup(10, 6)
    first = prover.create_proof(params, pk, circuit, seed=1)
    other = prover.fresh_witness(circuit, 7)
    second = prover.create_proof(params, pk, other, seed=1)
    assert all(first["checks"].values()) and all(second["checks"].values())
    assert first["proof"] != second["proof"] and first["instance"] != second["instance"]
    assert _verify(params, pk, second["instance"], second["proof"], pairing=False)
    assert _rejects(params, pk, first["instance"], second["proof"])


def test_config3_proof_k20_verifies(lib):
    """The benchmark proof: k = 20, 16-bit range table, fixed-base SRS tables. Accepted by
    the verifier (pairing and tau forms); a flipped byte is rejected."""
    from halo2_amd import prover

    params, pk, circuit = _setup(20, 5, table_bits=16, fixed_base=True)
    print("k=20 keys ready", flush=True)
    out = prover.create_proof(params, pk, circuit, seed=11)
    assert all(out["checks"].values()), out["checks"]
    print("k=20 proof", out["times_ms"], flush=True)
    assert _verify(params, pk, out["instance"], out["proof"])
    bad = bytearray(out["proof"])
    bad[len(bad) // 2] ^= 1
    assert _rejects(params, pk, out["instance"], bytes(bad))
