"""The device proving pipeline (halo2_amd.prover, the phase order of create_proof,
src/plonk/prover.rs:141-629): on the standard-gate circuit with copy cycles and a range
lookup, the permutation and lookup grand products close (z = 1 on the last usable row)
and h's numerator is divisible by X^n - 1; a witness broken in a copied cell, a gate or
the lookup breaks exactly the matching checks. Same seed, same proof bytes."""
import pytest

pytestmark = pytest.mark.gpu


def _prove(k, seed, mutate=None, fixed_base=False):
    from halo2_amd import fields as F, msm, prover

    circuit = prover.synthesize_standard(k, seed, table_bits=8)
    if mutate:
        mutate(circuit)
    pk = prover.keygen(circuit)
    params = msm.ParamsKZG(k, 0x5EED_0F_7A0 % F.R_MOD, fixed_base=fixed_base)
    return prover.create_proof(params, pk, circuit, seed=seed)


def test_proof_pipeline_self_checks(lib):
    out = _prove(10, 3)
    assert out["checks"] == {"permutation_product_closes": True, "lookup_product_closes": True,
                             "h_is_a_polynomial": True}
    assert len(out["proof"]) > 0 and set(out["times_ms"]) >= {"advice", "quotient", "opening"}
    assert _prove(10, 3)["proof"] == out["proof"]  # deterministic under one seed
    # the fixed-base tables give the same commitments, hence the same transcript
    assert _prove(10, 3, fixed_base=True)["proof"] == out["proof"]


def test_proof_pipeline_detects_bad_witness(lib):
    from halo2_amd import device as dv, fields as F

    def break_copy(c):
        col, row = c.cycles[0][1]
        c.advice[col].view(-1, 4)[row].add_(1)

    def break_gate(c):
        c.advice[2].view(-1, 4)[5, 0] ^= 1  # c on an active row

    def break_lookup(c):
        big = dv.to_device(F.mont_limbs(12345678).reshape(1, 4))  # outside the 8-bit table
        c.advice[3].view(-1, 4)[7].copy_(big.view(-1, 4)[0])

    out = _prove(10, 4, break_copy)
    assert not out["checks"]["permutation_product_closes"] and not out["checks"]["h_is_a_polynomial"]
    out = _prove(10, 4, break_gate)
    assert out["checks"]["permutation_product_closes"] and not out["checks"]["h_is_a_polynomial"]
    out = _prove(10, 4, break_lookup)
    assert not out["checks"]["lookup_product_closes"] and not out["checks"]["h_is_a_polynomial"]


def test_fresh_witness_same_key(lib):
    """A fresh witness for the same circuit proves under the same proving key; the
    proof differs."""
    import dataclasses

    from halo2_amd import fields as F, msm, prover

    circuit = prover.synthesize_standard(10, 6, table_bits=8)
    pk = prover.keygen(circuit)
    params = msm.ParamsKZG(10, 77 % F.R_MOD)
    first = prover.create_proof(params, pk, circuit, seed=1)
    other = dataclasses.replace(circuit, advice=prover.fresh_witness(circuit, 7, 1 << 8))
    second = prover.create_proof(params, pk, other, seed=1)
    assert all(first["checks"].values()) and all(second["checks"].values())
    assert first["proof"] != second["proof"]
