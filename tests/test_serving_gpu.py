"""The continuous-batching proving service (halo2_amd.serving): every result is the
same alone and in a mixed batch (the same witness and blinding seeds give the same
proof bytes whether the proof runs by itself or beside proofs of other sizes on other
streams), also with larger proofs yielding to smaller ones and with the largest size
run one at a time, every proof verifies, and admission respects the HBM budget (a
budget for one proof serialises the service; a budget for two never has more than two
in flight)."""
import pytest

pytestmark = pytest.mark.gpu

TAU = 0x5EED_5E4


def _setups(ks):
    from halo2_amd import serving

    return {k: serving.prepare_setup(k, 40 + k, TAU + k, table_bits=8) for k in ks}


def test_same_result_alone_and_in_mix(lib):
    import plonk_verifier as pv

    from halo2_amd import serving

    setups = _setups((10, 11, 12))
    for k, s in setups.items():
        print(f"k={k} footprint {s.footprint} bytes", flush=True)
        assert s.footprint > 0
    mix = [(0.0, k) for k in (12, 10, 11, 10, 12, 11, 10, 10)]
    alone = {}
    for i, (_, k) in enumerate(mix):  # each request by itself (one slot, one request)
        r = serving.run_trace(setups, [(0.0, k)], slots=1, seed=1000 + 7 * i)["results"][0]
        alone[i] = r.result["proof"]
    # the same requests together: same seeds as above (run_trace derives them from the
    # index, so the mixed run re-creates them by index through a per-request seed base)
    for yield_ms in (0.0, 20.0):  # plain, and larger proofs pausing between phases for smaller ones
        svc = serving.ProvingService(setups, slots=4, yield_ms=yield_ms)
        reqs = [svc.submit(serving.Request(k, witness_seed=1000 + 7 * i + 1, blind_seed=1000 + 7 * i + 3,
                                           arrival=0.0)) for i, (_, k) in enumerate(mix)]
        for r in reqs:
            r.event.wait()
        svc.close()
        assert svc.max_in_flight > 1, "the mix never overlapped"
        for i, r in enumerate(reqs):
            assert r.error is None, r.error
            assert r.result["proof"] == alone[i], f"request {i} (k={r.k}) differs in the mix (yield {yield_ms})"
            s = setups[r.k]
            assert pv.verify_proof(s.pk.vk, r.result["instance"], r.result["proof"], tau=s.params.tau), i


def test_large_proofs_one_at_a_time(lib):
    """exclusive_k: proofs at or above it never overlap; smaller ones still run beside
    them, and every proof is the one it would be alone."""
    from halo2_amd import serving

    setups = _setups((10, 11))
    trace = [(0.0, 11), (0.0, 11), (0.0, 10), (0.0, 11), (0.0, 10), (0.0, 10)]
    free = serving.run_trace(setups, trace, slots=4, seed=9)
    one = serving.run_trace(setups, trace, slots=4, seed=9, yield_ms=20.0, exclusive_k=11)
    big = sorted((r for r in one["results"] if r.k == 11), key=lambda r: r.start)
    for a, b in zip(big, big[1:]):
        assert b.start >= a.done, "two k = 11 proofs ran at once"
    assert one["max_in_flight"] >= 2 and one["all_checks_hold"]
    assert [r.result["proof"] for r in one["results"]] == [r.result["proof"] for r in free["results"]]


def test_admission_respects_budget(lib):
    from halo2_amd import serving

    setups = _setups((10, 11))
    big = max(s.footprint for s in setups.values())
    trace = [(0.0, 10), (0.0, 11), (0.0, 10), (0.0, 11), (0.0, 10)]
    one = serving.run_trace(setups, trace, slots=4, seed=5, hbm_budget=big)
    assert one["max_in_flight"] == 1 and one["all_checks_hold"]
    two = serving.run_trace(setups, trace, slots=4, seed=5, hbm_budget=2 * big)
    assert 1 <= two["max_in_flight"] <= 2 and two["all_checks_hold"]
    # same seeds, same proofs whatever the admission
    assert [r.result["proof"] for r in one["results"]] == [r.result["proof"] for r in two["results"]]
    with pytest.raises(ValueError):
        serving.ProvingService(setups, slots=2, hbm_budget=big // 2)
