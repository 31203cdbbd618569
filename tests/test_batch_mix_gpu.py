"""The uneven batch of the batching benchmark, exactly: 32 polynomials with k drawn from
{14..22} (numpy seed 2024, as `bench.py` draws them), each committed against one SRS and
extended to its coset domain (quotient degree 4), as one batched MSM pipeline and one
batched NTT launch sequence. Every batched commitment equals the closed form
(sum c_i s_i) G for the SRS bases s_i G, and every batched extension equals the
single-polynomial coeff_to_extended (itself tested against the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_uneven_batch_of_32(lib, co):
    import torch

    from halo2_amd import device as dv, msm
    from halo2_amd.domain import EvaluationDomain, coeff_to_extended_batch_device

    seed, count = 2024, 32
    ks = [int(x) for x in np.random.default_rng(seed).integers(14, 23, count)]
    print("ks", ks, flush=True)
    nmax = 1 << max(ks)
    s = msm.random_fr(nmax, seed=seed + 1)
    srs = msm.g1_mul_generator(s, nmax)
    polys = [msm.random_fr(1 << k, seed=seed + 10 + i) for i, k in enumerate(ks)]
    got = msm.multiexp_batch_device(polys, srs)
    hs = dv.to_host(s)
    g = co.g1_generator()
    for i, (k, p) in enumerate(zip(ks, polys)):
        exp = co.g1_mul(g, co.fr_inner_product(dv.to_host(p), hs[: 1 << k]))
        assert (got[i] == exp).all(), f"commitment {i} (k={k})"
    print("commitments ok", flush=True)
    doms = [EvaluationDomain(5, k) for k in ks]
    batched = coeff_to_extended_batch_device(doms, polys)
    for i, (d, p) in enumerate(zip(doms, polys)):
        one = d.coeff_to_extended_many_device([p])[0]
        assert torch.equal(one, batched[i]), f"extension {i} (k={ks[i]})"
