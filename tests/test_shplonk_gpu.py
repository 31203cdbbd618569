"""SHPLONK multi-opening on the device (src/poly/kzg/multiopen/shplonk/prover.rs:97-441
of the reference, challenges given) against the same steps composed from the oracle's
evaluation, kate division and field operations, plus properties that pin the wiring
independently: l_x(u) = 0 (the reference's own debug check), every rotation set's
n_x vanishes on its points, and the commitments equal f(tau) G for the SRS tau^i G."""
import random

import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def test_shplonk_create_proof(lib, co):
    from halo2_amd import device as dv, fields as F, msm
    from halo2_amd.shplonk import ProverQuery, create_proof_device, lagrange_interpolate

    R = F.R_MOD
    M = F.mont_limbs
    k, tau = 10, 0x5EED7A0
    n = 1 << k
    params = msm.ParamsKZG(k, tau)
    polys = [rand_fr(co, n, 2000 + i) for i in range(5)]
    d_polys = [dv.to_device(p) for p in polys]
    rng = random.Random(5)
    x = rng.randrange(R)
    w = F.omega(k)
    xw, xwi = x * w % R, x * pow(w, -1, R) % R
    plan = [(0, [x]), (1, [x, xw]), (2, [xw, x]), (3, [xwi, x, xw]), (4, [x])]
    queries = [ProverQuery(p, d_polys[i], i) for i, pts in plan for p in pts]
    y, v, u = (rng.randrange(R) for _ in range(3))
    out = create_proof_device(params, queries, y, v, u)
    # u squeezed after h is written (a function of h's commitment) gives the same proof
    seen = []
    again = create_proof_device(params, queries, y, v, lambda h: seen.append(h.copy()) or u)
    assert len(seen) == 1 and (seen[0] == out["h"]).all() and (again["h_final"] == out["h_final"]).all()

    # the same steps on the oracle
    val = lambda row: F.value_of(row)  # noqa: E731
    ev = {(i, p): val(co.eval_poly(polys[i], M(p))) for i, pts in plan for p in pts}
    sets = []  # (points, [poly index])
    for i, pts in plan:
        for s in sets:
            if set(s[0]) == set(pts):
                s[1].append(i)
                break
        else:
            sets.append((list(pts), [i]))
    super_pts = [x, xw, xwi]
    tile = lambda c: np.tile(M(c), (n, 1))  # noqa: E731
    zero = np.zeros((n, 4), dtype=np.uint64)
    qs, pxs, zs = [], [], []
    for pts, idx in sets:
        p_x, r_x = zero.copy(), [0] * len(pts)
        for j, i in enumerate(idx):
            yj = pow(y, j, R)
            p_x = co.field_add(0, p_x, co.field_mul(0, polys[i], tile(yj)))
            for d, c in enumerate(lagrange_interpolate(pts, [ev[(i, p)] for p in pts])):
                r_x[d] = (r_x[d] + c * yj) % R
        n_x = p_x.copy()
        n_x[:len(r_x)] = co.field_sub(0, p_x[:len(r_x)], np.stack([M(c) for c in r_x]))
        for p in pts:  # exact division: n_x vanishes on the set
            assert val(co.eval_poly(n_x, M(p))) == 0
        q = n_x
        for p in pts:
            q = co.kate_division(q, M(p))
        qs.append(np.concatenate([q, np.zeros((n - len(q), 4), dtype=np.uint64)]))
        z_i = 1
        for p in super_pts:
            if p not in pts:
                z_i = z_i * (u - p) % R
        r_i = 0
        for j, i in enumerate(idx):
            r_i = (r_i + val(co.eval_poly(np.stack([M(c) for c in lagrange_interpolate(
                pts, [ev[(i, p)] for p in pts])]), M(u))) * pow(y, j, R)) % R
        l_x = p_x.copy()
        l_x[0] = co.field_sub(0, p_x[:1], M(r_i).reshape(1, 4))[0]
        pxs.append(co.field_mul(0, l_x, tile(z_i)))
        zs.append(z_i)
    h_x = zero.copy()
    l_tot = zero.copy()
    for j, (q, lx) in enumerate(zip(qs, pxs)):
        h_x = co.field_add(0, h_x, co.field_mul(0, q, tile(pow(v, j, R))))
        l_tot = co.field_add(0, l_tot, co.field_mul(0, lx, tile(pow(v, j, R))))
    zt = 1
    for p in super_pts:
        zt = zt * (u - p) % R
    l_x = co.field_sub(0, l_tot, co.field_mul(0, h_x, tile(zt)))
    assert val(co.eval_poly(l_x, M(u))) == 0
    h_final = co.field_mul(0, co.kate_division(l_x, M(u)), np.tile(M(pow(zs[0], -1, R)), (n - 1, 1)))

    assert (dv.to_host(out["h_x"]) == h_x).all()
    assert (dv.to_host(out["l_x"]) == l_x).all()
    assert (dv.to_host(out["h_x_final"]) == h_final).all()
    for name, poly in (("h", h_x), ("h_final", h_final)):
        exp = co.g1_mul_gen(co.eval_poly(poly, M(tau)).reshape(1, 4))[0]
        assert (out[name][:8] == exp).all(), name
