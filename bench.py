#!/usr/bin/env python3
"""Benchmark of the MI355X hot path: BN254 Fr coset NTT and G1 MSM at 2^20..2^24
on one GPU, inputs resident in HBM (the "single large primitives on one GPU"
workload), with the CPU restatement of the reference path timed beside it.

Headline (`value`): coset NTT elements/s at 2^24. A step is one
`_halo2_fft_normal_to_device(CosetFFT_Part, 2^24)` call (coset factor zeta):
twiddle-table build + two LDS-staged Stockham passes, uniform seeded Fr input.
MSM throughput at 2^20, 2^22, 2^24 against a seeded-tau SRS is reported under
`extra`.

Prints ONE JSON line. `--gpus N` with N > 1 exits: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "halo2-gpu_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# HBM bytes per ntt_pass launch at 2^24 measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# scaled by the 1 GiB copy calibration (FETCH x2.0, WRITE x1.0); profiles/ntt_pass_coset_2_24.md.
NTT_PASS_TRAFFIC_2_24 = 1.101e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--msm-logs", default="20,22,24")
    ap.add_argument("--msm-reps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the CPU baseline: the job's CPU share (16 CPUs per GPU on the GPU box)")
    ap.add_argument("--no-msm", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rows", action="store_true", help="skip the per-row measurements (products, quotient, ...)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the per-size MSM width / NTT tile measurements against the default rules")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven batch (32 commit MSMs + coeff_to_extended, per poly vs batched)")
    ap.add_argument("--proof-k", type=int, default=20, help="k of the standard-gate proof measured under extra")
    ap.add_argument("--proof-reps", type=int, default=10)
    ap.add_argument("--no-proof", action="store_true", help="skip the end-to-end proof measurement")
    ap.add_argument("--proof-phase-gap-ms", type=float, default=0.0,
                    help="profiling aid: one extra, uncounted proof last that pauses this long after each phase")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the Poisson-arrival serving run (0 skips it)")
    ap.add_argument("--serve-rates", default="4,28", help="arrival rates (per second) of the serving runs")
    ap.add_argument("--serve-modes", default="fifo,continuous,continuous_yield,continuous_sized",
                    help="serving modes to run (continuous_yield: larger proofs pause between phases "
                         "while a smaller one runs; continuous_sized: that, and proofs of the largest k "
                         "one at a time)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving measurement (e.g. --serve-seconds 600 --serve-rates 4)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this size first (FETCH_SIZE/WRITE_SIZE calibration)")
    return ap.parse_args()


def progress(msg: str) -> None:
    """One progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _timed(fn, reps, torch):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def row_measurements(lib, dv, F, msm, torch, stream) -> dict:
    """Device-resident timings of the other hot-path entry points at one large size each.
    `gbs` uses the algorithmic bytes (32 B per element read or written, as the
    reference's formats hold them); every row is VALU-bound, so `gbs` is far below HBM."""
    from halo2_amd import _lib
    from halo2_amd.domain import EvaluationDomain
    from halo2_amd.quotient import Advice, Fixed, Product, Sum, custom_gates_evaluator, quotient_device

    import numpy as np

    out = {}

    def rec(name, n, dt, bytes_per_elem):
        out[name] = {"n": n, "ms": round(dt * 1e3, 4), "elems_per_s": n / dt,
                     "gbs": round(n * bytes_per_elem / dt / 1e9, 2)}

    one = dv.to_device(F.mont_limbs(1).reshape(1, 4))
    k = 24
    n = 1 << k
    d = msm.random_fr(n, seed=301)
    rec("batch_invert_2^24", n, _timed(lambda: _lib.check(lib._halo2_batch_invert(dv.ptr(d), 0, n, stream)), 5,
                                       torch), 64)
    rec("grand_product_2^24", n, _timed(lambda: _lib.check(lib._halo2_grand_product_device_inputs(
        ctypes.byref(_lib.obj(dv.ptr(d))), ctypes.byref(_lib.obj(dv.ptr(one))), n, 0, stream)), 5, torch), 64)
    del d
    k = 22
    n = 1 << k
    cols = [msm.random_fr(n, seed=310 + i) for i in range(5)]
    beta, gamma = dv.to_device(F.mont_limbs(11).reshape(1, 4)), dv.to_device(F.mont_limbs(13).reshape(1, 4))
    o = dv.empty_u64(4 * n)
    objs = [_lib.obj(dv.ptr(x)) for x in [o] + cols[:4]]
    rec("lookup_product_2^22", n, _timed(lambda: _lib.check(lib._halo2_commit_product_device_inputs(
        *[ctypes.byref(x) for x in objs], dv.ptr(beta), dv.ptr(gamma), n, 0, stream)), 5, torch), 32 * 5)
    perms = [msm.random_fr(n, seed=320 + i) for i in range(4)]
    sc = [dv.to_device(F.mont_limbs(v).reshape(1, 4)) for v in (11, 13, 17, F.omega(k), 1)]
    zmod = dv.empty_u64(4 * n)
    pws = int(lib._halo2_permutation_product_device_inputs_workspace_size(n))
    pscr = dv.empty_bytes(pws)

    def perm():
        zmod.copy_(cols[4])
        _lib.check(lib._halo2_permutation_product_device_inputs(
            ctypes.byref(_lib.obj(dv.ptr(zmod))), _lib.obj_array([dv.ptr(x) for x in perms]),
            _lib.obj_array([dv.ptr(x) for x in cols[:4]]), *[dv.ptr(x) for x in sc], n, 0, 4, dv.ptr(pscr), pws,
            stream))
    rec("permutation_product_2^22_4cols", n, _timed(perm, 5, torch), 32 * 10)
    # custom-gate quotient over one extended part: q_a a + q_b b + q_m a b + q_c c, rotations scaled by 4
    a_, b_, c_ = Advice(0), Advice(1), Advice(2)
    gate = Sum(Sum(Sum(Product(Fixed(0), a_), Product(Fixed(1), b_)), Product(Product(Fixed(2), a_), b_)),
               Product(Fixed(3), c_))
    ev = custom_gates_evaluator([gate])
    fixed = [msm.random_fr(n, seed=330 + i) for i in range(4)]
    rec("quotient_rules_standard_gate_2^22", n, _timed(
        lambda: quotient_device(ev, fixed, [], cols[:3], [], 5, n, 4), 5, torch), 32 * (7 + 1))
    # lookup and permutation quotient identities over one extended part (2 sets of 2 columns)
    qv = msm.random_fr(n, seed=335)
    lcols = [msm.random_fr(n, seed=336 + i) for i in range(7)]  # table, z, a', s', l0, l_last, l_active
    y = dv.to_device(F.mont_limbs(19).reshape(1, 4))
    rec("quotient_lookups_2^22", n, _timed(lambda: _lib.check(lib._halo2_quotient_lookups(
        dv.ptr(qv), *[dv.ptr(x) for x in lcols], dv.ptr(beta), dv.ptr(gamma), dv.ptr(y), n, stream)), 5, torch),
        32 * (7 + 2))
    zs2 = [lcols[1], lcols[2]]
    ptab = [torch.tensor([dv.ptr(x) for x in group], dtype=torch.int64, device=dv.device())
            for group in (zs2, perms, cols[:4])]
    host_sc = [np.ascontiguousarray(F.mont_limbs(v).reshape(1, 4)) for v in (11, 13, 19, 5, 11, 7, F.omega(k))]
    rec("quotient_permutation_2^22_2sets", n, _timed(lambda: _lib.check(lib._halo2_quotient_permutation(
        dv.ptr(qv), *[dv.ptr(x) for x in lcols[4:7]], *[dv.ptr(t) for t in ptab],
        *[ctypes.c_void_p(x.ctypes.data) for x in host_sc], 2, 2, 4, -1, 4, n, n, stream)), 5, torch),
        32 * (2 + 4 + 4 + 3 + 2))
    del perms, fixed, zmod, pscr, qv, lcols, zs2, ptab
    # coeff_to_extended 2^20 -> 2^22 (pad, zeta twist, NTT of the extended size)
    k = 20
    dom = EvaluationDomain(5, k)  # quotient degree 4: extended domain 2^22
    poly = msm.random_fr(1 << k, seed=340)
    rec("coeff_to_extended_2^20_to_2^22", 1 << k,
        _timed(lambda: dom.coeff_to_extended_many_device([poly]), 5, torch), 160)
    del cols, poly
    # opening phase: evaluation at a point and division by (X - u) of a 2^22-coefficient polynomial
    from halo2_amd.lookup import permute_expression_pair_device

    k = 22
    n = 1 << k
    poly = msm.random_fr(n, seed=360)
    u = dv.to_device(F.mont_limbs(0x1234567).reshape(1, 4))
    ews = int(lib._halo2_eval_polynomial_workspace_size(n))
    escr = dv.empty_bytes(ews)
    res = ctypes.c_void_p()
    pt = np.ascontiguousarray(F.mont_limbs(0x1234567).reshape(1, 4))
    rec("eval_polynomial_2^22", n, _timed(lambda: _lib.check(lib._halo2_eval_polynomial(
        dv.ptr(poly), ctypes.byref(_lib.obj(pt.ctypes.data)), ctypes.byref(res), n, dv.ptr(escr), ews, stream)),
        5, torch), 32)
    q = dv.empty_u64(4 * n)
    kws = int(lib._halo2_kate_division_workspace_size(n))
    kscr = dv.empty_bytes(kws)
    rec("kate_division_2^22", n, _timed(lambda: _lib.check(lib._halo2_kate_division_device(
        dv.ptr(poly), dv.ptr(q), dv.ptr(u), n, dv.ptr(kscr), kws, stream)), 5, torch), 64)
    # lookup permutation pair: mixed duplicates, usable rows 2^22 - 6
    alphabet = msm.random_fr(n // 3, seed=361)
    g = torch.Generator(device="cpu").manual_seed(362)
    inp = alphabet.view(-1, 4)[torch.randint(0, n // 3, (n,), generator=g).to(alphabet.device)].reshape(-1)
    tab = torch.cat([alphabet, alphabet.view(-1, 4)[torch.randint(
        0, n // 3, (n - n // 3,), generator=g).to(alphabet.device)].reshape(-1)])
    pin, ptab = dv.empty_u64(4 * n), dv.empty_u64(4 * n)
    rec("permute_expression_pair_2^22", n, _timed(
        lambda: permute_expression_pair_device(inp, tab, pin, ptab, n, n - 6), 3, torch), 128)
    del poly, q, inp, tab, pin, ptab, alphabet
    # MSM at 2^16 (the small-primitive size)
    m = 1 << 16
    pts = msm.g1_mul_generator(msm.random_fr(m, seed=350), m)
    scal = msm.random_fr(m, seed=351)
    rec("msm_2^16", m, _timed(lambda: msm.multiexp_device_scalars_device_bases(scal, pts, m), 5, torch), 96)
    return out


def batch_measurements(msm, torch, seed: int = 2024, count: int = 32) -> dict:
    """Uneven batch: `count` polynomials with k drawn from {14..22}, each with its KZG
    commit MSM (against one SRS) and its coeff_to_extended (quotient degree 4), run one
    call per polynomial and as one batched pipeline each; the results are compared."""
    import numpy as np

    from halo2_amd.domain import EvaluationDomain, coeff_to_extended_batch_device

    rng = np.random.default_rng(seed)
    ks = [int(x) for x in rng.integers(14, 23, count)]
    nmax = 1 << max(ks)
    srs = msm.g1_mul_generator(msm.random_fr(nmax, seed=seed + 1), nmax)
    polys = [msm.random_fr(1 << k, seed=seed + 10 + i) for i, k in enumerate(ks)]
    doms = [EvaluationDomain(5, k) for k in ks]

    def msm_one():
        return [msm.multiexp_device_scalars_device_bases(p, srs, p.numel() // 4) for p in polys]

    def msm_batch():
        return msm.multiexp_batch_device(polys, srs)

    def c2e_one():
        return [d.coeff_to_extended_many_device([p])[0] for d, p in zip(doms, polys)]

    def c2e_batch():
        return coeff_to_extended_batch_device(doms, polys)

    a, b = msm_one(), msm_batch()
    same_msm = all((a[i] == b[i]).all() for i in range(count))
    a, b = c2e_one(), c2e_batch()
    same_c2e = all(bool((x == y).all()) for x, y in zip(a, b))
    del a, b
    out = {"ks": ks, "elements": int(sum(1 << k for k in ks)), "same_results": bool(same_msm and same_c2e)}
    for name, f in (("msm_per_poly_ms", msm_one), ("msm_batched_ms", msm_batch), ("c2e_per_poly_ms", c2e_one),
                    ("c2e_batched_ms", c2e_batch)):
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            r = f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - ta)
            del r
        out[name] = round(min(ts) * 1e3, 2)
    out["batched_speedup"] = round((out["msm_per_poly_ms"] + out["c2e_per_poly_ms"])
                                   / (out["msm_batched_ms"] + out["c2e_batched_ms"]), 3)
    return out


def proof_measurements(msm, torch, k: int, reps: int, phase_gap_ms: float = 0.0) -> dict:
    """End-to-end proofs of the standard-gate circuit at 2^k rows (halo2_amd.prover:
    5 advice columns, copy cycles, a 16-bit range lookup, SHPLONK opening), one after
    another on one GPU with the proving key, the SRS and its fixed-base tables resident
    (built by the untimed warm-up proof): proofs/s, p50/p99 latency
    and the per-phase split of the median proof. Every proof's self-checks must hold.
    With `phase_gap_ms`, one more proof (not counted) runs last, 0.3 s after the
    others, pausing that long after each phase, so that a kernel trace shows its phases
    apart."""
    import numpy as np

    from halo2_amd import prover

    circuit = prover.synthesize_standard(k, 5)
    params = msm.ParamsKZG(k, 0x5EED)
    pk = prover.keygen(circuit, params)
    prover.create_proof(params, pk, circuit, seed=99)  # warm-up: builds g_lagrange and the fixed-base tables
    lat, runs = [], []
    ta = time.perf_counter()
    for i in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = prover.create_proof(params, pk, circuit, seed=i + 1)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
        runs.append(out)
    wall = time.perf_counter() - ta
    med = int(np.argsort(lat)[len(lat) // 2])
    out = {"k": k, "proofs": reps, "proofs_per_s": round(reps / wall, 3),
           "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
           "phases_ms": runs[med]["times_ms"], "proof_bytes": len(runs[med]["proof"]),
           "all_checks_hold": all(all(r["checks"].values()) for r in runs)}
    for streams in (2, 4):
        out[f"in_flight_{streams}"] = proofs_in_flight(prover, params, pk, circuit, torch, streams,
                                                       max(2, reps // 2))
    if phase_gap_ms > 0:
        torch.cuda.synchronize()
        time.sleep(0.3)
        prover.create_proof(params, pk, circuit, seed=reps + 1, phase_hook=lambda _: time.sleep(phase_gap_ms / 1e3))
        torch.cuda.synchronize()
    return out


def serving_measurements(torch, seconds: float, rates=(4.0, 28.0), ks=(16, 18, 20, 22),
                         weights=(0.4, 0.3, 0.2, 0.1), seed: int = 55, slots: int = 4,
                         modes=("fifo", "continuous", "continuous_yield", "continuous_sized")) -> dict:
    """The proving service (halo2_amd.serving) on open-loop Poisson traces: requests for
    the standard-gate circuit at k drawn with `weights`, each with a fresh witness, every
    k's SRS, fixed-base tables and keys resident. Each trace is served in each mode: one
    proof at a time (FIFO, the plain GPU path); continuous batching (`slots` proofs in
    flight, admitted by their measured HBM footprint); that with larger proofs pausing
    between phases while a smaller one runs (yield); and that with the largest k's
    proofs one at a time (sized). The first rate is the config's
    4/s (the device mostly idle); the second is just above the one-at-a-time capacity
    (~25/s for this mix: a mean service time of ~40 ms), so FIFO falls behind and its
    queue grows while the batched service, whose capacity is higher, keeps up."""
    from halo2_amd import serving

    setups = {k: serving.prepare_setup(k, seed + k, 0x5EED + k) for k in ks}
    out = {"seconds": seconds, "ks": list(ks), "weights": list(weights), "slots": slots,
           "footprint_mib": {str(k): round(s.footprint / 2**20, 1) for k, s in setups.items()}, "runs": {}}
    for rate in rates:
        trace = serving.poisson_trace(seconds, rate, ks, weights, seed + int(rate * 10))
        runs = {}
        for name, nslots, yield_ms, excl in (("fifo", 1, 0.0, None), ("continuous", slots, 0.0, None),
                                             ("continuous_yield", slots, 40.0, None),
                                             ("continuous_sized", slots, 40.0, max(ks))):
            if name not in modes:
                continue
            progress(f"serving rate {rate:g}/s {name}: {len(trace)} requests")
            r = serving.run_trace(setups, trace, nslots, seed=seed * 1000, progress=progress, yield_ms=yield_ms,
                                  exclusive_k=excl)
            r.pop("results")
            runs[name] = r
        out["runs"][f"rate_{rate:g}"] = runs
    del setups
    torch.cuda.empty_cache()
    return out


def proofs_in_flight(prover, params, pk, circuit, torch, streams: int, per_stream: int) -> dict:
    """Serving mode: `streams` host threads, each proving on its own HIP stream, so one
    proof's latency-bound kernels (bucket-reduction trees, scans) overlap another's
    bandwidth- and VALU-bound ones; proofs/s over the whole run and per-proof latency."""
    import threading

    import numpy as np

    def run(idx, count, sink):
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            for i in range(count):
                t = time.perf_counter()
                r = prover.create_proof(params, pk, circuit, seed=1000 + 100 * idx + i)
                st.synchronize()
                sink.append(((time.perf_counter() - t) * 1e3, all(r["checks"].values())))

    def round_(count):
        sinks = [[] for _ in range(streams)]
        th = [threading.Thread(target=run, args=(i, count, sinks[i])) for i in range(streams)]
        torch.cuda.synchronize()
        t = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        torch.cuda.synchronize()
        return time.perf_counter() - t, [v for sk in sinks for v in sk]

    round_(1)  # warm-up: each stream's allocations
    wall, res = round_(per_stream)
    lat = [v[0] for v in res]
    assert len(res) == streams * per_stream, "a proving thread failed"
    return {"streams": streams, "proofs": len(res), "proofs_per_s": round(len(res) / wall, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "all_checks_hold": all(v[1] for v in res)}


def _median_time(fn, reps: int) -> float:
    fn()  # warm-up
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(sorted(ts)[len(ts) // 2])


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_measurements(msm, dv, F, k: int, max_threads: int):
    """The C restatement of the reference CPU path (oracle/, the checker's code, timed
    here as the baseline) on the same seeded inputs: the headline 2^k coset NTT
    (distribute c^i + radix-2 FFT) on the job's CPU share, and beside it the MSM
    (multiexp_serial / best_multiexp_cpu, src/cpu/arithmetic.rs:14-156) at 2^16 and
    2^20, the 2^16 NTT, and single-thread figures. Each is the median of 5 after a
    warm-up (3 for the single-thread ones)."""
    from oracle import coracle as co

    avail = len(os.sched_getaffinity(0))
    threads = max(1, min(max_threads, avail))
    w = F.mont_limbs(F.omega(k))
    c = F.mont_limbs(F.ZETA)
    a = dv.to_host(msm.random_fr(1 << k, seed=2024))

    def coset_ntt(arr, kk, th):
        t = co.distribute_powers_inplace(arr.copy(), c, th)
        co.fft(t, kk, F.mont_limbs(F.omega(kk)), th)

    dt = _median_time(lambda: coset_ntt(a, k, threads), 5)
    cpu = {"value": (1 << k) / dt, "unit": "elems/s", "cores": threads, "kind": "port",
           "sample": f"median of 5 (after a warm-up) full coset NTTs of 2^{k} (distribute c^i + radix-2 DIT "
                     f"FFT; C restatement of the reference CPU path) on {threads} threads"}
    del a
    host = {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": avail, "threads_used": threads,
            "note": "threads = the job's CPU share (16 on the one-GPU box), not every core the host reports"}
    a16 = co.chacha_fr(31, 0, 1 << 16)
    a20 = co.chacha_fr(32, 0, 1 << 20)
    host["ntt_coset_2^16_elems_per_s"] = (1 << 16) / _median_time(lambda: coset_ntt(a16, 16, threads), 5)
    host["ntt_coset_2^20_elems_per_s"] = (1 << 20) / _median_time(lambda: coset_ntt(a20, 20, threads), 5)
    host["ntt_coset_2^20_1thread_elems_per_s"] = (1 << 20) / _median_time(lambda: coset_ntt(a20, 20, 1), 3)
    for lk in (16, 20):
        m = 1 << lk
        s_ = co.chacha_fr(40 + lk, 0, m)
        pts = co.g1_mul_gen(co.chacha_fr(50 + lk, 0, m), threads)
        host[f"msm_2^{lk}_elems_per_s"] = m / _median_time(lambda: co.msm(s_, pts, threads), 5)
        if lk == 16:
            host["msm_2^16_1thread_elems_per_s"] = m / _median_time(lambda: co.msm(s_, pts, 1), 3)
    # The k = 20 proof on the CPU, estimated from these rates (no CPU prover runs here):
    # its MSMs (5 advice, 2 permuted lookup columns, 2 permutation sets + 1 lookup
    # product, 3 quotient pieces, 2 SHPLONK commitments: 15 of 2^20 points) and its NTTs
    # (11 inverse transforms to coefficients, 4 extended parts x 11 coset transforms, the
    # 2^22 inverse of h: ~59 transforms of 2^20 elements).
    msm_s = 15 * (1 << 20) / host["msm_2^20_elems_per_s"]
    ntt_s = 59 * (1 << 20) / host["ntt_coset_2^20_elems_per_s"]
    host["proof_k20_estimate_s"] = round(msm_s + ntt_s, 2)
    host["proof_k20_estimate_note"] = ("estimate, not a run: 15 MSMs of 2^20 points and ~59 NTTs of 2^20 "
                                       "elements at the measured CPU rates; quotient, products and opening "
                                       "arithmetic left out")
    return cpu, host


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): MSM chunks, NTT batches and "
              f"quotient parts would shard across GPUs without cross-GPU kernels, which this library "
              f"does not implement", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from halo2_amd import _lib, device as dv, fields as F, msm

    lib = _lib.load()
    torch.cuda.set_device(0)
    stream = dv.stream()
    if args.serve_only:  # a long serving run on its own (its result line is not the benchmark line)
        rates = tuple(float(x) for x in args.serve_rates.split(",") if x)
        modes = tuple(x for x in args.serve_modes.split(",") if x)
        print(json.dumps({"serving": serving_measurements(torch, args.serve_seconds, rates=rates, modes=modes)}),
              flush=True)
        return 0
    k = args.log_n
    n = 1 << k

    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 20, dtype=torch.uint8, device="cuda")
        src.fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---------------------------------------------------------------- NTT
    progress(f"coset NTT 2^{k}")
    d_in = msm.random_fr(n, seed=2024)
    d_out = dv.empty_u64(4 * n)
    ws = int(lib._halo2_fft_normal_to_device_workspace_size(6, k, k))
    scratch = dv.empty_bytes(ws)
    d_omega = dv.to_device(F.mont_limbs(F.omega(k)).reshape(1, 4))
    coset = ctypes.create_string_buffer(F.mont_bytes(F.ZETA), 32)

    def ntt_step():
        e = lib._halo2_fft_normal_to_device(6, k, dv.ptr(d_in), dv.ptr(d_out), dv.ptr(d_omega), coset,
                                            dv.ptr(scratch), ws, stream)
        if e.code:
            _lib.check(e, "_halo2_fft_normal_to_device")

    for _ in range(args.warmup):
        ntt_step()
    torch.cuda.synchronize()
    lib.h2amd_profile_reset()
    lib.h2amd_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ntt_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.h2amd_profile_enable(0)
    launches, pass_ms = _lib.profile_read("ntt_pass")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3)
    avg_pass_ms = pass_ms / max(1, launches)
    passes_per_step = launches / max(1, args.steps)
    # algorithmic bytes: 64 N per transform (one compulsory 32-byte read and write per
    # element); a launch (one of the transform's passes) is credited its share of them
    transform_bytes = 64 * n
    pass_bytes = transform_bytes / passes_per_step
    achieved = pass_bytes / (avg_pass_ms * 1e-3) / 1e9

    # VALU view: field multiplications of the transform vs the measured multiply-add peak
    probe = dv.empty_bytes(2048 * 256 * 8)
    peak = ctypes.c_double(0.0)
    _lib.check(lib.h2amd_mad_peak_probe(dv.ptr(probe), 2048, 2000, ctypes.byref(peak), stream))
    mads_per_mul = 136  # 32x32->64 multiply-adds of one 8-limb Montgomery product
    ntt_muls = n * (k / 2.0 + 2 * (passes_per_step - 1) + 2)  # butterflies + inter-pass twiddles + coset twist
    valu_frac = ntt_muls * mads_per_mul / (ms_per_step * 1e-3) / peak.value

    extra = {
        "ntt_2^%d_ms" % k: round(ms_per_step, 4),
        "ntt_pass_avg_ms": round(avg_pass_ms, 4),
        "ntt_pass_launches_per_step": passes_per_step,
        "ntt_whole_step_hbm_frac": round(transform_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        # each pass streams the whole array once: 64 N bytes moved per launch
        "ntt_pass_bytes_moved_per_launch": 64 * n,
        "ntt_pass_moved_gbs": round(64 * n / (avg_pass_ms * 1e-3) / 1e9, 1),
        "mad_u64_u32_peak_per_s": peak.value,
        "ntt_valu_frac_of_mad_peak": round(valu_frac, 4),
    }
    del d_in, d_out, scratch

    # smaller NTTs for the size sweep
    for kk in (20, 22):
        if kk >= k:
            continue
        nn = 1 << kk
        a = msm.random_fr(nn, seed=7)
        o = dv.empty_u64(4 * nn)
        w2 = int(lib._halo2_fft_normal_to_device_workspace_size(6, kk, kk))
        s2 = dv.empty_bytes(w2)
        om = dv.to_device(F.mont_limbs(F.omega(kk)).reshape(1, 4))
        for _ in range(3):
            _lib.check(lib._halo2_fft_normal_to_device(6, kk, dv.ptr(a), dv.ptr(o), dv.ptr(om), coset, dv.ptr(s2),
                                                       w2, stream))
        torch.cuda.synchronize()
        reps = 20
        ta = time.perf_counter()
        for _ in range(reps):
            _lib.check(lib._halo2_fft_normal_to_device(6, kk, dv.ptr(a), dv.ptr(o), dv.ptr(om), coset, dv.ptr(s2),
                                                       w2, stream))
        torch.cuda.synchronize()
        extra["ntt_2^%d_elems_per_s" % kk] = nn / ((time.perf_counter() - ta) / reps)

    # ---------------------------------------------------------------- MSM
    progress("MSM")
    if not args.no_msm:
        logs = [int(x) for x in args.msm_logs.split(",") if x]
        kmax = max(logs)
        tau = 0x5EED_0F_7A0_2024 % F.R_MOD
        srs = msm.g1_mul_generator(msm.omega_powers(tau, kmax), 1 << kmax)  # G_i = tau^i G
        for lk in logs:
            m = 1 << lk
            sc = msm.random_fr(m, seed=100 + lk)
            msm.multiexp_device_scalars_device_bases(sc, srs[: 8 * m], m)  # warm-up
            torch.cuda.synchronize()
            lib.h2amd_profile_reset()
            lib.h2amd_profile_enable(1)
            ta = time.perf_counter()
            for _ in range(args.msm_reps):
                msm.multiexp_device_scalars_device_bases(sc, srs[: 8 * m], m)
            dt = (time.perf_counter() - ta) / args.msm_reps
            lib.h2amd_profile_enable(0)
            c_ref = int(math.log2(m)) // 2 + 2  # the reference's window (multiexp.cu:70)
            mads = m * math.ceil(254 / c_ref) * 10 * 136
            extra["msm_2^%d_ms" % lk] = round(dt * 1e3, 3)
            extra["msm_2^%d_elems_per_s" % lk] = m / dt
            extra["msm_2^%d_valu_frac_of_mad_peak" % lk] = round(mads / dt / peak.value, 4)
            for name in ("msm_sort", "msm_buckets", "msm_reduce"):
                cnt, tot = _lib.profile_read(name)
                if cnt:
                    extra["msm_2^%d_%s_ms" % (lk, name)] = round(tot / cnt, 3)
            # the same MSM through the fixed-base table of the resident SRS (built once per SRS)
            torch.cuda.synchronize()
            ta = time.perf_counter()
            table = msm.msm_precompute(srs[: 8 * m], m)
            torch.cuda.synchronize()
            extra["msm_fixed_base_table_2^%d_ms" % lk] = round((time.perf_counter() - ta) * 1e3, 1)
            msm.multiexp_precomputed(sc, table, m, m)  # warm-up
            ta = time.perf_counter()
            for _ in range(args.msm_reps):
                msm.multiexp_precomputed(sc, table, m, m)
            dt = (time.perf_counter() - ta) / args.msm_reps
            extra["msm_fixed_base_2^%d_ms" % lk] = round(dt * 1e3, 3)
            extra["msm_fixed_base_2^%d_elems_per_s" % lk] = m / dt
            del sc, table
        del srs

    # ----------------------------------------- the other hot-path rows
    progress("rows")
    if not args.no_rows:
        extra["rows"] = row_measurements(lib, dv, F, msm, torch, stream)

    # ------------------------------- uneven batch (32 polys, k in 14..22)
    progress("uneven batch")
    if not args.no_batch:
        extra["batch_uneven_32"] = batch_measurements(msm, torch)

    # ------------------------- end-to-end proof of the standard-gate circuit
    progress("proofs")
    if not args.no_proof:
        extra[f"proof_k{args.proof_k}"] = proof_measurements(msm, torch, args.proof_k, args.proof_reps,
                                                             args.proof_phase_gap_ms)

    # ---------------- proving service: Poisson arrivals over k in {16..22}
    progress("serving")
    if not args.no_proof and args.serve_seconds > 0:
        extra["serving"] = serving_measurements(torch, args.serve_seconds,
                                                rates=tuple(float(x) for x in args.serve_rates.split(",") if x))

    # ------------------ tuner: MSM width / NTT tile per size vs the default rules
    progress("tuner")
    if not args.no_autotune:
        from halo2_amd import autotune

        tuned = autotune.tune(msm_logs=(16, 20, 22, 24), ntt_logs=(16, 18, 20, 22), table_logs=(16, 20, 22),
                              apply=False)
        extra["autotune"] = {
            key: {f"2^{lg}": {"default_c": v["default"], "best_c": v["best"],
                              "default_ms": v["ms"][v["default"]], "best_ms": v["ms"][v["best"]]}
                  for lg, v in tuned[key].items()} for key in ("msm", "msm_table")}
        extra["autotune"].update({
            "ntt": {f"2^{lg}": {"best_tile_log2": v["best"], "default_ms": v["ms"][0],
                                "best_ms": v["ms"][v["best"]]} for lg, v in tuned["ntt"].items()},
        })

    # ------------------------------------------------------ CPU baseline
    progress("CPU baseline")
    cpu = None
    if not args.no_cpu:
        cpu, extra["cpu_baselines"] = cpu_measurements(msm, dv, F, k, args.cpu_threads)

    line = {
        "metric": "MSM/NTT elems/s at 2^16–2^24; proofs/s + p99 latency at k=20 vs plain GPU",
        "value": value,
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (BN254 Fr: 32-byte Montgomery words in HBM, 9x29-bit limbs in registers)",
        "data": "synthetic (ChaCha20-seeded uniform Fr generated on device; SRS tau^i G from a seeded tau)",
        "config": {"workload": f"coset NTT (CosetFFT_Part, c = zeta) of 2^{k} Fr, device-resident",
                   "log_n": k, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": NTT_PASS_TRAFFIC_2_24,
                     "kernel": "ntt_pass", "algorithmic_bytes_per_launch": pass_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
