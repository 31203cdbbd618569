"""A proving service on one GPU with continuous batching.

Requests name a circuit size k (one resident setup per k: SRS with its fixed-base
tables, proving key with its cached coset parts, the circuit's fixed part) and carry a
witness seed and a blinding seed. Up to `slots` proofs run at once, each on its own
host thread and HIP stream, so one proof's latency-bound phases (reduction trees,
scans, small launches, host transcript work) overlap another's VALU- and HBM-bound
kernels. A request is admitted when a slot is free and its HBM footprint fits the
budget: the waiting queue is scanned in arrival order and the first request that fits
is taken (a large request at the head does not block smaller ones behind it while its
memory is not yet free). Each admitted request draws its fresh witness on its own
stream (prover.fresh_witness) and proves it (prover.create_proof).

The footprint of one proof at k is measured once per setup: the caching allocator's
peak above its resident level during a warm-up proof (advice and instance columns,
their coefficient forms, the grand products, the extended-domain quotient, MSM
scratch). The budget is the device's memory minus what the resident setups hold and a
reserve. With slots = 1 the service is the plain one-proof-at-a-time FIFO path.

The reference has no service loop: its prover runs one proof at a time on one compute
stream (src/cuda/utils.rs:28-48, src/cuda/funcs/multiexp.rs:102-108).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import prover


@dataclass
class Setup:
    params: object
    pk: object
    circuit: object
    footprint: int = 0  # HBM bytes one in-flight proof needs beyond the resident setup


@dataclass
class Request:
    k: int
    witness_seed: int
    blind_seed: int
    arrival: float | None = None  # seconds on the service clock (submit time when None)
    start: float | None = None
    done: float | None = None
    result: dict | None = None
    error: BaseException | None = None
    event: threading.Event = field(default_factory=threading.Event)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3


def prepare_setup(k: int, seed: int, tau: int, table_bits: int = 16, fixed_base: bool = True) -> Setup:
    """One resident setup: circuit, SRS (with fixed-base tables), keys, and the measured
    footprint of one proof (from a warm-up proof, which also builds the lazy tables and
    the key's coset parts)."""
    from . import msm

    circuit = prover.synthesize_standard(k, seed, table_bits=table_bits)
    params = msm.ParamsKZG(k, tau, fixed_base=fixed_base)
    pk = prover.keygen(circuit, params)
    prover.create_proof(params, pk, circuit, seed=1)  # builds the tables and key parts
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    w = prover.fresh_witness(circuit, seed + 1)
    prover.create_proof(params, pk, w, seed=2)
    torch.cuda.synchronize()
    footprint = torch.cuda.max_memory_allocated() - base
    del w
    return Setup(params, pk, circuit, int(footprint))


class ProvingService:
    """`yield_ms`: between its phases, a proof waits (at most this long per phase) while a
    proof of a smaller circuit is running and no request is waiting, so a short proof
    that arrives during a long one is not stretched by sharing the device with it. The
    pause applies only while the service keeps up: under a backlog every proof runs
    flat out, since pausing then only delays the queue. 0 disables.

    `exclusive_k`: proofs of circuits at or above this size run one at a time (a later
    one waits while smaller requests behind it are admitted): one such proof's launches
    already fill the device, and two side by side stretch each other to twice their
    latency. None disables."""

    def __init__(self, setups: dict, slots: int = 4, hbm_budget: int | None = None, reserve: int = 4 << 30,
                 yield_ms: float = 0.0, exclusive_k: int | None = None):
        self.setups = setups
        self.slots = slots
        self.yield_s = yield_ms / 1e3
        self.exclusive_k = exclusive_k
        if hbm_budget is None:
            total = torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory
            hbm_budget = total - torch.cuda.memory_allocated() - reserve
        self.budget = int(hbm_budget)
        for k, s in setups.items():
            if s.footprint > self.budget:
                raise ValueError(f"a proof at k={k} needs {s.footprint} bytes, more than the budget {self.budget}")
        self.cv = threading.Condition()
        self._running_k: dict = {}  # id(request) -> k of the proofs on the device
        self.waiting: deque = deque()
        self.in_flight = 0
        self.in_flight_bytes = 0
        self.max_in_flight = 0
        self.closed = False
        self.t0 = time.perf_counter()
        self.device = torch.cuda.current_device()
        self.workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(slots)]
        for w in self.workers:
            w.start()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> Request:
        with self.cv:
            req.arrival = self.now() if req.arrival is None else req.arrival
            self.waiting.append(req)
            self.cv.notify_all()
        return req

    def close(self) -> None:
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()

    def _take(self) -> Request | None:
        """Under the lock: the first waiting request (arrival order) whose footprint fits
        (and, with exclusive_k, that is not a second large proof)."""
        big_running = self.exclusive_k is not None and any(k >= self.exclusive_k for k in self._running_k.values())
        for i, req in enumerate(self.waiting):
            if big_running and req.k >= self.exclusive_k:
                continue
            need = self.setups[req.k].footprint
            if self.in_flight_bytes + need <= self.budget:
                del self.waiting[i]
                self._running_k[id(req)] = req.k
                self.in_flight += 1
                self.in_flight_bytes += need
                self.max_in_flight = max(self.max_in_flight, self.in_flight)
                return req
        return None

    def _worker(self) -> None:
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream()
        while True:
            with self.cv:
                req = self._take()
                while req is None:
                    if self.closed and not self.waiting:
                        return
                    self.cv.wait()
                    req = self._take()
            s = self.setups[req.k]
            req.start = self.now()
            hook = (lambda _name, r=req: self._yield_to_smaller(r)) if self.yield_s > 0 else None
            try:
                with torch.cuda.stream(stream):
                    circuit = prover.fresh_witness(s.circuit, req.witness_seed)
                    req.result = prover.create_proof(s.params, s.pk, circuit, seed=req.blind_seed, phase_hook=hook)
                    del circuit
                stream.synchronize()
            except BaseException as e:  # reported on the request; the service keeps running
                req.error = e
            req.done = self.now()
            with self.cv:
                self._running_k.pop(id(req), None)
                self.in_flight -= 1
                self.in_flight_bytes -= s.footprint
                self.cv.notify_all()
            req.event.set()

    def _yield_to_smaller(self, req: Request) -> None:
        """Between two phases of `req`: wait while a smaller circuit's proof runs and
        nothing waits for a slot (bounded)."""
        deadline = time.perf_counter() + self.yield_s
        with self.cv:
            while not self.waiting and any(k < req.k for key, k in self._running_k.items() if key != id(req)):
                left = deadline - time.perf_counter()
                if left <= 0:
                    return
                self.cv.wait(left)


def poisson_trace(seconds: float, rate: float, ks, weights, seed: int) -> list:
    """(arrival time, k) pairs of a Poisson process of `rate` per second."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > seconds:
            return out
        out.append((t, int(rng.choice(ks, p=weights))))


def run_trace(setups: dict, trace: list, slots: int, seed: int = 0, hbm_budget: int | None = None,
              progress=None, progress_every_s: float = 30.0, yield_ms: float = 0.0,
              exclusive_k: int | None = None) -> dict:
    """Open-loop run: every request is submitted at its arrival time whatever the
    service's state; each gets a fresh witness (seed + index) and blinders. Returns
    throughput, latency percentiles (arrival to finished proof), per-k latencies and the
    peak number of proofs in flight. `progress(message)`, when given, is called about
    every `progress_every_s` seconds of the run."""
    svc = ProvingService(setups, slots=slots, hbm_budget=hbm_budget, yield_ms=yield_ms, exclusive_k=exclusive_k)
    reqs = []
    torch.cuda.synchronize()
    svc.t0 = time.perf_counter()
    last = 0.0
    for i, (ta, k) in enumerate(trace):
        wait = ta - svc.now()
        if wait > 0:
            time.sleep(wait)
        reqs.append(svc.submit(Request(k, witness_seed=seed + 7 * i + 1, blind_seed=seed + 7 * i + 3, arrival=ta)))
        if progress is not None and svc.now() - last >= progress_every_s:
            last = svc.now()
            done = sum(r.event.is_set() for r in reqs)
            progress(f"t={last:.0f}s submitted {len(reqs)} done {done} in flight {svc.in_flight}")
    for r in reqs:
        while not r.event.wait(progress_every_s if progress is not None else None):
            done = sum(x.event.is_set() for x in reqs)
            progress(f"t={svc.now():.0f}s draining: done {done} of {len(reqs)}")
    svc.close()
    errors = [r.error for r in reqs if r.error is not None]
    if errors:
        raise errors[0]
    lat = np.array([r.latency_ms for r in reqs]) if reqs else np.zeros(0)
    end = max((r.done for r in reqs), default=0.0)
    by_k = {}
    for k in sorted({r.k for r in reqs}):
        v = np.array([r.latency_ms for r in reqs if r.k == k])
        by_k[str(k)] = {"n": int(len(v)), "p50_ms": round(float(np.percentile(v, 50)), 2),
                        "p99_ms": round(float(np.percentile(v, 99)), 2)}
    return {"slots": slots, "yield_ms": yield_ms, "exclusive_k": exclusive_k, "requests": len(reqs),
            "makespan_s": round(end, 3),
            "proofs_per_s": round(len(reqs) / max(end, trace[-1][0] if trace else 1e-9), 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 2) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 2) if len(lat) else None,
            "max_in_flight": svc.max_in_flight, "by_k": by_k,
            "all_checks_hold": all(all(r.result["checks"].values()) for r in reqs),
            "results": reqs}
