"""Per-size tuning of the MSM window width and the NTT work-group tile, by
measurement on the device (window width and NTT tile picked per size).

The library keeps two process-global tables (h2amd_msm_set_widths,
h2amd_ntt_set_tiles); every choice they hold gives the same results, only the
time changes. `tune()` measures the candidates around the default rule for
each size, keeps the fastest, applies the tables and returns them; `save` /
`load` keep them in a JSON file so a service can tune once per machine.
"""
from __future__ import annotations

import json
import time

import numpy as np
import torch

from . import _lib
from . import device as dv


def _median_ms(fn, reps: int) -> float:
    fn()  # warm-up (and workspace sizing under the current tables)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def set_msm_widths(widths: dict[int, int], table_widths: dict[int, int] | None = None) -> None:
    """Plain MSM widths per log2 n and precomputed-table widths per log2 table_n (tables
    already built stay valid only under the width they were built with)."""
    table = np.zeros(64, dtype=np.uint8)
    for lg, c in widths.items():
        table[lg] = c
    for lg, c in (table_widths or {}).items():
        table[32 + lg] = c
    _lib.check(_lib.load().h2amd_msm_set_widths(table.ctypes.data, 64), "h2amd_msm_set_widths")


def set_ntt_tiles(tiles: dict[int, int]) -> None:
    table = np.zeros(29, dtype=np.uint8)
    for lg, t in tiles.items():
        table[lg] = t
    _lib.check(_lib.load().h2amd_ntt_set_tiles(table.ctypes.data, 29), "h2amd_ntt_set_tiles")


def tune_msm_widths(logs=(16, 18, 20, 22, 24), spread: int = 2, reps: int = 3, seed: int = 77) -> dict:
    """For each 2^lg: time plain MSMs (device scalars and bases) at widths default +- spread;
    returns {lg: {"default": c0, "best": c, "ms": {c: ms}}}. The tables are left cleared."""
    from . import msm

    lib = _lib.load()
    set_msm_widths({})
    nmax = 1 << max(logs)
    bases = msm.g1_mul_generator(msm.random_fr(nmax, seed), nmax)
    out = {}
    for lg in logs:
        n = 1 << lg
        scal = msm.random_fr(n, seed + lg)
        c0 = int(lib.h2amd_msm_width(n))
        times = {}
        for c in range(max(4, c0 - spread), min(20, c0 + spread) + 1):
            set_msm_widths({lg: c})
            times[c] = round(_median_ms(lambda: msm.multiexp_device_scalars_device_bases(scal, bases, n), reps), 3)
        set_msm_widths({})
        out[lg] = {"default": c0, "best": min(times, key=times.get), "ms": times}
        del scal
    return out


def tune_table_widths(logs=(16, 18, 20, 22), spread: int = 3, reps: int = 5, seed: int = 79) -> dict:
    """For each 2^lg: build the fixed-base table at widths default - spread .. default and
    time the precomputed MSM over all its points; returns {lg: {"default": c0, "best": c,
    "ms": {c: ms}}}. A table serves only the width it was built with, so a chosen width
    applies to tables built after it is set. The tables are left cleared."""
    from . import msm

    lib = _lib.load()
    set_msm_widths({})
    out = {}
    for lg in logs:
        n = 1 << lg
        bases = msm.g1_mul_generator(msm.random_fr(n, seed), n)
        scal = msm.random_fr(n, seed + lg)
        c0 = int(lib.h2amd_msm_table_width(n))
        times = {}
        for c in range(max(4, c0 - spread), c0 + 1):
            set_msm_widths({}, {lg: c})
            table = msm.msm_precompute(bases, n)
            times[c] = round(_median_ms(lambda: msm.multiexp_precomputed(scal, table, n, n), reps), 3)
            del table
        set_msm_widths({})
        out[lg] = {"default": c0, "best": min(times, key=times.get), "ms": times}
        del bases, scal
    return out


def tune_ntt_tiles(logs=(16, 18, 20, 22), candidates=(8, 9, 10, 11, 12), reps: int = 5, seed: int = 78) -> dict:
    """For each 2^lg: time the CosetFFT_Part transform (device in/out) with each tile;
    returns {lg: {"best": tile (0 = default), "ms": {tile: ms}}}. The table is left cleared."""
    from . import fields as F, msm, ntt

    set_ntt_tiles({})
    out = {}
    for lg in logs:
        a = msm.random_fr(1 << lg, seed + lg)
        o = dv.empty_u64(4 << lg)
        times = {}
        for t in (0,) + tuple(candidates):
            set_ntt_tiles({lg: t} if t else {})
            times[t] = round(_median_ms(lambda: ntt.fft_normal_device(6, lg, a, o, F.omega(lg), F.ZETA), reps), 4)
        set_ntt_tiles({})
        out[lg] = {"best": min(times, key=times.get), "ms": times}
        del a, o
    return out


def tune(msm_logs=(16, 18, 20, 22, 24), ntt_logs=(16, 18, 20, 22), table_logs=(16, 18, 20, 22),
         apply: bool = True) -> dict:
    """All three tuners; with apply, the best choices become the library's tables."""
    res = {"msm": tune_msm_widths(msm_logs), "ntt": tune_ntt_tiles(ntt_logs),
           "msm_table": tune_table_widths(table_logs)}
    if apply:
        apply_tables(res)
    return res


def apply_tables(res: dict) -> None:
    set_msm_widths({int(lg): v["best"] for lg, v in res["msm"].items() if v["best"] != v["default"]},
                   {int(lg): v["best"] for lg, v in res.get("msm_table", {}).items() if v["best"] != v["default"]})
    set_ntt_tiles({int(lg): v["best"] for lg, v in res["ntt"].items() if v["best"]})


def save(res: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(res, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        res = json.load(f)
    apply_tables(res)
    return res
