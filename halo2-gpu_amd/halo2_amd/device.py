"""Device-memory plumbing over torch (allocation, copies, the current stream).

Field elements travel as (n, 4) uint64 numpy arrays on the host and as int64
torch tensors (same bits) on the device; points as (n, 8).
"""
from __future__ import annotations

import threading

import numpy as np
import torch

_side = threading.local()


def side_streams(count: int) -> list:
    """Up to `count` side streams of this host thread (created once, default priority):
    for independent launch chains of one proof that are short or latency-bound."""
    pool = getattr(_side, "pool", None)
    if pool is None:
        pool = _side.pool = []
    while len(pool) < count:
        pool.append(torch.cuda.Stream())
    return pool[:count]


def device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def empty_u64(n_words: int) -> torch.Tensor:
    return torch.empty(max(1, int(n_words)), dtype=torch.int64, device=device())


def empty_bytes(n_bytes: int) -> torch.Tensor:
    """uint8 scratch rounded up to 256 bytes (hipMalloc alignment is >= 256)."""
    n = max(256, (int(n_bytes) + 255) // 256 * 256)
    return torch.empty(n, dtype=torch.uint8, device=device())


def to_device(a: np.ndarray) -> torch.Tensor:
    """Host words to a new device tensor, stream-ordered and without waiting: the words
    are staged in pinned memory (torch's caching host allocator keeps the staging block
    until the copy has run), so the host does not drain the stream as a pageable copy
    would -- the small tables and scalars of every launch sequence stay off the
    critical path."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    staged = torch.from_numpy(a.view(np.int64).reshape(-1)).pin_memory()
    return staged.to(device(), non_blocking=True)


def to_host(t: torch.Tensor, cols: int = 4) -> np.ndarray:
    return t.detach().cpu().numpy().view(np.uint64).reshape(-1, cols).copy()


def ptr(t: torch.Tensor) -> int:
    return int(t.data_ptr())


def stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def sync() -> None:
    torch.cuda.synchronize()
