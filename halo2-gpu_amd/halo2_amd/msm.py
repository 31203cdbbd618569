"""MSM wrappers over the C ABI, mirroring src/cuda/funcs/multiexp.rs and
src/arithmetic.rs::best_multiexp of the reference, plus the KZG parameter
object of src/poly/kzg/commitment.rs that keeps the SRS resident in HBM.

Results are the reference's 96-byte projective-out layout ((n,) uint64[12]):
(x, y, 1) or (0, 1, 0) in Montgomery form.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _lib, g1
from . import device as dv
from .fields import R_MOD, mont_limbs


def _out_obj():
    out = np.zeros(12, dtype=np.uint64)
    return out, _lib.obj(out.ctypes.data)


def multiexp_device_scalars_device_bases(d_scalars: torch.Tensor, d_bases: torch.Tensor, n: int,
                                         stream: int | None = None) -> np.ndarray:
    """`_halo2_multiexp_device_scalars_device_bases` (multiexp.rs:215-256): one chunk."""
    lib = _lib.load()
    ws = int(lib._halo2_multiexp_device_scalars_device_bases_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    out, obj = _out_obj()
    _lib.check(lib._halo2_multiexp_device_scalars_device_bases(dv.ptr(d_scalars), dv.ptr(d_bases),
                                                               ctypes.byref(obj), n, dv.ptr(scratch), ws,
                                                               stream if stream is not None else dv.stream()),
               "_halo2_multiexp_device_scalars_device_bases")
    return out


def multiexp_device_bases(h_scalars: np.ndarray, d_bases: torch.Tensor) -> np.ndarray:
    """`_halo2_multiexp_device_bases` (multiexp.rs:122-170): host scalars, resident bases."""
    lib = _lib.load()
    s = np.ascontiguousarray(h_scalars, dtype=np.uint64).reshape(-1, 4)
    n = len(s)
    ws = int(lib._halo2_multiexp_device_bases_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    out, obj = _out_obj()
    sobj = _lib.obj(s.ctypes.data)
    _lib.check(lib._halo2_multiexp_device_bases(ctypes.byref(sobj), dv.ptr(d_bases), ctypes.byref(obj), n,
                                                dv.ptr(scratch), ws, dv.stream()), "_halo2_multiexp_device_bases")
    return out


def multiexp(h_scalars: np.ndarray, h_points: np.ndarray) -> np.ndarray:
    """`_halo2_multiexp` (multiexp.rs:35-108 single chunk): host scalars and bases."""
    lib = _lib.load()
    s = np.ascontiguousarray(h_scalars, dtype=np.uint64).reshape(-1, 4)
    pts = np.ascontiguousarray(h_points, dtype=np.uint64).reshape(-1, 8)
    assert len(s) == len(pts)
    n = len(s)
    ws = int(lib._halo2_multiexp_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    out, obj = _out_obj()
    so, po = _lib.obj(s.ctypes.data), _lib.obj(pts.ctypes.data)
    _lib.check(lib._halo2_multiexp(ctypes.byref(so), ctypes.byref(po), ctypes.byref(obj), n, dv.ptr(scratch), ws,
                                   dv.stream()), "_halo2_multiexp")
    return out


def multiexp_gpu_device_scalars_device_bases_chunked(d_scalars: torch.Tensor, d_bases: torch.Tensor, n: int,
                                                     chunk_size: int) -> np.ndarray:
    """multiexp.rs:258-308: split into chunks, fold the per-chunk points on the host."""
    acc = None
    for off in range(0, n, chunk_size):
        m = min(chunk_size, n - off)
        part = multiexp_device_scalars_device_bases(d_scalars[4 * off: 4 * (off + m)],
                                                    d_bases[8 * off: 8 * (off + m)], m)
        acc = g1.add(acc, g1.from_out12(part))
    return g1.to_out12(acc)


def best_multiexp(h_scalars: np.ndarray, h_points: np.ndarray) -> np.ndarray:
    """src/arithmetic.rs:73-81. The reference switches to its CPU path below 2^14
    points; here every size runs on the GPU (the result is the same point), with the
    free-memory chunking of multiexp.rs:47-99."""
    lib = _lib.load()
    s = np.ascontiguousarray(h_scalars, dtype=np.uint64).reshape(-1, 4)
    pts = np.ascontiguousarray(h_points, dtype=np.uint64).reshape(-1, 8)
    n = len(s)
    free, _ = torch.cuda.mem_get_info()
    max_len = int(lib._halo2_msm_max_length(max(0, free - (256 << 20))))
    if max_len == 0:
        raise _lib.Halo2GpuError(2, "not enough device memory for any MSM chunk")
    if n <= max_len:
        return multiexp(s, pts)
    acc = None
    for off in range(0, n, max_len):
        acc = g1.add(acc, g1.from_out12(multiexp(s[off: off + max_len], pts[off: off + max_len])))
    return g1.to_out12(acc)


def msm_precompute(d_bases: torch.Tensor, n: int) -> torch.Tensor:
    """Fixed-base table of n device bases (`h2amd_msm_precompute`): W rows of 2^(c w) P_i."""
    lib = _lib.load()
    table = torch.empty(max(1, int(lib.h2amd_msm_table_bytes(n)) // 8), dtype=torch.int64, device=d_bases.device)
    ws = int(lib.h2amd_msm_precompute_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib.h2amd_msm_precompute(dv.ptr(table), dv.ptr(d_bases), n, dv.ptr(scratch), ws, dv.stream()),
               "h2amd_msm_precompute")
    return table


def multiexp_precomputed(d_scalars: torch.Tensor, table: torch.Tensor, table_n: int, n: int) -> np.ndarray:
    """Fixed-base MSM over the first n <= table_n points (`h2amd_multiexp_precomputed`)."""
    lib = _lib.load()
    ws = int(lib.h2amd_multiexp_precomputed_workspace_size(n, table_n))
    scratch = dv.empty_bytes(ws)
    out, obj = _out_obj()
    _lib.check(lib.h2amd_multiexp_precomputed(dv.ptr(d_scalars), dv.ptr(table), table_n, ctypes.byref(obj), n,
                                              dv.ptr(scratch), ws, dv.stream()), "h2amd_multiexp_precomputed")
    return out


def multiexp_batch_device(d_scalars: list[torch.Tensor], d_bases: torch.Tensor,
                          lengths: list[int] | None = None) -> np.ndarray:
    """`h2amd_multiexp_batch`: MSM j of the first lengths[j] (default: all) scalars of
    d_scalars[j] against the shared bases, all in one pipeline; (count, 12) uint64."""
    lib = _lib.load()
    count = len(d_scalars)
    if lengths is None:
        lengths = [x.numel() // 4 for x in d_scalars]
    lens = np.ascontiguousarray(lengths, dtype=np.uint64)
    ptrs = (ctypes.c_void_p * max(1, count))(*[dv.ptr(x) for x in d_scalars])
    ws = int(lib.h2amd_multiexp_batch_workspace_size(lens.ctypes.data, count))
    scratch = dv.empty_bytes(ws)
    out = np.zeros((max(1, count), 12), dtype=np.uint64)
    _lib.check(lib.h2amd_multiexp_batch(ptrs, lens.ctypes.data, count, dv.ptr(d_bases), out.ctypes.data,
                                        dv.ptr(scratch), ws, dv.stream()), "h2amd_multiexp_batch")
    return out[:count]


def multiexp_batch_precomputed(d_scalars: list[torch.Tensor], table: torch.Tensor, table_n: int,
                               lengths: list[int] | None = None) -> np.ndarray:
    """`h2amd_multiexp_batch_precomputed`: the batch of multiexp_batch_device over a
    fixed-base table of table_n bases (msm_precompute); (count, 12) uint64."""
    lib = _lib.load()
    count = len(d_scalars)
    if lengths is None:
        lengths = [x.numel() // 4 for x in d_scalars]
    lens = np.ascontiguousarray(lengths, dtype=np.uint64)
    ptrs = (ctypes.c_void_p * max(1, count))(*[dv.ptr(x) for x in d_scalars])
    ws = int(lib.h2amd_multiexp_batch_precomputed_workspace_size(lens.ctypes.data, count, table_n))
    scratch = dv.empty_bytes(ws)
    out = np.zeros((max(1, count), 12), dtype=np.uint64)
    _lib.check(lib.h2amd_multiexp_batch_precomputed(ptrs, lens.ctypes.data, count, dv.ptr(table), table_n,
                                                    out.ctypes.data, dv.ptr(scratch), ws, dv.stream()),
               "h2amd_multiexp_batch_precomputed")
    return out[:count]


def g_to_lagrange_device(d_g: torch.Tensor, k: int, omega: int | None = None) -> torch.Tensor:
    """`h2amd_g1_to_lagrange` (halo2-axiom g_to_lagrange, used by ParamsKZG::from_parts and
    downsize, commitment.rs:205-228, :436-450): the Lagrange bases n^-1 sum_j w^-ij g_j of
    2^k device affine points, w the 2^k-th root of unity (default: the domain's)."""
    from .fields import inv, omega as root

    lib = _lib.load()
    n = 1 << k
    assert d_g.numel() >= 8 * n
    w_inv = inv(root(k) if omega is None else omega)
    wb = np.ascontiguousarray(mont_limbs(w_inv))
    out = dv.empty_u64(8 * n)
    ws = int(lib.h2amd_g1_to_lagrange_workspace_size(k))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib.h2amd_g1_to_lagrange(dv.ptr(out), dv.ptr(d_g), k, wb.ctypes.data, dv.ptr(scratch), ws,
                                        dv.stream()), "h2amd_g1_to_lagrange")
    return out


def g1_mul_generator(d_scalars: torch.Tensor, n: int) -> torch.Tensor:
    """d_out[i] = s_i * G (affine, 8 words per point) on the device."""
    lib = _lib.load()
    out = dv.empty_u64(8 * n)
    ws = int(lib.h2amd_g1_mul_generator_workspace_size(n))
    scratch = dv.empty_bytes(ws)
    _lib.check(lib.h2amd_g1_mul_generator(dv.ptr(out), dv.ptr(d_scalars), n, dv.ptr(scratch), ws, dv.stream()),
               "h2amd_g1_mul_generator")
    return out


def random_fr(n: int, seed: int, start: int = 0) -> torch.Tensor:
    lib = _lib.load()
    out = dv.empty_u64(4 * n)
    _lib.check(lib.h2amd_random_fr(dv.ptr(out), n, seed, start, dv.stream()), "h2amd_random_fr")
    return out


def omega_powers(base: int, log_n: int) -> torch.Tensor:
    """d[i] = base^i for i < 2^log_n (`_halo2_generate_omega_powers`)."""
    lib = _lib.load()
    out = dv.empty_u64(4 << log_n)
    b = dv.to_device(mont_limbs(base % R_MOD).reshape(1, 4))
    _lib.check(lib._halo2_generate_omega_powers(dv.ptr(out), dv.ptr(b), log_n, None, 0, dv.stream()),
               "_halo2_generate_omega_powers")
    return out


class ParamsKZG:
    """src/poly/kzg/commitment.rs:27-582: KZG parameters with the G1 SRS kept resident
    in HBM. Built from a toxic-waste tau (`setup`, :118-197): g[i] = tau^i G and the
    Lagrange basis g_lagrange[i] = L_i(tau) G = G * ((tau^n - 1)/n) * w^i / (tau - w^i),
    both generated on the device. The lazy device mirrors of the reference
    (`g_device`, `g_lagrange_device`, :246-304) are the resident buffers themselves.
    g2 and s_g2 = tau G2 (affine over Fq2, Python integers) are computed on the host
    when first asked for; `from_parts` builds the object from stored points (an SRS
    file, halo2_amd.srs)."""

    def __init__(self, k: int, tau: int, omega: int | None = None, fixed_base: bool = True):
        from .fields import omega as root

        self.k = k
        self.n = 1 << k
        self.tau = tau % R_MOD
        self.omega = root(k) if omega is None else omega % R_MOD
        self.g_device = g1_mul_generator(omega_powers(self.tau, k), self.n)
        self._g_lagrange_device = None
        # fixed-base tables (rows 2^(c w) of each basis), built on first commit
        self.fixed_base = fixed_base
        self._tables: dict[str, torch.Tensor] = {}
        self._table_width: dict[str, int] = {}
        self._g2 = self._s_g2 = None
        self._lock = threading.RLock()

    @classmethod
    def from_parts(cls, k: int, g_device: torch.Tensor, g_lagrange_device: torch.Tensor | None, g2, s_g2,
                   fixed_base: bool = True) -> "ParamsKZG":
        """commitment.rs:205-228: from device affine bases (n each); without g_lagrange it is
        computed from g (g_to_lagrange), as the reference does."""
        from .fields import omega as root

        assert g_device.numel() == 8 << k
        if g_lagrange_device is None:  # from_parts without g_lagrange: g_to_lagrange
            g_lagrange_device = g_to_lagrange_device(g_device, k)
        assert g_lagrange_device.numel() == 8 << k
        self = cls.__new__(cls)
        self.k, self.n = k, 1 << k
        self.tau = None
        self.omega = root(k)
        self.g_device = g_device
        self._g_lagrange_device = g_lagrange_device
        self.fixed_base = fixed_base
        self._tables = {}
        self._table_width = {}
        self._g2, self._s_g2 = g2, s_g2
        self._lock = threading.RLock()
        return self

    def downsize(self, k: int) -> None:
        """commitment.rs:436-450: keep the first 2^k bases of g and rebuild g_lagrange for
        the smaller domain (g_to_lagrange of the truncated g); cached tables are dropped."""
        from .fields import omega as root

        assert k <= self.k
        self.k, self.n = k, 1 << k
        self.omega = root(k)
        self.g_device = self.g_device[: 8 * self.n].clone()
        self._g_lagrange_device = g_to_lagrange_device(self.g_device, k)
        self._tables = {}
        self._table_width = {}

    @property
    def g2(self):
        """commitment.rs:306-309: the G2 generator."""
        if self._g2 is None and self.tau is not None:
            from .srs import G2_GENERATOR

            self._g2 = G2_GENERATOR
        return self._g2

    @property
    def s_g2(self):
        """commitment.rs:311-314: tau G2 (setup, :191-192)."""
        if self._s_g2 is None and self.tau is not None:
            from .srs import g2_mul

            self._s_g2 = g2_mul(self.g2, self.tau)
        return self._s_g2

    def _commit_resident(self, d_scalars: torch.Tensor, which: str) -> np.ndarray:
        n = d_scalars.numel() // 4
        assert n <= self.n
        bases = self.g_device if which == "g" else self.g_lagrange_device
        if not self.fixed_base:
            return multiexp_device_scalars_device_bases(d_scalars, bases[: 8 * n], n)
        return multiexp_precomputed(d_scalars, self._table(which), self.n, n)

    def _table(self, which: str) -> torch.Tensor:
        """The fixed-base table of g or g_lagrange, built on first use under the current
        table width; a width changed since (h2amd_msm_set_widths) is refused."""
        width = int(_lib.load().h2amd_msm_table_width(self.n))
        if which not in self._tables:
            with self._lock:  # concurrent first commits: one builder, readers see a finished table
                if which not in self._tables:
                    bases = self.g_device if which == "g" else self.g_lagrange_device
                    table = msm_precompute(bases, self.n)
                    # other threads launch on other streams: publish only once the table is written
                    torch.cuda.current_stream().synchronize()
                    self._table_width[which] = width
                    self._tables[which] = table
        if self._table_width[which] != width:
            raise RuntimeError(f"the {which} table was built with window width {self._table_width[which]}, "
                               f"the library now uses {width}: rebuild the parameters")
        return self._tables[which]

    @property
    def g_lagrange_device(self) -> torch.Tensor:
        """commitment.rs:157-183 (scalars) and :223-241 (device mirror, built on first use)."""
        if self._g_lagrange_device is not None:
            return self._g_lagrange_device
        with self._lock:
            if self._g_lagrange_device is None:
                self._g_lagrange_device = self._build_g_lagrange()
        return self._g_lagrange_device

    def _build_g_lagrange(self) -> torch.Tensor:
        """L_i(tau) G for every i, on the current stream, synchronised before it is
        published (other threads read it from other streams)."""
        from . import polyops

        n = self.n
        roots = omega_powers(self.omega, self.k)
        diff = dv.empty_u64(4 * n)
        polyops.poly_fill_one_device(diff)  # diff = tau - w^i
        polyops.poly_scale_device_with_d_s_minus_one(diff, _scalar_dev(self.tau - 1))
        polyops.poly_sub_short_out_of_place_device(diff, diff, roots)
        _lib.check(_lib.load()._halo2_batch_invert(dv.ptr(diff), 0, n, dv.stream()), "_halo2_batch_invert")
        polyops.poly_elementwise_multiply_device(diff, diff, roots)
        mult = (pow(self.tau, n, R_MOD) - 1) * pow(n, -1, R_MOD) % R_MOD
        polyops.poly_scale_device_with_d_s_minus_one(diff, _scalar_dev(mult - 1))
        out = g1_mul_generator(diff, n)
        torch.cuda.current_stream().synchronize()
        return out

    def g0_out12(self) -> np.ndarray:
        """g[0] (the generator G) in the 96-byte projective-out layout, read once."""
        if getattr(self, "_g0", None) is None:
            from .fields import P_MOD

            out = np.zeros(12, dtype=np.uint64)
            out[:8] = dv.to_host(self.g_device[:8], cols=8)[0]
            out[8:12] = mont_limbs(1, P_MOD)
            self._g0 = out
        return self._g0

    def commit_device(self, d_poly: torch.Tensor) -> np.ndarray:
        """commitment.rs:560-577: MSM of device coefficients against the resident SRS
        (through its fixed-base table unless fixed_base is off)."""
        return self._commit_resident(d_poly, "g")

    def commit(self, h_poly: np.ndarray) -> np.ndarray:
        """commitment.rs:531-541: host coefficients, resident SRS."""
        h = np.ascontiguousarray(h_poly, dtype=np.uint64).reshape(-1, 4)
        if self.fixed_base:
            return self._commit_resident(dv.to_device(h), "g")
        return multiexp_device_bases(h, self.g_device[: 8 * len(h)])

    def commit_lagrange_device(self, d_values: torch.Tensor) -> np.ndarray:
        """commitment.rs:467-485: MSM of device Lagrange values against the resident
        Lagrange basis."""
        return self._commit_resident(d_values, "g_lagrange")

    def _commit_many(self, d_polys: list[torch.Tensor], which: str) -> np.ndarray:
        bases = self.g_device if which == "g" else self.g_lagrange_device
        outs = []
        for i in range(0, len(d_polys), 64):
            part = d_polys[i:i + 64]
            outs.append(multiexp_batch_precomputed(part, self._table(which), self.n) if self.fixed_base
                        else multiexp_batch_device(part, bases))
        return np.concatenate(outs) if outs else np.zeros((0, 12), dtype=np.uint64)

    def commit_many_device(self, d_polys: list[torch.Tensor]) -> np.ndarray:
        """KZG commitments of many device polynomials (any lengths <= n) against the
        resident SRS in one batched MSM pipeline (over the fixed-base table unless
        fixed_base is off), at most 64 per launch sequence; the same points
        commit_device gives one by one."""
        return self._commit_many(d_polys, "g")

    def commit_lagrange_many_device(self, d_values: list[torch.Tensor]) -> np.ndarray:
        """commit_lagrange of many device Lagrange vectors in one batched pipeline."""
        return self._commit_many(d_values, "g_lagrange")

    def commit_lagrange(self, h_values: np.ndarray) -> np.ndarray:
        """commitment.rs:456-466: host Lagrange values, resident Lagrange basis."""
        h = np.ascontiguousarray(h_values, dtype=np.uint64).reshape(-1, 4)
        if self.fixed_base:
            return self._commit_resident(dv.to_device(h), "g_lagrange")
        return multiexp_device_bases(h, self.g_lagrange_device[: 8 * len(h)])


def _scalar_dev(x: int) -> torch.Tensor:
    return dv.to_device(mont_limbs(x % R_MOD).reshape(1, 4))
