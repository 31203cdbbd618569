"""EvaluationDomain, mirroring the GPU methods of src/poly/domain.rs (reference)
over the C ABI. Domain constants follow halo2-axiom's EvaluationDomain::new(j, k)
(the reference re-exports that type and reads its fields: omega, omega_inv,
extended_omega(_inv), g_coset(_inv), ifft_divisor, extended_ifft_divisor,
t_evaluations, quotient_poly_degree).
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as dv
from . import ntt
from .fields import R_MOD, TWO_ADICITY, ZETA, mont_limbs, omega as root


class EvaluationDomain:
    def __init__(self, j: int, k: int):
        self.k = k
        self.n = 1 << k
        self.quotient_poly_degree = j - 1
        extended_k = k
        while (1 << extended_k) < self.n * self.quotient_poly_degree:
            extended_k += 1
        if extended_k > TWO_ADICITY:
            raise ValueError("extended domain exceeds the 2-adicity of Fr")
        self.extended_k = extended_k
        self.extended_omega = root(extended_k)
        self.omega = pow(self.extended_omega, 1 << (extended_k - k), R_MOD)
        self.omega_inv = pow(self.omega, -1, R_MOD)
        self.extended_omega_inv = pow(self.extended_omega, -1, R_MOD)
        self.g_coset = ZETA
        self.g_coset_inv = ZETA * ZETA % R_MOD
        self.ifft_divisor = pow(self.n, -1, R_MOD)
        self.extended_ifft_divisor = pow(1 << extended_k, -1, R_MOD)
        # t(X) = X^n - 1 on the coset zeta * <extended_omega>, repeats with period 2^(ext_k - k); inverted
        orig = pow(ZETA, self.n, R_MOD)
        step = pow(self.extended_omega, self.n, R_MOD)
        t = []
        cur = orig
        while True:
            t.append(cur)
            cur = cur * step % R_MOD
            if cur == orig:
                break
        self.t_evaluations = [pow((v - 1) % R_MOD, -1, R_MOD) for v in t]

    @property
    def extended_len(self) -> int:
        return 1 << self.extended_k

    # ---------------------------------------------------------------- iFFT

    def lagrange_to_coeff_many_device_inputs(self, polys: list[torch.Tensor]) -> list[torch.Tensor]:
        """domain.rs:330-358: device Lagrange -> device coefficients (batched iFFT)."""
        if not polys:
            return []
        return ntt.fft_many_to_device(ntt.IFFT, polys, self.k, self.k, self.omega_inv, self.ifft_divisor)

    def lagrange_to_coeff_device_input(self, poly: torch.Tensor) -> torch.Tensor:
        """domain.rs:387-408."""
        return self.lagrange_to_coeff_many_device_inputs([poly])[0]

    def lagrange_to_coeff_many_device(self, polys: list[np.ndarray]) -> list[torch.Tensor]:
        """domain.rs:414-434: host Lagrange -> device coefficients."""
        if not polys:
            return []
        return ntt.fft_many_to_device(ntt.IFFT, polys, self.k, self.k, self.omega_inv, self.ifft_divisor)

    # ----------------------------------------------------------- coset FFT

    def coeff_to_extended_part_many_device(self, polys, extended_omega_factor: int) -> list[torch.Tensor]:
        """domain.rs:502-566: FFT(f(g_coset * factor * X), n) for each poly, device outputs."""
        if not polys:
            return []
        return ntt.fft_many_to_device(ntt.COSET_FFT_PART, polys, self.k, self.k, self.omega,
                                      self.g_coset * extended_omega_factor % R_MOD)

    def coeff_to_extended_part_many(self, polys: list[np.ndarray], extended_omega_factor: int) -> list[np.ndarray]:
        """domain.rs:568-595 (host in/out)."""
        if not polys:
            return []
        return ntt.fft_many_host(ntt.COSET_FFT_PART, polys, self.k, self.k, self.omega,
                                 self.g_coset * extended_omega_factor % R_MOD)

    def coeff_to_extended_many(self, polys: list[np.ndarray]) -> list[np.ndarray]:
        """domain.rs:597-622 -> cosetfft_many: pad to 2^extended_k, zeta twist, DFT (host in/out)."""
        if not polys:
            return []
        return ntt.fft_many_host(ntt.COSET_FFT, polys, self.k, self.extended_k, self.extended_omega, 1)

    def coeff_to_extended_many_device(self, polys) -> list[torch.Tensor]:
        """Device-output variant of coeff_to_extended_many (same transform, no D2H)."""
        if not polys:
            return []
        return ntt.fft_many_to_device(ntt.COSET_FFT, polys, self.k, self.extended_k, self.extended_omega, 1)

    # ------------------------------------------------------ back from coset

    def extended_to_coeff_device(self, d_ext: torch.Tensor) -> torch.Tensor:
        """domain.rs:650-691: iFFT at extended_k, distribute [g_coset_inv, g_coset], truncate to
        n * quotient_poly_degree."""
        assert d_ext.numel() == 4 * self.extended_len
        ntt.fft_normal_device(ntt.IFFT, self.extended_k, d_ext, d_ext, self.extended_omega_inv,
                              self.extended_ifft_divisor)
        ntt.distribute_powers_zeta_device(d_ext, [self.g_coset_inv, self.g_coset])
        target = self.n * self.quotient_poly_degree
        if target == self.extended_len:
            return d_ext
        return d_ext[: 4 * target].clone()

    def divide_by_vanishing_poly_device(self, d_ext: torch.Tensor) -> torch.Tensor:
        """domain.rs:702-721."""
        assert d_ext.numel() == 4 * self.extended_len
        t = dv.to_device(np.stack([mont_limbs(v) for v in self.t_evaluations]))
        ntt.divide_by_vanishing_poly_device(d_ext, t)
        return d_ext

    def extended_from_lagrange_vec_device(self, parts: list[torch.Tensor]) -> torch.Tensor:
        """domain.rs:173-200: interleave the extended parts, out[i * P + p] = parts[p][i]."""
        out = dv.empty_u64(4 * self.n * len(parts))
        ntt.extended_from_lagrange_vec_device(out, parts, self.n)
        return out


def coeff_to_extended_batch_device(domains: list[EvaluationDomain], polys: list[torch.Tensor]) -> list[torch.Tensor]:
    """coeff_to_extended (domain.rs:597-613) of polynomials from domains of different sizes,
    every NTT pass of all of them in one launch: poly i (n_i device coefficients) becomes
    its domain's extended evaluations, as domains[i].coeff_to_extended_many_device would
    give one by one."""
    assert len(domains) == len(polys)
    return ntt.fft_batch_to_device(ntt.COSET_FFT, polys, [d.k for d in domains], [d.extended_k for d in domains],
                                   [d.extended_omega for d in domains], [1] * len(domains))
