"""MI355X (gfx950) hot path of the halo2-axiom-gpu prover.

Host-side mirror of the reference's wrappers (src/cuda/funcs/*.rs,
src/poly/domain.rs, src/arithmetic.rs, src/poly/kzg/commitment.rs) over the C ABI
declared in include/halo2_gpu.h and implemented by libhalo2_gpu.so (HIP).
"""
from ._lib import Halo2GpuError, load  # noqa: F401

__all__ = ["Halo2GpuError", "load"]
