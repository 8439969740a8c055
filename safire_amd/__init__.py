"""MI355X-native phaseless-AFQMC walker step (Cholesky Hamiltonian) behind a C ABI.

The product is libsafire_afq.so (HIP kernels for gfx950 + the C ABI of include/afq.h);
safire_amd.afqmc is its Python mirror and safire_amd.synth builds the synthetic workloads.
"""
from . import synth  # noqa: F401

__all__ = ["synth", "afqmc"]
