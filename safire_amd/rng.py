"""Field and comb random numbers in step with the reference's generators (parity mode).

The reference draws every random number on the host (execute.cpp:124-125): the walker
generator (stream 0: one uniform per population control, population_control.hpp:41-42) and the
field generator (stream 1: W x rng_block_size uniforms per step, walker by walker,
AFQMCBasePropagator.cpp:708-716), both std::mt19937 seeded by split_seed(seed, rank, stream)
(utilities/Random.cpp:7-20), uniforms from std::uniform_real_distribution<double>. The step turns
the uniforms into N(0,1) fields with Acklam's probit (construct_X_kernel.hpp:71).

``HostRandomGenerator`` wraps the library's engine (afq_rng_*); ``FieldSampler`` draws a step's
uniforms on the host, copies them to the device and applies the probit there (afq_probit),
giving the [W, K] field array that ``Propagator.Propagate`` takes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import afqmc as A

WALKER_STREAM, FIELD_STREAM = 0, 1


def split_seed(seed: int, rank: int = 0, stream: int = 0) -> int:
    return int(A.lib().afq_split_seed(int(seed), int(rank), int(stream)))


class HostRandomGenerator:
    """std::mt19937 + uniform_real_distribution<double>(0, 1) (utilities/Random.hpp:36-49)."""

    def __init__(self, seed: int):
        h = C.c_void_p()
        A.check(A.lib().afq_rng_create(C.c_ulonglong(seed), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            A.lib().afq_rng_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def raw(self, n: int) -> np.ndarray:
        out = np.empty(n, np.uint32)
        A.check(A.lib().afq_rng_raw(self.handle, n, out.ctypes.data))
        return out

    def uniform(self, n: int, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(n, np.float64)
        A.check(A.lib().afq_rng_uniform(self.handle, n, out.ctypes.data))
        return out


def probit(u, out=None, stream=None):
    """Acklam's probit of a device array of uniforms (afq_probit); returns out."""
    torch = A._torch()
    if out is None:
        out = torch.empty_like(u)
    A.check(A.lib().afq_probit(u.numel(), A.ptr(u), A.ptr(out), A.stream_ptr(stream)))
    return out


class FieldSampler:
    """The fields of one step from the field generator: W x block uniforms drawn on the host
    (block = rng_block_size, the number of Cholesky vectors by default), copied to the device,
    probit on the device; returns the [W, K] view of the first K of each walker's block."""

    def __init__(self, seed: int, W: int, K: int, rank: int = 0, block: int | None = None):
        torch = A._torch()
        self.W, self.K = W, K
        self.block = K if block is None else block
        if self.block < K:
            raise ValueError("Error in assemble_X: rng_block_size < nCV.")
        self.rng = HostRandomGenerator(split_seed(seed, rank, FIELD_STREAM))
        self.host = torch.empty(W * self.block, dtype=torch.float64, pin_memory=True)
        self.dev = torch.empty(W * self.block, dtype=torch.float64, device="cuda")
        self.xi = torch.empty(W * self.block, dtype=torch.float64, device="cuda")

    def __call__(self, step=None):
        torch = A._torch()
        # the copy below reads the pinned buffer asynchronously: wait for the previous one
        torch.cuda.current_stream().synchronize()
        self.rng.uniform(self.host.numel(), self.host.numpy())
        self.dev.copy_(self.host, non_blocking=True)
        probit(self.dev, self.xi)
        xi = self.xi.view(self.W, self.block)
        return xi if self.block == self.K else xi[:, :self.K].contiguous()


class CombUniforms:
    """One uniform per population control from the walker generator (population_control.hpp:42)."""

    def __init__(self, seed: int, rank: int = 0):
        self.rng = HostRandomGenerator(split_seed(seed, rank, WALKER_STREAM))

    def __call__(self, step=None) -> float:
        return float(self.rng.uniform(1)[0])
