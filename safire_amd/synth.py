"""Synthetic closed-shell Cholesky Hamiltonians, RHF trials and auxiliary fields.

The five workloads of the walker-step benchmark, built deterministically from fixed seeds
(numpy PCG64). The construction, fixed once here:

* H1 real symmetric: off-diagonal 0.1 * (A + A^T)/sqrt(2), A ~ N(0,1); diagonal linspace(-2, 0, M).
* Cholesky vectors L^n real symmetric: s_n * (B_n + B_n^T)/sqrt(2), B_n ~ N(0,1),
  s_n = 0.5 * exp(-n / decay).  Stored as Likn[i][j][n] (n fastest), the reference layout of
  Real3IndexFactorization::Likn (Real3IndexFactorization.hpp:453-454).
* Trial: the first N columns of Q from QR of an M x M N(0,1) matrix; PsiT = Q[:, :N]^H is
  [N, M] complex (NOMSD stores orbitals as the conjugate transpose, NOMSD.cpp:133-147).
* E0 = 0, dt = 0.005, every walker starts at the trial with weight 1.
* Fields xi[step][w][n] ~ N(0,1), one seeded stream per workload (they replace probit(U) of
  construct_X_kernel.hpp:71; U is uniform so probit(U) ~ N(0,1) too).
* c5 (PHMSD): the same Hamiltonian construction at M = 64, 10 + 10 electrons, Nchol = 320; the
  trial orbitals are all M columns of Q (as rows: Q^H), with the CI expansion of
  phmsd_expansion (1000 determinants of rank 1-4).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

WORKLOADS = {
    # name: (M, N, K, W, steps, decay, ortho_interval)
    "c1": (24, 6, 96, 64, 50, 24.0, 5),
    "c2": (64, 16, 320, 512, 100, 24.0, 5),
    "c3": (128, 32, 640, 1024, 200, 64.0, 5),
    "c5": (64, 10, 320, 256, 50, 24.0, 5),
}
SEEDS = {"c1": 1001, "c2": 1002, "c3": 1003, "c4": 1004, "c5": 1005}
DT = 0.005


@dataclass
class System:
    M: int
    N: int
    K: int
    h: np.ndarray            # [M, M] float64
    Likn: np.ndarray         # [M, M, K] float64
    PsiT: np.ndarray         # [N, M] complex128
    E0: float = 0.0
    dt: float = DT
    seed: int = 0
    name: str = ""
    meta: dict = field(default_factory=dict)


def make_system(M: int, N: int, K: int, seed: int, decay: float = 24.0, scale: float = 0.5,
                name: str = "") -> System:
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((M, M))
    h = 0.1 * (A + A.T) / np.sqrt(2.0)
    h[np.diag_indices(M)] = np.linspace(-2.0, 0.0, M)
    B = rng.standard_normal((K, M, M))
    Ln = (B + B.transpose(0, 2, 1)) / np.sqrt(2.0)
    s = scale * np.exp(-np.arange(K) / decay)
    Ln *= s[:, None, None]
    Likn = np.ascontiguousarray(Ln.transpose(1, 2, 0))
    Q, _ = np.linalg.qr(rng.standard_normal((M, M)))
    PsiT = np.ascontiguousarray(Q[:, :N].conj().T).astype(np.complex128)
    return System(M=M, N=N, K=K, h=np.ascontiguousarray(h), Likn=Likn, PsiT=PsiT, seed=seed,
                  name=name, meta={"orbitals": Q})


def phmsd_expansion(M: int, N: int, ndet: int, seed: int, max_rank: int = 4):
    """The c5 CI expansion over the RHF reference (orbitals 0..N-1 of each spin): determinant 0 is
    the reference; every other determinant has a total excitation rank r uniform in 1..max_rank,
    split between the spins uniformly over the feasible (r_up, r_dn), with holes drawn uniformly
    from the occupied orbitals and particles from the virtual ones; duplicates are redrawn.
    c_k ~ N(0,1) exp(-r_k), scaled so that c_0 = 0.9 and sum_k c_k^2 = 1.
    Returns (occs [ndet][2N] int32, sorted, down spin offset by M; ci [ndet] complex128)."""
    rng = np.random.default_rng([seed, 5])
    occs = [list(range(N)) + list(range(M, M + N))]
    ranks = [0]
    seen = {tuple(occs[0])}
    while len(occs) < ndet:
        r = int(rng.integers(1, max_rank + 1))
        ru = int(rng.integers(max(0, r - N), min(r, N) + 1))
        conf = []
        for nexc, base in ((ru, 0), (r - ru, M)):
            occ = list(range(base, base + N))
            holes = rng.choice(N, nexc, replace=False)
            parts = rng.choice(np.arange(base + N, base + M), nexc, replace=False)
            for hh, p in zip(holes, parts):
                occ[hh] = int(p)
            conf += sorted(occ)
        if tuple(conf) in seen:
            continue
        seen.add(tuple(conf))
        occs.append(conf)
        ranks.append(r)
    ranks = np.array(ranks)
    c = rng.standard_normal(ndet) * np.exp(-ranks)
    c[1:] *= np.sqrt(1.0 - 0.81) / np.linalg.norm(c[1:])
    c[0] = 0.9
    return np.array(occs, np.int32), c.astype(np.complex128)


def batch_workload(count: int = 32, seed: int = SEEDS["c4"]) -> tuple[list, int, int]:
    """The uneven batch (c4): `count` independent systems with NMO drawn uniformly from
    {16, 24, ..., 96}, N = round(NMO / 4) electrons per spin, Nchol = 5 NMO, walkers uniform in
    [32, 256], each with its own seeded H1, Cholesky vectors (decay 24) and RHF trial.
    Returns ([(system, walkers), ...], steps, ortho_interval)."""
    rng = np.random.default_rng([seed, 0])
    out = []
    for i in range(count):
        M = int(rng.choice(np.arange(16, 97, 8)))
        N = max(1, int(round(M / 4)))
        W = int(rng.integers(32, 257))
        s = make_system(M, N, 5 * M, seed * 1000 + i, decay=24.0, name=f"c4.{i}")
        out.append((s, W))
    return out, 50, 5


def workload(name: str) -> tuple[System, int, int, int]:
    """Returns (system, walkers, steps, ortho_interval) of a named single-system workload."""
    M, N, K, W, steps, decay, ortho = WORKLOADS[name]
    return make_system(M, N, K, SEEDS[name], decay=decay, name=name), W, steps, ortho


def fields(seed: int, steps: int, W: int, K: int, first_step: int = 0) -> np.ndarray:
    """xi[step][w][n] ~ N(0,1); step s always draws the same values whatever the window."""
    out = np.empty((steps, W, K))
    for s in range(steps):
        rng = np.random.default_rng([seed, 7919, first_step + s])
        out[s] = rng.standard_normal((W, K))
    return out
