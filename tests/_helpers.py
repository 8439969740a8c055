"""Shared helpers for the tests (not collected)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def cplx(a):
    a = np.asarray(a, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def tiny_system(rng, M=6, N=2, K=5, scale=0.5):
    A = rng.standard_normal((M, M))
    h = 0.1 * (A + A.T)
    B = rng.standard_normal((K, M, M))
    Ln = scale * (B + B.transpose(0, 2, 1)) / np.sqrt(2)
    Likn = np.ascontiguousarray(Ln.transpose(1, 2, 0))
    Q, _ = np.linalg.qr(rand_c(rng, M, M))
    PsiT = np.ascontiguousarray(Q[:, :N].conj().T)
    return h, Likn, PsiT
