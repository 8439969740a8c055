"""The oracle's BackPropagate (AFQMCBasePropagator.cpp:403-532) pinned by algebra rather than by
restating it: the adjoint propagator of each step is the Hermitian conjugate of the forward one
(built column by column with the oracle's forward propagate, which the propagation tests pin),
the ring is read newest first with wrap-around, and intermediate orthogonalisations change
neither the final Q (unique: R has a positive real diagonal) nor the total log determinant."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._helpers import rand_c


def system(rng, M, K):
    L = rng.standard_normal((M, M, K)) * 0.3
    L = 0.5 * (L + L.transpose(1, 0, 2))
    P1 = np.eye(M) + 0.1 * rand_c(rng, M, M)
    return L, P1


def forward_matrices(P1, V, order):
    W, M, _ = V.shape
    eye = np.broadcast_to(np.eye(M, dtype=np.complex128), (W, M, M)).copy()
    return O.propagate(eye, P1, V, order)


def wrap(d):
    d = np.array(d)
    d.imag = (d.imag + np.pi) % (2 * np.pi) - np.pi
    return d


@pytest.mark.parametrize("pos,nbp", [(6, 3), (2, 2), (4, 4)])
def test_back_propagation_is_the_adjoint_of_the_forward_steps(pos, nbp):
    rng = np.random.default_rng(pos * 10 + nbp)
    W, M, N, K, nf, nrefs, dt, order = 3, 6, 2, 5, 4, 2, 0.05, 6
    L, P1 = system(rng, M, K)
    fields = rand_c(rng, W, nf, K)
    Refs = rand_c(rng, W, nrefs, M, N)
    out, ld = O.back_propagate(fields, pos, nbp, 100, Refs, P1, L, dt, order)
    # newest first: slots pos-1, pos-2, ... (mod nf)
    S = Refs.copy()
    for j in range(nbp):
        slot = (pos - 1 - j) % nf
        B = forward_matrices(P1, O.vhs(np.ascontiguousarray(fields[:, slot]), L, dt), order)
        S = np.einsum("wji,wrjk->wrik", B.conj(), S)
    for nr in range(nrefs):
        Q = out[:, nr]
        np.testing.assert_allclose(np.einsum("wji,wjk->wik", Q.conj(), Q),
                                   np.broadcast_to(np.eye(N), (W, N, N)), atol=1e-12)
        R = np.einsum("wji,wjk->wik", Q.conj(), S[:, nr])
        np.testing.assert_allclose(np.einsum("wij,wjk->wik", Q, R), S[:, nr],
                                   rtol=1e-11, atol=1e-12 * np.abs(S).max())
        np.testing.assert_allclose(np.linalg.det(R), np.exp(ld[:, nr]), rtol=1e-10)


def test_intermediate_orthogonalisation_changes_nothing():
    rng = np.random.default_rng(7)
    W, M, N, K, nf = 2, 7, 3, 4, 5
    L, P1 = system(rng, M, K)
    fields = rand_c(rng, W, nf, K)
    Refs = rand_c(rng, W, 1, M, N)
    a, la = O.back_propagate(fields, 7, 5, 1, Refs, P1, L, 0.05)
    b, lb = O.back_propagate(fields, 7, 5, 100, Refs, P1, L, 0.05)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)
    assert np.abs(wrap(la - lb)).max() < 1e-10
