"""Pins the oracle's det_ops restatement (density_matrix.hpp, determinants_impl.hpp,
nda_functions.hpp, orthogonalize.hpp, propagate.hpp) against the reference's literal
known-answer values and against independent numpy/scipy/brute-force routes."""
import itertools
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from tests._helpers import cplx, load_json, rand_c, relerr

KAT = load_json("sdet_ops_kat.json")


def kat_inputs(nwalk=5):
    A = cplx(KAT["A"])
    B = np.broadcast_to(cplx(KAT["B"]), (nwalk, 4, 3)).copy()
    return A, B


# ---- test_sdet_ops.cpp:99-144 (Log_Overlap, full and sub-range) ------------------------------
def test_log_overlap_kat():
    A, B = kat_inputs()
    ov = O.log_overlap(A, B)
    ref = np.log(cplx(KAT["ov"])[0])
    np.testing.assert_allclose(ov, ref, rtol=1e-8, atol=1e-8)
    ov2 = O.log_overlap(A[:2, :3], B[:, :3, :2].copy())
    np.testing.assert_allclose(ov2, np.log(cplx(KAT["ov2"])[0]), rtol=1e-8, atol=1e-8)


# ---- test_sdet_ops.cpp:146-205 (MixedDensityMatrix full / compact / sub-range) ---------------
def test_mixed_density_matrix_kat():
    A, B = kat_inputs()
    G, ov = O.mixed_density_matrix(A, B, compact=False)
    np.testing.assert_allclose(G, np.broadcast_to(cplx(KAT["g_ref"]), G.shape), atol=1e-8,
                               rtol=1e-8)
    np.testing.assert_allclose(ov, np.log(cplx(KAT["ov"])[0]), atol=1e-8)
    Gc, ov = O.mixed_density_matrix(A, B, compact=True)
    np.testing.assert_allclose(Gc, np.broadcast_to(cplx(KAT["gc_ref"]), Gc.shape), atol=1e-8,
                               rtol=1e-8)
    G2, _ = O.mixed_density_matrix(A[:2, :3], B[:, :3, :2].copy(), compact=False)
    np.testing.assert_allclose(G2, np.broadcast_to(cplx(KAT["g_ref_2"]), G2.shape), atol=1e-8)
    Gc2, _ = O.mixed_density_matrix(A[:2, :3], B[:, :3, :2].copy(), compact=True)
    np.testing.assert_allclose(Gc2, np.broadcast_to(cplx(KAT["gc_ref_2"]), Gc2.shape),
                               atol=1e-8)


# ---- test_sdet_ops.cpp:333-356 (inverse_logdet; literals given to 8 decimals) ----------------
def test_inverse_logdet_kat():
    A = np.broadcast_to(cplx(KAT["inv_A"]), (5, 4, 4)).copy()
    inv, ld = O.inverse_logdet(A)
    np.testing.assert_allclose(inv, np.broadcast_to(cplx(KAT["inv_ref"]), inv.shape), atol=2e-8)
    np.testing.assert_allclose(ld, cplx(KAT["logdet_ref"])[0], atol=1e-8)


def _det_permutation(A):
    n = A.shape[0]
    s = 0j
    for p in itertools.permutations(range(n)):
        inv = sum(1 for i in range(n) for j in range(i + 1, n) if p[i] > p[j])
        s += (-1) ** inv * np.prod([A[i, p[i]] for i in range(n)])
    return s


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_getrf_logdet_vs_permutation_expansion(n):
    rng = np.random.default_rng(10 + n)
    for _ in range(20):
        A = rand_c(rng, n, n)
        LU, ipiv, info = O.getrf(A)
        assert info == 0
        ld = O.log_det_from_getrf(LU, ipiv)
        assert -np.pi <= ld.imag <= np.pi
        det = _det_permutation(A)
        np.testing.assert_allclose(np.exp(ld), det, rtol=1e-12)
        # principal branch: identical to log(det)
        np.testing.assert_allclose(ld, np.log(det), rtol=1e-12, atol=1e-12)


def test_getrf_pivot_is_lapack_cabs1():
    # column 0 has |re|+|im| maximal at row 2 (0.1+0.95i: 1.05) although |.| is larger at row 1
    A = np.array([[0.1, 1, 2], [0.8 + 0.0j, 3, 1], [0.1 + 0.95j, 1, 5]], np.complex128)
    LU, ipiv, info = O.getrf(A)
    assert ipiv[0] == 3
    P, L, U = scipy.linalg.lu(A)  # LAPACK getrf underneath
    np.testing.assert_allclose(np.abs(np.diag(LU)), np.abs(np.diag(U)), rtol=1e-12)


def test_accumulated_logdet_wraps_to_pi():
    # ov accumulates; the wrap acts on the accumulated value (determinants_impl.hpp:56-73)
    A = np.array([[-2.0 + 0.5j]])
    LU, ipiv, _ = O.getrf(A)
    ld = O.log_det_from_getrf(LU, ipiv, res=0.5 + 3.0j)
    raw = 0.5 + 3.0j + np.log(-2.0 + 0.5j)
    assert -np.pi <= ld.imag <= np.pi
    np.testing.assert_allclose(ld, raw - 2j * np.pi, rtol=1e-14)


def test_getri_or_zero_singular():
    A = np.array([[1, 2], [2, 4]], np.complex128)
    LU, ipiv, info = O.getrf(A)
    assert info == 2
    inv, info2 = O.getri_or_zero(LU, ipiv)
    assert info2 != 0 and np.all(inv == 0)
    # singular overlap: G is zero, log-overlap finite (DBL_MIN guard)
    PsiT = np.array([[1, 0, 0], [1, 0, 0]], np.complex128)
    Phi = np.array([[[1, 0], [0, 1], [0, 0]]], np.complex128)
    G, ov = O.mixed_density_matrix(PsiT, Phi)
    assert np.all(G == 0) and np.isfinite(ov[0].real) and ov[0].real < -700


def test_mixed_density_matrix_vs_numpy():
    rng = np.random.default_rng(3)
    M, N, nb = 9, 4, 3
    PsiT = rand_c(rng, N, M)
    Phi = rand_c(rng, nb, M, N)
    G, ov = O.mixed_density_matrix(PsiT, Phi, compact=False)
    Gc, ovc = O.mixed_density_matrix(PsiT, Phi, compact=True)
    for w in range(nb):
        T = PsiT @ Phi[w]
        # <c+_i c_j> = [Phi T^-1 PsiT]_{ji}
        Gref = (Phi[w] @ np.linalg.inv(T) @ PsiT).T
        assert relerr(G[w], Gref) < 1e-12
        assert relerr(Gc[w], (Phi[w] @ np.linalg.inv(T)).T) < 1e-12
        np.testing.assert_allclose(np.exp(ov[w]), np.linalg.det(T), rtol=1e-12)
    np.testing.assert_allclose(ov, ovc)


def test_propagate_vs_expm():
    rng = np.random.default_rng(5)
    M, N, nb = 8, 3, 2
    Phi = rand_c(rng, nb, M, N)
    P1 = rand_c(rng, M, M) * 0.3
    V = 0.02 * rand_c(rng, nb, M, M)
    out = O.propagate(Phi, P1, V, order=6)
    for w in range(nb):
        taylor = sum(np.linalg.matrix_power(1j * V[w], k) / math.factorial(k) for k in range(7))
        ex = scipy.linalg.expm(1j * V[w])
        ref = P1 @ ex @ P1 @ Phi[w]
        # truncation error of the order-6 Taylor series is ~||V||^7/7!
        assert relerr(out[w], ref) < 1e-10
        assert relerr(out[w], P1 @ taylor @ P1 @ Phi[w]) < 1e-13
    # order 1 keeps only I + iV
    out1 = O.propagate(Phi, P1, V, order=1)
    ref1 = P1 @ (np.eye(M) + 1j * V[0]) @ P1 @ Phi[0]
    assert relerr(out1[0], ref1) < 1e-13


def test_orthogonalize_unique_qr():
    rng = np.random.default_rng(7)
    M, N, nb = 10, 4, 3
    Phi = rand_c(rng, nb, M, N)
    Q, ldet = O.orthogonalize(Phi)
    for w in range(nb):
        q, r = np.linalg.qr(Phi[w])
        ph = np.diag(r) / np.abs(np.diag(r))
        qref = q * ph[None, :]          # the QR with a positive real diagonal is unique
        assert relerr(Q[w], qref) < 1e-12
        np.testing.assert_allclose(Q[w].conj().T @ Q[w], np.eye(N), atol=1e-13)
        np.testing.assert_allclose(ldet[w].real, np.sum(np.log(np.abs(np.diag(r)))), rtol=1e-12)
    # test_sdet_ops.cpp:294-304: Log_Overlap(Q, Q) = 0
    ov = O.log_overlap(Q[0].conj().T.copy(), Q)
    np.testing.assert_allclose(ov[0], 0.0, atol=1e-12)


def test_orthogonalize_kat_invariant():
    A, B = kat_inputs()
    Q, _ = O.orthogonalize(B)
    ov = O.log_overlap(Q[0].conj().T.copy(), Q)
    np.testing.assert_allclose(ov, 0.0, atol=1e-8)


def test_exp_hermitian_vs_expm():
    rng = np.random.default_rng(11)
    for n in (1, 3, 7, 16):
        X = rand_c(rng, n, n)
        H = 0.3 * (X + X.conj().T)
        np.testing.assert_allclose(O.exp_hermitian(H), scipy.linalg.expm(H), rtol=1e-12,
                                   atol=1e-12)
