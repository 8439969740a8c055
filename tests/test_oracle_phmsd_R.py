"""The oracle's PHMSD excitation overlaps and R matrices (calculate_overlaps,
get_compact_ph_R_matrices, calculate_R of phmsd_helpers.hpp) against what linear algebra fixes:
the overlap ratio of an excitation is det A with A(p, q) = T[e_(nex+p)][e_q] (numpy det); the
compact R row p is, at the orbital each electron occupies after the excitation,
Q(p, q0) - delta(p, q0) for the excited electron e_q0 and -(Q T[targets, :])(p, i) for the
others, with Q = A^-1 (numpy inv); calculate_R is the weighted sum of the reference term, the
identity terms and the compact rows."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._helpers import rand_c


def shells_for(rng, nelec, nact, ndet, nex):
    ex = np.zeros((ndet, 2 * nex), np.int32)
    for d in range(ndet):
        ex[d, :nex] = rng.choice(nelec, nex, replace=False)
        ex[d, nex:] = nelec + rng.choice(nact - nelec, nex, replace=False)
    return ex


@pytest.mark.parametrize("nex", [1, 2, 3, 4, 5, 6])
def test_overlaps_are_subdeterminants(nex):
    rng = np.random.default_rng(nex)
    nw, nelec, nact, ndet = 3, 7, 16, 9
    T = rand_c(rng, nw, nact, nelec)
    ex = shells_for(rng, nelec, nact, ndet, nex)
    ov = O.ph_overlaps(ex, T)
    for d in range(ndet):
        for w in range(nw):
            A = T[w][np.ix_(ex[d, nex:], ex[d, :nex])]
            assert abs(ov[d, w] - np.linalg.det(A)) < 1e-12 * max(1, abs(np.linalg.det(A)))


def test_singular_excitation_gives_zero():
    nw, nelec, nact = 1, 3, 6
    T = np.ones((nw, nact, nelec), np.complex128)          # every 2x2 minor vanishes
    ex = np.array([[0, 1, 3, 4]], np.int32)
    assert O.ph_overlaps(ex, T)[0, 0] == 0
    assert not np.any(O.ph_compact_R(ex, np.arange(nelec), T))


@pytest.mark.parametrize("nex", [1, 2, 3, 4])
def test_compact_R_closed_form(nex):
    rng = np.random.default_rng(10 + nex)
    nw, nelec, nact, ndet = 2, 6, 14, 5
    T = rand_c(rng, nw, nact, nelec)
    refc = np.arange(nelec)
    ex = shells_for(rng, nelec, nact, ndet, nex)
    R = O.ph_compact_R(ex, refc, T)
    for w in range(nw):
        for d in range(ndet):
            e, v = ex[d, :nex], ex[d, nex:]
            Q = np.linalg.inv(T[w][np.ix_(v, e)])
            QT = Q @ T[w][v, :]                              # [nex][nelec]
            ref = np.zeros((nex, nact), np.complex128)
            orbs = refc.copy()
            orbs[e] = v
            for i in range(nelec):
                for p in range(nex):
                    ref[p, orbs[i]] += -QT[p, i]
            for q0 in range(nex):
                ref[:, v[q0]] += Q[:, q0]
            # -QT at an excited electron plus Q(p, q0) = Q(p, q0) - delta(p, q0)
            np.testing.assert_allclose(R[w, d], ref, rtol=1e-12, atol=1e-12)
            for q0 in range(nex):
                np.testing.assert_allclose(R[w, d, :, v[q0]], Q[:, q0] - np.eye(nex)[:, q0],
                                           rtol=1e-11, atol=1e-12)


def test_calculate_R_is_the_weighted_sum():
    rng = np.random.default_rng(3)
    nw, nelec, nact = 2, 5, 12
    T = rand_c(rng, nw, nact, nelec)
    refc = np.arange(nelec)
    shells = [shells_for(rng, nelec, nact, n, k + 1) for k, n in enumerate([4, 3, 2])]
    nd = 1 + sum(len(s) for s in shells)
    wgt = rand_c(rng, nd, nw)
    R = O.ph_calculate_R(shells, refc, T, wgt)
    ref = np.zeros((nw, nelec, nact), np.complex128)
    for i in range(nelec):
        ref[:, i, refc[i]] += wgt[0]
    row = 1
    for sh in shells:
        Rc = O.ph_compact_R(sh, refc, T)
        for d, e in enumerate(sh):
            nex = len(e) // 2
            orbs = refc.copy()
            orbs[e[:nex]] = e[nex:]
            for w in range(nw):
                for i in range(nelec):
                    ref[w, i, orbs[i]] += wgt[row, w]
                for p in range(nex):
                    ref[w, e[p]] += wgt[row, w] * Rc[w, d, p]
            row += 1
    np.testing.assert_allclose(R, ref, rtol=1e-12, atol=1e-12)


def test_woodbury_mixed_density_matrix_closed_forms():
    """MixedDensityMatrixForWoodbury: TNN = (A Phi)[ref rows]; ov accumulates log det TNN
    (wrapped imaginary part); QQ0 = (A Phi) TNN^-1; compact G = TNN^-T Phi^T. numpy's det, inv
    and products give the same numbers independently."""
    rng = np.random.default_rng(11)
    nb, nmo, nel, nact = 4, 20, 5, 12
    A = rand_c(rng, nact, nmo)
    ref = np.array([0, 2, 3, 7, 9])
    Phi = rand_c(rng, nb, nmo, nel)
    ov0 = rand_c(rng, nb) * 0.1
    G, ov, QQ0 = O.ph_woodbury(A, ref, Phi, ov0)
    for w in range(nb):
        TAB = A @ Phi[w]
        TNN = TAB[ref]
        inv = np.linalg.inv(TNN)
        np.testing.assert_allclose(QQ0[w], TAB @ inv, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(G[w], inv.T @ Phi[w].T, rtol=1e-12, atol=1e-12)
        d = ov[w] - ov0[w] - np.log(np.linalg.det(TNN))
        d = d.real + 1j * ((d.imag + np.pi) % (2 * np.pi) - np.pi)
        assert abs(d) < 1e-12
        assert -np.pi <= ov[w].imag <= np.pi
        # the reference rows of QQ0 are the identity (TNN TNN^-1)
        np.testing.assert_allclose(QQ0[w][ref], np.eye(nel), atol=1e-12)


def random_csr(rng, nu, nd, density=0.3):
    """A CSR matrix [nu][nd] with random complex values, some empty rows, and a dense twin."""
    C = np.zeros((nu, nd), complex)
    for u in range(nu):
        if u % 5 == 4:
            continue                                   # an up excitation with no partner
        cols = rng.choice(nd, max(1, int(density * nd)), replace=False)
        C[u, cols] = rng.standard_normal(len(cols)) + 1j * rng.standard_normal(len(cols))
    cp = np.zeros(nu + 1, np.int32)
    ci, cv = [], []
    for u in range(nu):
        nz = np.nonzero(C[u])[0]
        ci += list(nz)
        cv += list(C[u, nz])
        cp[u + 1] = len(ci)
    return C, cp, np.array(ci, np.int32), np.array(cv, complex)


def test_spin_weights_is_the_dense_product():
    """PHMSD.cpp:441-450: wgt = ratio_up * (C ratio_dn), Ov = sum over up excitations of wgt;
    pinned by numpy's dense product and, for a single coupling, by the closed form c r_u r_d."""
    rng = np.random.default_rng(11)
    nu, nd, nw = 23, 17, 9
    C, cp, ci, cv = random_csr(rng, nu, nd)
    rup, rdn = rand_c(rng, nu, nw), rand_c(rng, nd, nw)
    wgt, ov = O.ph_spin_weights(cp, ci, cv, rup, rdn)
    W = rup * (C @ rdn)
    np.testing.assert_allclose(wgt, W, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(ov, W.sum(axis=0), rtol=1e-12)
    assert np.all(wgt[4::5] == 0)
    # one determinant pair: the multi-determinant overlap is c * ratio_up * ratio_dn
    wgt1, ov1 = O.ph_spin_weights(np.array([0, 1], np.int32), np.array([0], np.int32),
                                  np.array([0.5 - 2j]), rup[:1], rdn[:1])
    np.testing.assert_allclose(ov1, (0.5 - 2j) * rup[0] * rdn[0], rtol=1e-15)
