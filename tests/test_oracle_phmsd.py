"""PHMSD energy corrections of excited determinants (detail/phmsd_impl.hpp), oracle side.

Pinned by the literal EX/EJ values of the reference's component test (closed-form sin/cos
inputs, nex = 1..5) and, for the parts those values cannot isolate, by brute force: the
2-body routine is a sum of per-determinant terms that a direct numpy evaluation of the same
formula reproduces on random inputs, and both routines are linear in the weights.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._helpers import load_json, rand_c


def kat_inputs(nex, ndet=50, nelec=5, nact=10, nchol=50, nw=1):
    """The closed-form inputs of the component test for nex excitations per determinant."""
    idx = np.arange(nw * nelec * nchol * nact, dtype=np.float64)
    Twina = (np.sin(idx) + 1j * np.cos(idx * idx)).reshape(nw, nelec, nchol, nact)
    idx = np.arange(nw * ndet * nex * nact, dtype=np.float64)
    R = (np.sin(3 * idx) + 1j * np.cos(idx + idx * idx)).reshape(nw, ndet, nex, nact)
    idx = np.arange(ndet * nw, dtype=np.float64)
    wgt = (np.sin(2 * idx) + 1j * np.cos(idx - 2 * idx * idx)).reshape(ndet, nw)
    iexcit = np.zeros((ndet, 2 * nex), np.int32)
    for d in range(ndet):
        taken = np.zeros(nelec, bool)
        for e in range(nex):
            v = (e * e) % nelec
            while taken[v]:
                v = (v + 1) % nelec
            taken[v] = True
            iexcit[d, e] = v
        taken = np.zeros(nact - nelec, bool)
        for e in range(nex):
            v = (e + e * e) % (nact - nelec)
            while taken[v]:
                v = (v + 1) % (nact - nelec)
            taken[v] = True
            iexcit[d, e + nex] = v + nelec
    refc = np.arange(nelec, dtype=np.int32)
    return iexcit, refc, Twina, R, wgt


@pytest.mark.parametrize("nex", [1, 2, 3, 4, 5])
def test_ph_2body_literal_values(nex):
    ref = np.array(load_json("phmsd_2body_kat.json")["EX_EJ_by_nex"][nex - 1])
    EX, EJ, _ = O.ph_excited_2body(*kat_inputs(nex))
    np.testing.assert_allclose(EX[0], ref[0, 0] + 1j * ref[0, 1], rtol=1e-11)
    np.testing.assert_allclose(EJ[0], ref[1, 0] + 1j * ref[1, 1], rtol=1e-11)


def brute_2body(iexcit, refc, Twina, R, wgt):
    """The same energy written as dense numpy contractions, determinant by determinant."""
    nw, ndet, nex, nact = R.shape
    nelec = Twina.shape[1]
    EX = np.zeros(nw, complex)
    EJ = np.zeros(nw, complex)
    for w in range(nw):
        T = Twina[w]                                      # [i][n][a]
        F0 = sum(T[i, :, refc[i]] for i in range(nelec))
        for d in range(ndet):
            p, v = iexcit[d, :nex], iexcit[d, nex:]
            occ = np.array(refc, copy=True)
            for e in range(nex):
                occ[p[e]] = v[e]
            Rd = R[w, d]
            Y = np.einsum("ena,fa->efn", T[p], Rd)            # Y[e][f] = T[p_e] R_f
            eX = np.einsum("een,een->", Y, Y)
            for e in range(nex):
                for f in range(e + 1, nex):
                    eX += 2 * np.sum(Y[f, e] * Y[e, f])
                for j in range(nelec):
                    eX += 2 * np.sum((T[j].T @ T[p[e], :, occ[j]]) * Rd[e])
            for i in range(nelec):
                if occ[i] != refc[i]:
                    eX += np.sum(T[i, :, occ[i]] ** 2 - T[i, :, refc[i]] ** 2)
                for j in range(i + 1, nelec):
                    if occ[i] != refc[i] or occ[j] != refc[j]:
                        eX += 2 * np.sum(T[i, :, occ[j]] * T[j, :, occ[i]]
                                         - T[i, :, refc[j]] * T[j, :, refc[i]])
            Kn = np.einsum("een->n", Y) + sum(T[i, :, occ[i]] for i in range(nelec))
            EX[w] -= 0.5 * wgt[d, w] * eX
            EJ[w] += 0.5 * wgt[d, w] * (np.sum(Kn * Kn) - np.sum(F0 * F0))
    return EX, EJ


@pytest.mark.parametrize("nex", [1, 3])
def test_ph_2body_brute_force_and_linearity(nex):
    rng = np.random.default_rng(nex)
    nw, ndet, nelec, nact, nchol = 3, 7, 4, 9, 6
    iexcit = np.stack([np.concatenate([rng.choice(nelec, nex, replace=False),
                                       nelec + rng.choice(nact - nelec, nex, replace=False)])
                       for _ in range(ndet)]).astype(np.int32)
    refc = np.arange(nelec, dtype=np.int32)
    Twina, R, wgt = rand_c(rng, nw, nelec, nchol, nact), rand_c(rng, nw, ndet, nex, nact), \
        rand_c(rng, ndet, nw)
    EX, EJ, KE = O.ph_excited_2body(iexcit, refc, Twina, R, wgt)
    bx, bj = brute_2body(iexcit, refc, Twina, R, wgt)
    np.testing.assert_allclose(EX, bx, rtol=1e-12)
    np.testing.assert_allclose(EJ, bj, rtol=1e-12)
    EX2, EJ2, KE2 = O.ph_excited_2body(iexcit, refc, Twina, R, 2.5j * wgt)
    np.testing.assert_allclose(EX2, 2.5j * EX, rtol=1e-13)
    np.testing.assert_allclose(KE2, KE, rtol=0, atol=0)   # KE does not depend on the weights


def test_ph_1body_brute_force():
    rng = np.random.default_rng(11)
    nw, ndet, nex, nelec, nact = 2, 5, 2, 4, 8
    iexcit = np.stack([np.concatenate([rng.choice(nelec, nex, replace=False),
                                       nelec + rng.choice(nact - nelec, nex, replace=False)])
                       for _ in range(ndet)]).astype(np.int32)
    refc = np.arange(nelec, dtype=np.int32)
    S, R, wgt = rand_c(rng, nw, nelec, nact), rand_c(rng, nw, ndet, nex, nact), rand_c(rng, ndet, nw)
    E = O.ph_excited_1body(iexcit, refc, S, R, wgt)
    ref = np.zeros(nw, complex)
    for w in range(nw):
        for d in range(ndet):
            p, v = iexcit[d, :nex], iexcit[d, nex:]
            ref[w] += wgt[d, w] * sum(S[w, p[e]] @ R[w, d, e] + S[w, p[e], v[e]]
                                      - S[w, p[e], refc[p[e]]] for e in range(nex))
    np.testing.assert_allclose(E, ref, rtol=1e-13)
