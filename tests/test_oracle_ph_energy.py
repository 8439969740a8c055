"""The oracle's PHMSD local energy (PHMSD::energy_alg1 restated) pinned by a brute-force
multi-determinant evaluation: every determinant of the CI expansion is built explicitly, its
overlap and mixed-estimator energy are evaluated by Wick's theorem from the full one-body density
matrix, and the CI-weighted average is compared component by component."""
import numpy as np
import pytest

from oracle import ph_energy as P
from tests._helpers import rand_c


def ph_system(rng, nmo, nup, ndn, nchol, ndet, max_rank, uhf=False):
    """A random COLLINEAR PHMSD system. Returns a dict with the Hamiltonian (h [nmo][nmo],
    L [nchol][nmo][nmo]), orbital rows A[s] [nmo][nmo], the half-rotated tensors, the CI
    expansion (occs, ci) and the derived structure."""
    X = rng.standard_normal((nmo, nmo))
    h = 0.1 * (X + X.T) + np.diag(np.linspace(-2, 0, nmo))
    B = rng.standard_normal((nchol, nmo, nmo))
    L = 0.5 * (B + B.transpose(0, 2, 1)) / np.sqrt(2) * np.exp(-np.arange(nchol) / 6)[:, None, None]
    A = []
    for s in range(2 if uhf else 1):
        Q, _ = np.linalg.qr(rand_c(rng, nmo, nmo))
        A.append(np.ascontiguousarray(Q.conj().T))           # orbitals as rows (PsiT)
    if not uhf:
        A.append(A[0])
    occs, dets = [], set()
    ref = (tuple(range(nup)), tuple(range(nmo, nmo + ndn)))
    occs.append(list(ref[0]) + list(ref[1]))
    dets.add(ref)
    while len(occs) < ndet:
        conf = []
        for s, (ne, base) in enumerate([(nup, 0), (ndn, nmo)]):
            r = rng.integers(0, min(max_rank, ne) + 1)
            occ = list(range(base, base + ne))
            holes = rng.choice(ne, r, replace=False)
            virt = rng.choice(np.arange(base + ne, base + nmo), r, replace=False)
            for hh, v in zip(holes, virt):
                occ[hh] = v
            conf.append(tuple(sorted(occ)))
        if tuple(conf) in dets:
            continue
        dets.add(tuple(conf))
        occs.append(list(conf[0]) + list(conf[1]))
    rank = np.array([sum(o not in ref[0] + ref[1] for o in oc) for oc in occs])
    ci = (rng.standard_normal(ndet) + 0.3j * rng.standard_normal(ndet)) * np.exp(-rank)
    ci[0] = 0.9
    occs = np.array(occs, np.int32)
    Lna = [np.ascontiguousarray(np.einsum("al,nlk->nak", A[s], L)) for s in range(2)]
    haj = [np.ascontiguousarray(A[s] @ h) for s in range(2)]
    return {"h": h, "L": L, "A": A, "Lna": Lna, "haj": haj, "occs": occs, "ci": ci,
            "nmo": nmo, "nup": nup, "ndn": ndn,
            "struct": P.build_ph_struct(ci, occs, nmo, nup, ndn)}


def brute_force(sysd, Phi_up, Phi_dn, E0):
    """sum_D conj(c_D) <D|phi> E_D / sum_D conj(c_D) <D|phi>, per energy component, with
    rho_s = Phi (A_D Phi)^-1 A_D: E1 = E0 + sum_s tr(h rho_s), EXX = -1/2 sum_ns tr(L rho L rho),
    EJ = 1/2 sum_n (sum_s tr(L_n rho_s))^2."""
    h, L, A, nmo, nup = sysd["h"], sysd["L"], sysd["A"], sysd["nmo"], sysd["nup"]
    nw = Phi_up.shape[0]
    num = np.zeros((nw, 3), complex)
    den = np.zeros(nw, complex)
    for occ, c in zip(sysd["occs"], sysd["ci"]):
        rows = [list(occ[:nup]), [o - nmo for o in occ[nup:]]]
        for w in range(nw):
            ov, E = 1.0, np.array([E0, 0.0, 0.0], complex)
            trL = np.zeros(L.shape[0], complex)
            for s, Phi in enumerate([Phi_up[w], Phi_dn[w]]):
                Ad = A[s][rows[s]]
                M = Ad @ Phi
                ov *= np.linalg.det(M)
                rho = Phi @ np.linalg.solve(M, Ad)
                E[0] += np.trace(h @ rho)
                LR = L @ rho
                E[1] += -0.5 * np.einsum("nij,nji->", LR, LR)
                trL += np.einsum("nii->n", LR)
            E[2] = 0.5 * np.sum(trL * trL)
            num[w] += np.conj(c) * ov * E
            den[w] += np.conj(c) * ov
    return num / den[:, None], den


def test_excitation_sign_and_positions():
    # orbital 1 (position 1) -> 5: in-place order [0, 5, 2] needs one exchange to sort
    holes, parts, sign = P.excitation([0, 1, 2], [0, 2, 5])
    assert holes == [1] and parts == [5] and sign == -1
    # 0 -> 3 and 2 -> 4: in place [3, 1, 4] -> sorted [1, 3, 4], one exchange
    holes, parts, sign = P.excitation([0, 1, 2], [1, 3, 4])
    assert holes == [0, 2] and parts == [3, 4] and sign == -1
    assert P.excitation([0, 1, 2], [0, 1, 7]) == ([2], [7], 1)


def test_struct_couplings_reproduce_the_expansion():
    rng = np.random.default_rng(3)
    s = ph_system(rng, 9, 3, 2, 4, 25, 3)
    st = s["struct"]
    nu, nd = st["nunique"]
    assert st["C"][0].shape == (nu, nd) and np.allclose(st["C"][1], st["C"][0].T)
    # sum over the couplings of |c| equals the expansion's (signs aside)
    assert np.isclose(np.abs(st["C"][0]).sum(), np.abs(s["ci"]).sum())
    assert st["C"][0][0, 0] == 0.9
    for spin in range(2):
        cp, ci, cv = st["csr"][spin]
        dense = np.zeros_like(st["C"][spin])
        for r in range(len(cp) - 1):
            dense[r, ci[cp[r]:cp[r + 1]]] += cv[cp[r]:cp[r + 1]]
        np.testing.assert_array_equal(dense, st["C"][spin])


@pytest.mark.parametrize("uhf", [False, True])
def test_energy_alg1_matches_brute_force(uhf):
    rng = np.random.default_rng(17 + uhf)
    s = ph_system(rng, 9, 4, 3, 7, 40, 3, uhf=uhf)
    nw = 5
    Phi_up, Phi_dn = rand_c(rng, nw, 9, 4), rand_c(rng, nw, 9, 3)
    E0 = 1.25
    E, logov = P.energy_alg1(s["struct"], s["A"], s["Lna"], s["haj"], E0, Phi_up, Phi_dn)
    Eb, ovb = brute_force(s, Phi_up, Phi_dn, E0)
    scale = np.abs(Eb).max()
    np.testing.assert_allclose(E, Eb, rtol=1e-10, atol=1e-11 * scale)
    np.testing.assert_allclose(np.exp(logov), ovb, rtol=1e-10)


def test_single_determinant_reduces_to_the_reference_energy():
    rng = np.random.default_rng(5)
    s = ph_system(rng, 7, 3, 3, 5, 1, 1)
    Phi_up, Phi_dn = rand_c(rng, 3, 7, 3), rand_c(rng, 3, 7, 3)
    E, logov = P.energy_alg1(s["struct"], s["A"], s["Lna"], s["haj"], 0.0, Phi_up, Phi_dn)
    Eb, ovb = brute_force(s, Phi_up, Phi_dn, 0.0)
    np.testing.assert_allclose(E, Eb, rtol=1e-11, atol=1e-12 * np.abs(Eb).max())
    np.testing.assert_allclose(np.exp(logov), 0.9 * ovb / 0.9, rtol=1e-11)


@pytest.mark.parametrize("uhf,ndet", [(False, 40), (True, 40), (False, 1)])
def test_energy_alg0_matches_brute_force(uhf, ndet):
    """energy_alg0 (one G and energy per unique configuration) against every determinant built
    explicitly; ndet = 1 is the single-determinant energy."""
    rng = np.random.default_rng(23 + uhf + ndet)
    s = ph_system(rng, 9, 4, 3, 7, ndet, 3, uhf=uhf)
    nw = 5
    Phi_up, Phi_dn = rand_c(rng, nw, 9, 4), rand_c(rng, nw, 9, 3)
    E0 = 1.25
    E, logov = P.energy_alg0(s["struct"], s["A"], s["Lna"], s["haj"], E0, Phi_up, Phi_dn)
    Eb, ovb = brute_force(s, Phi_up, Phi_dn, E0)
    scale = np.abs(Eb).max()
    np.testing.assert_allclose(E, Eb, rtol=1e-10, atol=1e-11 * scale)
    np.testing.assert_allclose(np.exp(logov), ovb, rtol=1e-10)


def test_energy_algorithms_agree():
    """The reference algorithm and the Woodbury one are the same local energy: rank-1..4
    excitations, 120 determinants, UHF orbitals."""
    rng = np.random.default_rng(41)
    s = ph_system(rng, 14, 5, 4, 9, 120, 4, uhf=True)
    Phi_up, Phi_dn = rand_c(rng, 6, 14, 5), rand_c(rng, 6, 14, 4)
    E0, ov0 = P.energy_alg0(s["struct"], s["A"], s["Lna"], s["haj"], 0.5, Phi_up, Phi_dn)
    E1, ov1 = P.energy_alg1(s["struct"], s["A"], s["Lna"], s["haj"], 0.5, Phi_up, Phi_dn)
    np.testing.assert_allclose(E0, E1, rtol=1e-10, atol=1e-11 * np.abs(E1).max())
    d = ov0 - ov1
    d.imag = (d.imag + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-10


def brute_force_density(sysd, Phi_up, Phi_dn):
    """The full multi-determinant mixed density matrix of each spin,
    sum_D conj(c_D) <D|phi> rho_D^T / sum_D conj(c_D) <D|phi> with rho_D = Phi (A_D Phi)^-1 A_D."""
    A, nmo, nup = sysd["A"], sysd["nmo"], sysd["nup"]
    nw = Phi_up.shape[0]
    num = np.zeros((nw, 2, nmo, nmo), complex)
    den = np.zeros(nw, complex)
    for occ, c in zip(sysd["occs"], sysd["ci"]):
        rows = [list(occ[:nup]), [o - nmo for o in occ[nup:]]]
        for w in range(nw):
            ov, rho = 1.0, []
            for s, Phi in enumerate([Phi_up[w], Phi_dn[w]]):
                Ad = A[s][rows[s]]
                M = Ad @ Phi
                ov *= np.linalg.det(M)
                rho.append(Phi @ np.linalg.solve(M, Ad))
            for s in range(2):
                num[w, s] += np.conj(c) * ov * rho[s].T
            den[w] += np.conj(c) * ov
    return num / den[:, None, None, None], den


@pytest.mark.parametrize("uhf", [False, True])
def test_mixed_density_matrix_and_vbias_match_brute_force(uhf):
    rng = np.random.default_rng(70 + uhf)
    s = ph_system(rng, 8, 3, 2, 6, 30, 2, uhf=uhf)
    nw = 4
    Phi_up, Phi_dn = rand_c(rng, nw, 8, 3), rand_c(rng, nw, 8, 2)
    Gf, logov = P.mixed_density_matrix(s["struct"], s["A"], Phi_up, Phi_dn, compact=False)
    Gb, ovb = brute_force_density(s, Phi_up, Phi_dn)
    np.testing.assert_allclose(Gf, Gb, rtol=1e-10, atol=1e-11 * np.abs(Gb).max())
    np.testing.assert_allclose(np.exp(logov), ovb, rtol=1e-10)
    # the compact form carries the same G through the orbital rows (A is square and unitary)
    Gc, _ = P.mixed_density_matrix(s["struct"], s["A"], Phi_up, Phi_dn)
    np.testing.assert_allclose(np.einsum("al,wak->wlk", s["A"][0], Gc[:, :8]), Gb[:, 0],
                               rtol=1e-10, atol=1e-11 * np.abs(Gb).max())
    # vbias: sqrt(dt) sum_s tr(L_n G_s^T) over the brute-force density
    dt = 0.01
    v = P.vbias(s["struct"], s["A"], s["Lna"], dt, Phi_up, Phi_dn)
    vb = np.sqrt(dt) * np.einsum("nlk,wslk->wn", s["L"], Gb)
    np.testing.assert_allclose(v, vb, rtol=1e-10, atol=1e-11 * np.abs(vb).max())


def test_vmf_single_determinant_closed_form():
    """One determinant: vMF_n = sqrt(dt) sum_s sum_{a occupied} (A_s L_n A_s^H)[a][a]."""
    rng = np.random.default_rng(90)
    s = ph_system(rng, 8, 3, 2, 5, 1, 1, uhf=True)
    v = P.vmf(s["struct"], s["A"], s["Lna"], 0.01)
    ref = np.zeros(5, complex)
    for sp, ne in ((0, 3), (1, 2)):
        M = np.einsum("al,nlk,bk->nab", s["A"][sp], s["L"], s["A"][sp].conj())
        ref += np.einsum("naa->n", M[:, :ne, :ne])
    np.testing.assert_allclose(v, 0.1 * ref, rtol=1e-12, atol=1e-14)


def test_vmf_is_the_trial_density_when_no_pair_differs_by_one_orbital():
    """Orthonormal orbitals and determinants that pairwise differ by two or more orbitals (in the
    same spin, with the same other-spin part): <Psi|a+a|Psi> / <Psi|Psi> is then
    sum_D |c_D|^2 rho_D / sum_D |c_D|^2 exactly, and vMF its vbias."""
    rng = np.random.default_rng(91)
    nmo, nup, ndn = 8, 3, 2
    X = rng.standard_normal((nchol := 4, nmo, nmo))
    L = (X + X.transpose(0, 2, 1)) / 2
    Q, _ = np.linalg.qr(rng.standard_normal((nmo, nmo)))
    A = [Q.T.copy(), Q.T.copy()]
    occs = np.array([[0, 1, 2, 8, 9], [0, 4, 5, 8, 9], [3, 6, 7, 8, 9], [0, 1, 2, 12, 13]], np.int32)
    ci = np.array([0.9, 0.3 - 0.1j, -0.2j, 0.25])
    st = P.build_ph_struct(ci, occs, nmo, nup, ndn)
    Lna = [np.einsum("al,nlk->nak", A[s], L) for s in range(2)]
    v = P.vmf(st, A, Lna, 0.04)
    rho = np.zeros((2, nmo, nmo))
    for occ, c in zip(occs, ci):
        for sp, orbs in ((0, occ[:nup]), (1, occ[nup:] - nmo)):
            rho[sp] += abs(c) ** 2 * sum(np.outer(Q[:, o], Q[:, o]) for o in orbs)
    rho /= np.sum(np.abs(ci) ** 2)
    ref = 0.2 * np.einsum("nlk,slk->n", L, rho)
    np.testing.assert_allclose(v, ref, rtol=1e-12, atol=1e-13)


def test_vmf_single_excitation_pair_matches_the_exact_density():
    """Two determinants of real orthonormal orbitals differing by the up orbital 2 -> 5 (no occupied
    orbital between the two in the sorted order, so the fermionic sign is +1 and the reference's
    unsigned pair term is exact): vMF is the vbias of the exact trial density matrix
    (|c1|^2 rho1 + |c2|^2 rho2 + c1 c2 (phi2 phi5^T + phi5 phi2^T)) / (|c1|^2 + |c2|^2)."""
    rng = np.random.default_rng(92)
    nmo, nup, ndn = 7, 3, 2
    X = rng.standard_normal((4, nmo, nmo))
    L = (X + X.transpose(0, 2, 1)) / 2
    Q, _ = np.linalg.qr(rng.standard_normal((nmo, nmo)))
    A = [Q.T.copy(), Q.T.copy()]
    occs = np.array([[0, 1, 2, 7, 8], [0, 1, 5, 7, 8]], np.int32)
    c1, c2 = 0.9, -0.4
    st = P.build_ph_struct(np.array([c1, c2], complex), occs, nmo, nup, ndn)
    Lna = [np.einsum("al,nlk->nak", A[s], L) for s in range(2)]
    v = P.vmf(st, A, Lna, 1.0)
    P_ = lambda orbs: sum(np.outer(Q[:, o], Q[:, o]) for o in orbs)
    up = c1 ** 2 * P_([0, 1, 2]) + c2 ** 2 * P_([0, 1, 5]) + c1 * c2 * (np.outer(Q[:, 2], Q[:, 5]) + np.outer(Q[:, 5], Q[:, 2]))
    dn = (c1 ** 2 + c2 ** 2) * P_([0, 1])
    rho = np.stack([up, dn]) / (c1 ** 2 + c2 ** 2)
    ref = np.einsum("nlk,slk->n", L, rho)
    np.testing.assert_allclose(v, ref, rtol=1e-12, atol=1e-13)
