"""Pins the oracle's Real3IndexFactorization / RealDenseHamiltonian / propagator restatement by
routes independent of the half-rotated contractions: Wick's theorem on the explicit
two-electron integrals v_ijkl = sum_n L^n_ij L^n_kl, closed forms on the 4x4 Hubbard Cholesky
vectors of tests/hubbard_factorizations.hpp:332-386 (L_n = sqrt(U) e_n e_n^T), scipy's expm,
and special cases of the walker update that fix each term."""
import numpy as np
import scipy.linalg

from oracle import oracle as O
from tests._helpers import rand_c, relerr, tiny_system


def full_G(PsiT, Phi):
    T = PsiT @ Phi
    return (Phi @ np.linalg.inv(T) @ PsiT).T   # G[i][j] = <c+_i c_j>


def wick_energy_closed(h, Likn, G):
    """Closed-shell mixed energy by Wick's theorem with both spins equal:
    E1 = 2 sum h_ij G_ij; E2 = 1/2 sum v_ijkl [4 G_ij G_kl - 2 G_il G_kj]."""
    v = np.einsum("ijn,kln->ijkl", Likn, Likn)
    E1 = 2 * np.einsum("ij,ij", h, G)
    EJ = 2 * np.einsum("ijkl,ij,kl", v, G, G)
    EXX = -np.einsum("ijkl,il,kj", v, G, G)
    return E1, EXX, EJ


def test_energy_vs_wick_bruteforce():
    rng = np.random.default_rng(21)
    for M, N, K in [(5, 2, 4), (7, 3, 6)]:
        h, Likn, PsiT = tiny_system(rng, M, N, K)
        haj, Lnak = O.half_rotate(h, Likn, PsiT)
        Phi = rand_c(rng, 3, M, N)
        Gc, _ = O.mixed_density_matrix(PsiT, Phi, compact=True)
        E = O.energy(Gc, haj, Lnak, E0=0.25, walker_type=1)
        for w in range(3):
            E1, EXX, EJ = wick_energy_closed(h, Likn, full_G(PsiT, Phi[w]))
            np.testing.assert_allclose(E[w, 0], 0.25 + E1, rtol=1e-12)
            np.testing.assert_allclose(E[w, 1], EXX, rtol=1e-12)
            np.testing.assert_allclose(E[w, 2], EJ, rtol=1e-12)


def test_energy_collinear_vs_wick():
    rng = np.random.default_rng(22)
    M, nup, ndn, K = 6, 3, 2, 5
    h, Likn, PsiTu = tiny_system(rng, M, nup, K)
    Q, _ = np.linalg.qr(rand_c(rng, M, M))
    PsiTd = np.ascontiguousarray(Q[:, :ndn].conj().T)
    haju, Lu = O.half_rotate(h, Likn, PsiTu)
    hajd, Ld = O.half_rotate(h, Likn, PsiTd)
    Phiu, Phid = rand_c(rng, 2, M, nup), rand_c(rng, 2, M, ndn)
    Gu, _ = O.mixed_density_matrix(PsiTu, Phiu)
    Gd, _ = O.mixed_density_matrix(PsiTd, Phid)
    G = np.concatenate([Gu, Gd], axis=1)
    E = O.energy(G, haju, Lu, walker_type=2, ndn=ndn, haj_dn=hajd, Lnak_dn=Ld)
    v = np.einsum("ijn,kln->ijkl", Likn, Likn)
    for w in range(2):
        Ga, Gb = full_G(PsiTu, Phiu[w]), full_G(PsiTd, Phid[w])
        E1 = np.einsum("ij,ij", h, Ga + Gb)
        EJ = 0.5 * np.einsum("ijkl,ij,kl", v, Ga + Gb, Ga + Gb)
        EXX = -0.5 * (np.einsum("ijkl,il,kj", v, Ga, Ga) + np.einsum("ijkl,il,kj", v, Gb, Gb))
        np.testing.assert_allclose(E[w], [E1, EXX, EJ], rtol=1e-12)


def test_vbias_vs_full_G():
    rng = np.random.default_rng(23)
    M, N, K, dt = 7, 3, 6, 0.01
    h, Likn, PsiT = tiny_system(rng, M, N, K)
    _, Lnak = O.half_rotate(h, Likn, PsiT)
    Phi = rand_c(rng, 4, M, N)
    Gc, _ = O.mixed_density_matrix(PsiT, Phi)
    v = O.vbias(Gc, Lnak, dt, walker_type=1)
    for w in range(4):
        ref = 2 * np.sqrt(dt) * np.einsum("ijn,ij->n", Likn, full_G(PsiT, Phi[w]))
        assert relerr(v[w], ref) < 1e-12


def hubbard_likn(U=4.0, M=4):
    # tests/hubbard_factorizations.hpp:332-386: L_n[i][j] = sqrt(U) delta_ni delta_nj
    L = np.zeros((M, M, M))
    for n in range(M):
        L[n, n, n] = np.sqrt(U)
    return L


def test_hubbard_closed_forms():
    rng = np.random.default_rng(24)
    M, N, U, dt = 4, 2, 4.0, 0.01
    Likn = hubbard_likn(U, M)
    h = np.zeros((M, M))
    Q, _ = np.linalg.qr(rand_c(rng, M, M))
    PsiT = np.ascontiguousarray(Q[:, :N].conj().T)
    haj, Lnak = O.half_rotate(h, Likn, PsiT)
    Phi = rand_c(rng, 2, M, N)
    Gc, _ = O.mixed_density_matrix(PsiT, Phi)
    v = O.vbias(Gc, Lnak, dt, walker_type=1)
    E = O.energy(Gc, haj, Lnak, walker_type=1)
    for w in range(2):
        G = full_G(PsiT, Phi[w])
        d = np.diag(G)
        np.testing.assert_allclose(v[w], 2 * np.sqrt(dt) * np.sqrt(U) * d, rtol=1e-12)
        # on-site U sum_i n_i^2 split: EJ = 2 U sum_i G_ii^2, EXX = -U sum_i G_ii^2
        np.testing.assert_allclose(E[w, 2], 2 * U * np.sum(d * d), rtol=1e-12)
        np.testing.assert_allclose(E[w, 1], -U * np.sum(d * d), rtol=1e-12)
    # v0 = -1/2 sum_j v_ijjl: diagonal -U/2
    np.testing.assert_allclose(O.exchange_v0(Likn), -0.5 * U * np.eye(M), atol=1e-14)
    # vHS(vbias(G)) is basis independent: sum_n L_n v_n = 2 sqrt(dt) U diag(G) here
    V = O.vhs(v, Likn, dt)
    np.testing.assert_allclose(np.diag(V[0]), np.sqrt(dt) * np.sqrt(U) * v[0], rtol=1e-12)


def test_v0_and_vhs_vs_eri():
    rng = np.random.default_rng(25)
    M, K, dt = 6, 5, 0.02
    h, Likn, _ = tiny_system(rng, M, 2, K)
    v = np.einsum("ijn,kln->ijkl", Likn, Likn)
    np.testing.assert_allclose(O.exchange_v0(Likn), -0.5 * np.einsum("ijjl->il", v), rtol=1e-12)
    X = rand_c(rng, 3, K)
    V = O.vhs(X, Likn, dt)
    np.testing.assert_allclose(V, np.sqrt(dt) * np.einsum("ijn,wn->wij", Likn, X), rtol=1e-12)


def test_system_setup_P1_vs_expm():
    rng = np.random.default_rng(26)
    M, N, K, dt = 8, 3, 7, 0.005
    h, Likn, PsiT = tiny_system(rng, M, N, K)
    S = O.System(h, Likn, PsiT, 0.0, dt)
    T = S.tensors()
    # vMF = Re vbias(G_trial) with G_trial from Phi = PsiT^H
    Gt = full_G(PsiT, PsiT.conj().T)
    vmf = (2 * np.sqrt(dt) * np.einsum("ijn,ij->n", Likn, Gt)).real
    np.testing.assert_allclose(T["vMF"], vmf, rtol=1e-12)
    v0 = -0.5 * np.einsum("ijn,ljn->il", Likn, Likn)
    H1 = dt * (h + v0) + np.sqrt(dt) * np.einsum("ijn,n->ij", Likn, vmf)
    H1 = 0.5 * (H1 + H1.conj().T)
    np.testing.assert_allclose(T["P1"], scipy.linalg.expm(-0.5 * H1), rtol=1e-12, atol=1e-14)


def test_construct_X_special_cases():
    rng = np.random.default_rng(27)
    W, K = 3, 5
    vmf = rng.standard_normal(K) + 0j
    xi = rng.standard_normal((W, K))
    # vbias == vMF: no shift, no reweighting
    X, HW = O.construct_X(np.tile(vmf, (W, 1)), xi, vmf)
    np.testing.assert_allclose(X, xi, atol=0)
    np.testing.assert_allclose(HW, 0, atol=0)
    # a single shifted component: X = xi + i d, HW = -(i d)(X - i d/2) = d^2/2 - i d xi
    vb = np.tile(vmf, (W, 1))
    vb[:, 2] += 0.3
    X, HW = O.construct_X(vb, xi, vmf)
    np.testing.assert_allclose(X[:, 2], xi[:, 2] + 0.3j, rtol=1e-15)
    np.testing.assert_allclose(HW, 0.045 - 0.3j * xi[:, 2], rtol=1e-14)
    # clamp: |vbias| > bound is scaled onto the bound (construct_X_kernel.hpp:56)
    vb = np.zeros((1, K), complex)
    vb[0, 0] = 300 + 400j
    X, _ = O.construct_X(vb, np.zeros((1, K)), np.zeros(K, complex), vbias_bound=50.0)
    np.testing.assert_allclose(X[0, 0], 1j * (30 + 40j), rtol=1e-15)
    mf = O.meanfield_factor(X, vmf)
    np.testing.assert_allclose(mf[0], -0.5 * np.sum(vmf.real ** 2) + 1j * np.sum(X[0] * vmf))


def _one_walker_buf(M=2, N=1):
    buf = np.zeros((1, M * N + O.P_COUNT), complex)
    p = M * N
    buf[0, p + O.P_WEIGHT] = 1.0
    buf[0, p + O.P_PHASE] = 1.0
    return buf, p


def test_walker_update_phaseless_and_bounds():
    import ctypes as C
    L = O.lib()
    dt, E0, Es = 0.01, 0.5, -1.0
    buf, p = _one_walker_buf()
    buf[0, p + O.P_OVLP] = 0.2 + 0.1j
    buf[0, p + O.P_PSEUDO_ELOC] = -1.2
    new_ov = np.array([0.3 + 0.1j + np.pi / 3 * 1j])
    mf = np.array([0.05 + 0.0j])
    HW = np.array([0.01 + 0.0j])
    E = np.zeros((1, 3), complex)
    st = np.zeros(3, np.int64)
    wf = np.zeros(1, complex)      # back-propagation's weight factor of the step
    wn = np.zeros(1, complex)      # and the weight it records (after the update)
    L.orc_walker_update(1, buf.ctypes.data, buf.shape[1], p, 1, 0, 0, dt, Es, E0,
                        new_ov.ctypes.data, E.ctypes.data, mf.ctypes.data, HW.ctypes.data,
                        1.0, 10.0, st.ctypes.data, wf.ctypes.data, wn.ctypes.data)
    dth = np.pi / 3
    eloc = (mf[0] - HW[0] - (new_ov[0] - (0.2 + 0.1j))) / dt + E0
    np.testing.assert_allclose(buf[0, p + O.P_THETA], dth, rtol=1e-14)
    np.testing.assert_allclose(buf[0, p + O.P_PSEUDO_ELOC], eloc, rtol=1e-13)
    w = np.exp(-dt * (eloc - Es))
    np.testing.assert_allclose(buf[0, p + O.P_WEIGHT], np.cos(dth) * abs(w), rtol=1e-13)
    np.testing.assert_allclose(buf[0, p + O.P_PHASE], w / abs(w), rtol=1e-13)
    np.testing.assert_allclose(wf[0], w / abs(w) / np.cos(dth), rtol=1e-13)
    assert wn[0] == buf[0, p + O.P_WEIGHT]
    # |dtheta| > pi/2 kills the walker (max(0, cos))
    buf2, p = _one_walker_buf()
    new_ov = np.array([2.0j])
    L.orc_walker_update(1, buf2.ctypes.data, buf2.shape[1], p, 1, 0, 0, dt, Es, E0,
                        new_ov.ctypes.data, E.ctypes.data, np.zeros(1, complex).ctypes.data,
                        np.zeros(1, complex).ctypes.data, 1.0, 10.0, st.ctypes.data, wf.ctypes.data, wn.ctypes.data)
    assert buf2[0, p + O.P_WEIGHT] == 0
    assert wf[0] == 0 and wn[0] == 0       # scale 0: no factor
    # local-energy form: midpoint rule and the lower clamp at Eshift - sqrt(2/dt)
    buf3, p = _one_walker_buf()
    buf3[0, p + O.P_PSEUDO_ELOC] = -2.0
    E = np.array([[-50.0, -3.0, 1.0]], complex)
    st[:] = 0
    L.orc_walker_update(1, buf3.ctypes.data, buf3.shape[1], p, 0, 0, 0, dt, Es, E0,
                        np.zeros(1, complex).ctypes.data, E.ctypes.data,
                        np.zeros(1, complex).ctypes.data, np.zeros(1, complex).ctypes.data,
                        1.0, 10.0, st.ctypes.data, wf.ctypes.data, wn.ctypes.data)
    lo = Es - np.sqrt(2 / dt)
    np.testing.assert_allclose(buf3[0, p + O.P_PSEUDO_ELOC], lo, rtol=1e-14)
    np.testing.assert_allclose(buf3[0, p + O.P_WEIGHT], np.exp(-dt * (0.5 * (lo - 2.0) - Es)),
                               rtol=1e-13)
    assert list(st) == [1, 0, 1]
    np.testing.assert_allclose(buf3[0, p + O.P_E1], -50.0)
    # non-finite energy: weight 0, old energy kept
    buf4, p = _one_walker_buf()
    buf4[0, p + O.P_PSEUDO_ELOC] = -7.0
    E = np.array([[np.nan, 0, 0]], complex)
    L.orc_walker_update(1, buf4.ctypes.data, buf4.shape[1], p, 0, 0, 0, dt, Es, E0,
                        np.zeros(1, complex).ctypes.data, E.ctypes.data,
                        np.zeros(1, complex).ctypes.data, np.zeros(1, complex).ctypes.data,
                        1.0, 10.0, st.ctypes.data, wf.ctypes.data, wn.ctypes.data)
    assert buf4[0, p + O.P_WEIGHT] == 0 and buf4[0, p + O.P_PSEUDO_ELOC] == -7.0
    assert wf[0] == 0 and wn[0] == 0


def test_copy_plan_matches_population_control():
    """The source map of oracle.copy_plan moves rows exactly as the oracle's population control
    moves walker rows (the slot placement of swap_walkers_async)."""
    rng = np.random.default_rng(5)
    for W in (1, 7, 64, 301):
        buf = rng.standard_normal((W, 4 + O.P_COUNT)) + 1j * rng.standard_normal((W, 4 + O.P_COUNT))
        w = np.abs(rng.standard_normal(W))
        w[rng.random(W) < 0.3] = 0.0
        w[0] = max(w[0], 0.5)
        buf[:, 4 + O.P_WEIGHT] = w
        tag = np.arange(W, dtype=float)
        ref = buf.copy()
        off, rc = O.population_control(ref, 4, 0.37)
        assert rc == 0
        src = O.copy_plan(off)
        rows = buf.copy()
        O.apply_copy_plan(rows, src)
        O.apply_copy_plan(tag, src)
        keep = np.ones(buf.shape[1], bool)
        keep[4 + O.P_WEIGHT] = False
        assert np.array_equal(rows[:, keep], ref[:, keep])
        # every walker with k offspring appears k times
        assert np.array_equal(np.bincount(tag.astype(int), minlength=W), off)


def test_bound_weights():
    buf = np.zeros((3, 2 + O.P_COUNT), complex)
    buf[:, 2 + O.P_WEIGHT] = [1.0, 150.0 * np.exp(0.3j), 99.0]
    st = np.zeros(2, np.int64)
    O.lib().orc_bound_weights(3, buf.ctypes.data, buf.shape[1], 2, 100.0, 0.1, 3.0,
                              st.ctypes.data)
    np.testing.assert_allclose(buf[:, 2 + O.P_WEIGHT], [1.0, 100.0 * np.exp(0.3j), 99.0])
    assert list(st) == [3, 1]


def test_comb_closed_form():
    """population_control.hpp:34-65: teeth (ic + r) * total / n over the cumulative weights."""
    off, rc = O.comb([1.0, 0.0, 3.0], 0.5)
    # total 4, teeth 0.667, 2.0, 3.333 over cum (1, 1, 4)
    assert rc == 0 and list(off) == [1, 0, 2]
    off, _ = O.comb([0.25, 0.25, 0.25, 0.25], 0.999)
    assert list(off) == [1, 1, 1, 1]
    # the last tooth lands on the last walker with weight, not on a zero-weight one behind it
    off, _ = O.comb([1.0, 1.0, 0.0], 0.9999999999999999)
    assert list(off) == [1, 2, 0]
    assert O.comb([0.0, 1e-7], 0.3)[1] == 1          # collapsed total weight


def test_comb_is_unbiased():
    rng = np.random.default_rng(3)
    w = rng.random(50) ** 3
    acc = np.zeros(50)
    n = 4000
    for _ in range(n):
        off, _ = O.comb(w, rng.random())
        assert off.sum() == 50
        acc += off
    np.testing.assert_allclose(acc / n, 50 * w / w.sum(), atol=0.05)


def test_population_control_placement():
    # walkers 0..3 with offspring (0, 3, 1, 0): walker 1 keeps slot 1, its two extra copies go
    # to the zero-offspring slots 0 and 3 in order; weights become w/|w|
    W, P = 4, 2
    buf = np.zeros((W, P + O.P_COUNT), complex)
    buf[:, 0] = [10, 11, 12, 13]
    buf[:, P + O.P_WEIGHT] = [0.0, 2.0 * 1j, 0.9, 0.0]
    off, rc = O.population_control(buf, P, 0.1)
    # total 2.9, teeth (0.1, 1.1, 2.1, 3.1) * 2.9 / 4 over cum (0, 2, 2.9, 2.9)
    assert rc == 0 and list(off) == [0, 3, 1, 0]
    for slot, s in {0: 1, 1: 1, 2: 2, 3: 1}.items():
        assert buf[slot, 0] == 10 + s
    np.testing.assert_allclose(np.abs(buf[:, P + O.P_WEIGHT]), 1.0)


def test_one_body_propagator_sparse_truncation():
    """P1 as the reference's host path applies it (denseP1 = false): entries with |x| <= 1e-8
    dropped (to_csr, AFQMCBasePropagator.cpp:165-166) and a real P1 applied in real arithmetic
    (:174-211). Two decoupled blocks (block-diagonal h and Cholesky vectors) give a P1 whose
    off-block entries are roundoff: they must be exactly zero, the imaginary parts too, and the
    kept entries equal expm(-H1/2)."""
    import scipy.linalg
    rng = np.random.default_rng(12)
    M, K, dt = 10, 8, 0.01
    blocks = [np.arange(0, 6), np.arange(6, M)]
    h = np.zeros((M, M))
    L = np.zeros((K, M, M))
    for b in blocks:
        X = rng.standard_normal((len(b), len(b)))
        h[np.ix_(b, b)] = 0.3 * (X + X.T)
        for n in range(K):
            Y = rng.standard_normal((len(b), len(b)))
            L[n][np.ix_(b, b)] = 0.4 * (Y + Y.T)
    Likn = np.ascontiguousarray(L.transpose(1, 2, 0))
    vMF = rng.standard_normal(K) * 0.3 + 0j
    P1, H1 = O.one_body_propagator(h, Likn, vMF, dt)
    ref = scipy.linalg.expm(-0.5 * H1)
    off = np.ones((M, M), bool)
    for b in blocks:
        off[np.ix_(b, b)] = False
    assert np.abs(ref[off]).max() < 1e-8           # roundoff in the exact exponential
    assert np.all(P1[off] == 0.0)                   # dropped as the CSR conversion drops it
    assert np.all(P1.imag == 0.0)                   # real P1: real arithmetic
    keep = np.abs(ref.real) > 1e-8
    np.testing.assert_allclose(P1[keep].real, ref[keep].real, rtol=1e-12, atol=1e-14)
    assert np.all(P1[~keep] == 0.0)
