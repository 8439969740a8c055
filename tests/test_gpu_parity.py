"""GPU parity of the HIP walker step against the FP64 oracle, through the C ABI.

Plain FP64 form: every quantity within 1e-10 relative of the oracle on the same seeded inputs
(G, log-overlap, vbias, vHS, E1/EXX/EJ, propagated walkers, QR, whole steps).
"""
import numpy as np
import pytest

from oracle import oracle as O
from safire_amd import synth
from tests._helpers import cplx, load_json, rand_c, relerr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402

TOL = 1e-10


def dev(a, dtype=None):
    return A.to_device(a, dtype)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make(name_or_shape, seed=0):
    if isinstance(name_or_shape, str):
        s, W, steps, ortho = synth.workload(name_or_shape)
        return s, W
    M, N, K, W = name_or_shape
    return synth.make_system(M, N, K, seed=seed), W


@pytest.mark.parametrize("kind", [0, 1])
def test_mfma_lane_layout(kind):
    """The accumulator row maps the kernels use (acc_row in afq_internal.hpp)."""
    out = torch.zeros(64 * 4, dtype=torch.float64, device="cuda")
    A.check(A.lib().afq_probe_mfma(kind, A.ptr(out), A.stream_ptr()))
    d = host(out).reshape(64, 4)
    for lane in range(64):
        g, j = lane // 16, lane % 16
        for r in range(4):
            i = 4 * r + g if kind == 0 else 4 * g + r
            assert d[lane, r] == i + 100 * (j % 4), (lane, r, d[lane])


@pytest.mark.parametrize("kind", [2, 3, 4, 5])
def test_mfma_bf16_lane_layout(kind):
    """BF16 16x16x32 and int8 16x16x64: A and B pair their k slots lane group by lane group,
    the accumulator follows the FP32 layout (rows 4g..4g+3)."""
    out = torch.zeros(64 * 4, dtype=torch.float64, device="cuda")
    A.check(A.lib().afq_probe_mfma(kind, A.ptr(out), A.stream_ptr()))
    d = host(out).reshape(64, 4)
    for lane in range(64):
        g, j = lane // 16, lane % 16
        for r in range(4):
            i = 4 * g + r
            assert d[lane, r] == (i + 1) * (j + 1), (lane, r, d[lane])


def test_det_ops_kat_on_gpu():
    K = load_json("sdet_ops_kat.json")
    Am = cplx(K["A"])
    B = np.broadcast_to(cplx(K["B"]), (5, 4, 3)).copy()
    G = torch.zeros((5, 3, 4), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(5, dtype=torch.complex128, device="cuda")
    A.det_ops.MixedDensityMatrix(dev(Am), dev(B), G, ov)
    np.testing.assert_allclose(host(G), np.broadcast_to(cplx(K["gc_ref"]), (5, 3, 4)), atol=1e-8)
    np.testing.assert_allclose(host(ov), np.log(cplx(K["ov"])[0]), atol=1e-8)
    # sub-range (test_sdet_ops.cpp:121-131): A[:2,:3], B[:, :3, :2]
    ov2 = torch.zeros(5, dtype=torch.complex128, device="cuda")
    A.det_ops.Log_Overlap(dev(Am[:2, :3].copy()), dev(B[:, :3, :2].copy()), ov2)
    np.testing.assert_allclose(host(ov2), np.log(cplx(K["ov2"])[0]), atol=1e-8)


@pytest.mark.parametrize("M,N,nb", [(4, 3, 5), (24, 6, 33), (64, 16, 70), (128, 32, 9), (5, 1, 3)])
def test_mixed_density_matrix_parity(M, N, nb):
    rng = np.random.default_rng(M * 100 + N)
    PsiT = rand_c(rng, N, M)
    Phi = rand_c(rng, nb, M, N)
    Gref, ovref = O.mixed_density_matrix(PsiT, Phi)
    G = torch.zeros((nb, N, M), dtype=torch.complex128, device="cuda")
    ov = dev(np.full(nb, 0.25 + 0.5j))   # accumulates, like the reference
    A.det_ops.MixedDensityMatrix(dev(PsiT), dev(Phi), G, ov)
    assert relerr(host(G), Gref) < TOL
    ovexp = ovref + 0.25 + 0.5j
    ovexp.imag = (ovexp.imag + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(host(ov), ovexp, rtol=TOL, atol=1e-12)


def test_singular_overlap_gives_zero_G():
    PsiT = np.array([[1, 0, 0], [1, 0, 0]], np.complex128)
    Phi = np.array([[[1, 0], [0, 1], [0, 0]]], np.complex128)
    G = torch.ones((1, 2, 3), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(1, dtype=torch.complex128, device="cuda")
    A.det_ops.MixedDensityMatrix(dev(PsiT), dev(Phi), G, ov)
    _, ovref = O.mixed_density_matrix(PsiT, Phi)
    assert np.all(host(G) == 0)
    np.testing.assert_allclose(host(ov), ovref, rtol=1e-14)


@pytest.mark.parametrize("M,N", [(24, 5), (64, 16), (40, 12), (40, 17), (48, 24), (128, 32)])
def test_mixed_density_matrix_pivoting(M, N):
    """Overlap matrices that force row exchanges and ties in the pivot search: permutation
    matrices (one candidate per column, a swap at most steps) and small-integer matrices, whose
    |re|+|im| values tie often. Singular integer draws are kept: G is then zero, like getri_or_zero."""
    rng = np.random.default_rng(7 * M + N)
    nb = 24
    PsiT = np.zeros((N, M), np.complex128)
    PsiT[:, :N] = np.eye(N)
    Phi = rand_c(rng, nb, M, N)
    for w in range(nb):
        if w % 2 == 0:
            Phi[w, :N] = np.eye(N)[rng.permutation(N)] * (1 + 1j * (w % 4 == 0))
        else:
            Phi[w, :N] = rng.integers(-1, 2, (N, N)) + 1j * rng.integers(-1, 2, (N, N))
    Phi[2, 1] = Phi[2, 0]     # a repeated row: an exactly zero pivot part-way through
    Gref, ovref = O.mixed_density_matrix(PsiT, Phi)
    assert not np.any(Gref[2])
    G = torch.zeros((nb, N, M), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(nb, dtype=torch.complex128, device="cuda")
    A.det_ops.MixedDensityMatrix(dev(PsiT), dev(Phi), G, ov)
    Gh, ovh = host(G), host(ov)
    for w in range(nb):
        if not np.any(Gref[w]):
            assert not np.any(Gh[w]), w
        else:
            assert relerr(Gh[w], Gref[w]) < 1e-9, w
        d = ovh[w] - ovref[w]
        d = d.real + 1j * ((d.imag + np.pi) % (2 * np.pi) - np.pi)
        assert abs(d) < 1e-9 * max(1.0, abs(ovref[w])), (w, ovh[w], ovref[w])


@pytest.mark.parametrize("M,N,nb", [(8, 3, 2), (24, 6, 17), (64, 16, 5), (128, 32, 3)])
def test_propagate_parity(M, N, nb):
    rng = np.random.default_rng(7 + M)
    Phi = rand_c(rng, nb, M, N)
    P1 = rand_c(rng, M, M) * (0.5 / np.sqrt(M))
    V = 0.05 * rand_c(rng, nb, M, M) / np.sqrt(M)
    ref = O.propagate(Phi, P1, V, order=6)
    t = dev(Phi)
    A.det_ops.Propagate(t, dev(P1), dev(V), 6)
    assert relerr(host(t), ref) < TOL


@pytest.mark.parametrize("M,N,nb", [(4, 3, 5), (24, 6, 9), (64, 16, 5), (128, 32, 2)])
def test_orthogonalize_parity(M, N, nb):
    rng = np.random.default_rng(11 + M)
    Phi = rand_c(rng, nb, M, N)
    ref, ldref = O.orthogonalize(Phi)
    t = dev(Phi)
    ld = torch.zeros(nb, dtype=torch.complex128, device="cuda")
    A.det_ops.orthogonalize(t, ld)
    assert relerr(host(t), ref) < TOL
    np.testing.assert_allclose(host(ld), ldref, rtol=TOL)


@pytest.mark.parametrize("shape", ["c1", (16, 5, 40, 3), (40, 10, 200, 7), (64, 16, 320, 9),
                                   (96, 24, 480, 5), (128, 32, 640, 2), "c3"])
def test_setup_and_contractions_parity(shape):
    s, W = make(shape, seed=3)
    W = min(W, 40)
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.3, s.dt)
    ham.reserve(W)
    ref = O.System(s.h, s.Likn, s.PsiT, 0.3, s.dt)
    T, R = ham.tensors(), ref.tensors()
    for k in ("haj", "Lnak", "v0", "vMF", "H1", "P1"):
        assert relerr(T[k], R[k]) < 1e-12, k
    rng = np.random.default_rng(5)
    Phi = (s.PsiT.conj().T[None] + 0.3 * rand_c(rng, W, s.M, s.N))
    G, _ = O.mixed_density_matrix(s.PsiT, Phi)
    Gd = dev(G)
    v = torch.zeros((W, s.K), dtype=torch.complex128, device="cuda")
    ham.vbias(Gd, v)
    vref = O.vbias(G, R["Lnak"], s.dt)
    assert relerr(host(v), vref) < TOL
    X = rand_c(rng, W, s.K)
    V = torch.zeros((W, s.M, s.M), dtype=torch.complex128, device="cuda")
    ham.vHS(dev(X), V)
    assert relerr(host(V), O.vhs(X, s.Likn, s.dt)) < TOL
    E = torch.zeros((W, 3), dtype=torch.complex128, device="cuda")
    ham.energy(Gd, E)
    Eref = O.energy(G, R["haj"], R["Lnak"], E0=0.3)
    Eg = host(E)
    for c in range(3):
        assert relerr(Eg[:, c], Eref[:, c]) < TOL, c
    ham.close()


def run_resynced(shape, steps, hybrid, ortho=5, seed=9, walkers=None):
    """Runs the oracle trajectory; before every step the GPU gets the oracle's walker buffer, so
    each step is compared on identical inputs. Returns per step and walker the relative error of
    the Slater matrix, the error of the properties relative to (1 + |value|), and kappa(PsiT Phi)
    of the walker before the step (the conditioning that amplifies rounding through G).
    walkers: the first `walkers` walkers of the workload (all by default)."""
    s, W = make(shape, seed=seed)
    W = W if walkers is None else walkers
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=int(hybrid))
    prop = A.Propagator(ham)
    ref = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    buf = ref.new_walkers(W)
    ref.seed_properties(buf)
    seeded = A.WalkerSet(ham, W, ref.new_walkers(W))
    prop.seed(seeded)
    P = s.M * s.N
    assert relerr(seeded.host()[:, P:], buf[:, P:]) < TOL
    Eshift = float(ref.energy(buf)[0].sum(1).real.mean())
    xi = synth.fields(s.seed, steps, W, s.K)
    for t in range(steps):
        Phi = buf[:, :P].reshape(W, s.M, s.N)
        kappa = np.array([np.linalg.cond(s.PsiT @ Phi[w]) for w in range(W)])
        wset = A.WalkerSet(ham, W, buf.copy())
        prop.Propagate(wset, dev(xi[t]), Eshift)
        ref.step(buf, xi[t], Eshift, hybrid=hybrid)
        if t % ortho == 0:
            prop.Orthogonalize(wset)
            ref.orthogonalize(buf)
        g = wset.host()
        e_sm = np.abs(g[:, :P] - buf[:, :P]).max(1) / np.abs(buf[:, :P]).max(1)
        # Properties are compared on the scale they are computed from: the energies against
        # S = 1 + |E1| + |EXX| + |EJ| (E_loc is their sum), the phase and weight, which rotate
        # and scale by exp(-dt E), against 1 + dt S; the other slots against 1 + |value|.
        S = 1 + np.abs(buf[:, P + A.P_E1]) + np.abs(buf[:, P + A.P_EXX]) + np.abs(buf[:, P + A.P_EJ])
        scale = 1 + np.abs(buf[:, P:])
        for slot in (A.P_PSEUDO_ELOC, A.P_E1, A.P_EXX, A.P_EJ):
            scale[:, slot] = np.maximum(scale[:, slot], S)
        for slot in (A.P_PHASE, A.P_WEIGHT):
            scale[:, slot] = np.maximum(scale[:, slot], 1 + s.dt * S)
        e_slot = np.abs(g[:, P:] - buf[:, P:]) / scale
        e_pr = e_slot.max(1)
        run_resynced.last = (g, buf, e_slot, P)
        # a walker whose weight is exactly 0 stays 0 for good (the phaseless projection killed
        # it); it is still propagated, its energies grow without bound, and only its weight
        # carries meaning: it must be 0 on both sides
        dead = buf[:, P + A.P_WEIGHT] == 0
        assert np.array_equal(dead, g[:, P + A.P_WEIGHT] == 0), t
        e_sm[dead] = 0.0
        e_pr[dead] = 0.0
        e_slot[dead] = 0.0
        yield e_sm, e_pr, kappa
    ham.close()


SLOTS = ["WEIGHT", "PHASE", "PSEUDO_ELOC", "E1", "EXX", "EJ", "OVLP", "LOGSCL_UP", "LOGSCL_DN",
         "IS_UNITARY", "THETA"]


def worst(e, kappa):
    w = int(np.argmax(e))
    msg = f"walker {w}: err {e[w]:.3e}, kappa {kappa[w]:.3e}"
    last = getattr(run_resynced, "last", None)
    if last is not None:
        g, buf, e_slot, P = last
        s = int(np.argmax(e_slot[w]))
        msg += (f"; worst slot {SLOTS[s]} gpu {g[w, P + s]} ref {buf[w, P + s]}; "
                f"E ref {buf[w, P + 3:P + 6]}")
    return msg


@pytest.mark.parametrize("hybrid", [True, False])
def test_walker_steps_parity_c1(hybrid):
    """c1 (M=24, N=6, K=96): 64 walkers, 50 steps, QR every 5 steps, each step within 1e-10."""
    for t, (e_sm, e_pr, kappa) in enumerate(run_resynced("c1", 50, hybrid)):
        assert e_sm.max() < TOL, (t, worst(e_sm, kappa))
        assert e_pr.max() < TOL, (t, worst(e_pr, kappa))


@pytest.mark.parametrize("hybrid", [True, False])
def test_walker_steps_parity_c2(hybrid):
    """c2 (M=64, N=16, K=320), 512 walkers, 12 steps. Rounding is amplified by the conditioning
    of each walker's overlap matrix; the bound is 1e-10 + 1e-13 kappa_w per walker."""
    for t, (e_sm, e_pr, kappa) in enumerate(run_resynced("c2", 12, hybrid)):
        bound = TOL + 1e-13 * kappa
        assert np.all(e_sm < bound), (t, worst(e_sm / bound, kappa))
        assert np.all(e_pr < bound), (t, worst(e_pr / bound, kappa))


@pytest.mark.parametrize("hybrid", [True, False])
def test_walker_steps_parity_c3(hybrid):
    """c3 (M=128, N=32, K=640, Cholesky vectors decaying as exp(-n/64)): 16 walkers, 5 steps
    with QR after step 0, each step from the oracle's state, within 1e-10 + 1e-13 kappa_w."""
    for t, (e_sm, e_pr, kappa) in enumerate(run_resynced("c3", 5, hybrid, walkers=16)):
        bound = TOL + 1e-13 * kappa
        assert np.all(e_sm < bound), (t, worst(e_sm / bound, kappa))
        assert np.all(e_pr < bound), (t, worst(e_pr / bound, kappa))


@pytest.mark.parametrize("W", [1, 7, 65])
def test_ragged_walker_counts(W):
    for t, (e_sm, e_pr, kappa) in enumerate(run_resynced((24, 6, 96, W), 3, False)):
        assert e_sm.max() < TOL, (t, worst(e_sm, kappa))
        assert e_pr.max() < TOL, (t, worst(e_pr, kappa))


def test_bound_statistics_match():
    """Local-energy and weight bound counters agree with the oracle over a free-running c1 run."""
    s, W = make("c1")
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=0)
    wset = A.WalkerSet(ham, W)
    prop = A.Propagator(ham)
    prop.seed(wset)
    ref = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    buf = ref.new_walkers(W)
    ref.seed_properties(buf)
    Eshift = float(ref.energy(buf)[0].sum(1).real.mean())
    xi = synth.fields(s.seed, 8, W, s.K)
    for t in range(8):
        prop.Propagate(wset, dev(xi[t]), Eshift)
        ref.step(buf, xi[t], Eshift, hybrid=False)
        if t % 5 == 0:
            prop.Orthogonalize(wset)
            ref.orthogonalize(buf)
    e_gpu, w_gpu, _ = ham.stats()
    e_ref, w_ref = ref.stats()
    assert np.array_equal(e_gpu, e_ref) and np.array_equal(w_gpu, w_ref)
    assert relerr(wset.host(), buf) < TOL
    ham.close()


@pytest.mark.parametrize("W", [1, 5, 64, 512, 1000, 3000])
def test_population_control_parity(W):
    """Comb resampling and single-rank placement are bit-identical to the oracle's restatement
    of population_control.hpp (integer offspring, copied rows, w/|w|)."""
    s = synth.make_system(6, 2, 4, seed=1)
    rng = np.random.default_rng(W)
    ldw = s.M * s.N + A.P_COUNT
    P = s.M * s.N
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    prop = A.Propagator(ham)
    for trial in range(4):
        buf = rand_c(rng, W, ldw)
        w = np.abs(rng.standard_normal(W)) * np.exp(1j * rng.uniform(0, 6, W))
        w[rng.random(W) < 0.3] = 0.0                      # dead walkers are never copied
        if trial % 2:
            w = np.round(np.abs(w) * 4) / 4               # ties: teeth on exact boundaries
        if not np.any(w):
            w[0] = 1.0
        buf[:, P + A.P_WEIGHT] = w
        r = rng.random()
        wset = A.WalkerSet(ham, W, buf.copy())
        prop.population_control(wset, r)
        ref = buf.copy()
        off, rc = O.population_control(ref, P, r)
        assert rc == 0 and off.sum() == W
        got = wset.host()
        # rows are copied bit for bit; w/|w| may differ in the last ulp (host vs device hypot)
        keep = np.ones(ldw, bool)
        keep[P + A.P_WEIGHT] = False
        assert np.array_equal(got[:, keep], ref[:, keep]), trial
        np.testing.assert_allclose(got[:, P + A.P_WEIGHT], ref[:, P + A.P_WEIGHT], rtol=1e-15,
                                   atol=0)
    # a collapsed population is left alone and counted
    buf[:, P + A.P_WEIGHT] = 0.0
    wset = A.WalkerSet(ham, W, buf.copy())
    prop.population_control(wset, 0.5)
    assert np.array_equal(wset.host(), buf)
    _, _, pc = ham.stats()
    assert pc[0] == 5 and pc[1] == 1
    ham.close()


def test_one_body_propagator_sparse_truncation_gpu():
    """The GPU's P1 drops what the reference's sparse P1 drops (|x| <= 1e-8, imaginary roundoff
    of a real P1; AFQMCBasePropagator.cpp:165-211): on two decoupled blocks of orbitals its
    off-block entries are exactly zero and it equals the oracle's P1."""
    rng = np.random.default_rng(13)
    M, N, K = 12, 3, 10
    blocks = [np.arange(0, 7), np.arange(7, M)]
    h = np.zeros((M, M))
    L = np.zeros((K, M, M))
    for b in blocks:
        X = rng.standard_normal((len(b), len(b)))
        h[np.ix_(b, b)] = 0.3 * (X + X.T)
        for n in range(K):
            Y = rng.standard_normal((len(b), len(b)))
            L[n][np.ix_(b, b)] = 0.4 * (Y + Y.T)
    Likn = np.ascontiguousarray(L.transpose(1, 2, 0))
    Q, _ = np.linalg.qr(rand_c(rng, M, M))
    PsiT = np.ascontiguousarray(Q[:, :N].conj().T)
    ham = A.HamiltonianOperations(h, Likn, PsiT, 0.0, 0.01)
    P1g = ham.tensors()["P1"]
    P1o = O.System(h, Likn, PsiT, 0.0, 0.01).tensors()["P1"]
    off = np.ones((M, M), bool)
    for b in blocks:
        off[np.ix_(b, b)] = False
    assert np.all(P1g[off] == 0.0) and np.all(P1o[off] == 0.0)
    assert np.all(P1g.imag == 0.0) == np.all(P1o.imag == 0.0)
    assert relerr(P1g, P1o) < 1e-12
    ham.close()
