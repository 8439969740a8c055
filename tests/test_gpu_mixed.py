"""Reduced-precision forms of the Cholesky contractions against the FP64 oracle.

The mixed forms change only how vbias, vHS and the exchange contraction are computed; these
tests hold each form's energies (E1, EXX, EJ) and contractions to the error its arithmetic
allows, on the same seeded walker states as the oracle:

* FP32: operands rounded to FP32, FP32 MFMA accumulation. The rounding of G (u = 2^-24 of each
  entry) is shared by all Cholesky vectors and adds up coherently in EXX, so the energy error
  grows with the energy scale, as ~sqrt(N M) u over the shapes here (up to 64 u = 3.8e-6 at
  N M = 4096, measured up to 1.5e-6): bounded by 5e-6 * (1 + |E1| + |EXX| + |EJ|) per walker.
* BF16X3: three real products per complex product (Re Re, Im Im, (Re + Im)(Re + Im)), each
  operand split into hi + lo BF16 (~2^-16 relative), each real product hi hi' + hi lo' + lo hi'
  (the dropped lo lo' term is ~2^-16 of it), FP32 accumulation. The FP32 form's energy error is
  ~22 u (u = 2^-24) of the energy scale; with u = 2^-16 and the Gauss recombination
  Im = P3 - P1 - P2 adding up to ~3x, the bound here is 2e-3 * (1 + |E1| + |EXX| + |EJ|), and
  5e-4 of the largest vbias / vHS element for the contractions.
* INT8_SPLIT: row-scaled int8 slices (5 of G, 5 of Lnak, base 254) with exact int32
  accumulation: the truncation of the slices leaves ~1e-3 of the FP32 error, bounded here by
  1e-9 * (1 + |E1| + |EXX| + |EJ|) per walker.
"""
import numpy as np
import pytest

from oracle import oracle as O
from safire_amd import synth
from tests._helpers import rand_c

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402

# relative-to-energy-scale bound of each reduced form
FORMS = {"fp32": (A.FP32, 5e-6), "bf16x3": (A.BF16X3, 2e-3), "int8": (A.INT8_SPLIT, 1e-9)}
# relative bound of vbias / vHS (largest element) of each reduced form
CONTRACTION_BOUND = {"fp32": 1e-5, "bf16x3": 5e-4, "int8": 1e-5}


def states(s, W, amp, seed):
    """Walkers near the trial (amp small) to far from it (amp ~ 1: large mixed estimators)."""
    rng = np.random.default_rng(seed)
    return s.PsiT.conj().T[None] + amp * rand_c(rng, W, s.M, s.N)


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("shape", [(24, 6, 96), (40, 8, 200), (64, 10, 320), (64, 16, 320),
                                   (40, 20, 120), (64, 32, 96), (128, 24, 60), (128, 32, 40)])
def test_reduced_form_energy(form, shape):
    prec, bound = FORMS[form]
    M, N, K = shape
    s = synth.make_system(M, N, K, seed=M + N)
    W = 48
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.reserve(W)
    ham.set_precision(prec)
    R = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt).tensors()
    for amp in (0.05, 0.5, 2.0):
        Phi = states(s, W, amp, seed=int(amp * 100) + M)
        G, _ = O.mixed_density_matrix(s.PsiT, Phi)
        E = torch.zeros((W, 3), dtype=torch.complex128, device="cuda")
        ham.energy(A.to_device(G), E)
        torch.cuda.synchronize()
        Eg = E.cpu().numpy()
        Eref = O.energy(G, R["haj"], R["Lnak"], E0=0.0)
        scale = 1.0 + np.abs(Eref).sum(1)
        err = np.abs(Eg - Eref).max(1) / scale
        assert np.all(err < bound), (form, amp, float(err.max()))
        # E1 does not depend on the Cholesky vectors: it stays FP64 in every form
        np.testing.assert_allclose(Eg[:, 0], Eref[:, 0], rtol=1e-11, atol=1e-11)
    ham.close()


@pytest.mark.parametrize("form", list(FORMS))
def test_reduced_form_contractions(form):
    prec, bound = FORMS[form]
    s = synth.make_system(64, 16, 320, seed=11)
    W = 16
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.reserve(W)
    ham.set_precision(prec)
    R = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt).tensors()
    rng = np.random.default_rng(3)
    G, _ = O.mixed_density_matrix(s.PsiT, states(s, W, 0.3, 7))
    v = torch.zeros((W, s.K), dtype=torch.complex128, device="cuda")
    ham.vbias(A.to_device(G), v)
    torch.cuda.synchronize()
    vref = O.vbias(G, R["Lnak"], s.dt)
    # FP32 operands: error ~ 2^-24 sqrt(N M) relative to the row magnitude (int8 and FP32 forms:
    # FP32 GEMMs); BF16 hi/lo operands ~2^-16
    cb = CONTRACTION_BOUND[form]
    assert np.abs(v.cpu().numpy() - vref).max() <= cb * np.abs(vref).max()
    X = rand_c(rng, W, s.K)
    V = torch.zeros((W, s.M, s.M), dtype=torch.complex128, device="cuda")
    ham.vHS(A.to_device(X), V)
    torch.cuda.synchronize()
    Vref = O.vhs(X, s.Likn, s.dt)
    assert np.abs(V.cpu().numpy() - Vref).max() <= cb * np.abs(Vref).max()
    ham.close()


@pytest.mark.parametrize("prec,M,N", [(A.BF16X3, 136, 20), (A.INT8_SPLIT, 136, 20)])
def test_split_forms_reject_unsupported_shapes(prec, M, N):
    s = synth.make_system(M, N, 60, seed=2)        # M beyond 128
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    with pytest.raises(A.AfqError):
        ham.set_precision(prec)
    ham.set_precision(A.FP32)                       # other forms stay available
    ham.close()


def test_int8_split_edge_rows():
    """Zero rows of G (a walker with G = 0 after a singular overlap), a NaN walker, and rows
    spanning 200 orders of magnitude: zero stays zero, NaN stays NaN, scaled rows stay exact."""
    s = synth.make_system(64, 10, 320, seed=4)
    W = 6
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.reserve(W)
    G, _ = O.mixed_density_matrix(s.PsiT, states(s, W, 0.2, 9))
    G[1] = 0.0
    G[2, 3, 5] = np.nan
    G[3] *= 1e-100
    G[4] *= 1e100
    G[5, :, :] *= np.logspace(-8, 8, s.M)[None, :]
    E = {}
    for prec in (A.FP64, A.INT8_SPLIT):
        ham.set_precision(prec)
        Ed = torch.zeros((W, 3), dtype=torch.complex128, device="cuda")
        ham.energy(A.to_device(G), Ed)
        torch.cuda.synchronize()
        E[prec] = Ed.cpu().numpy()
    ham.close()
    e64, e8 = E[A.FP64], E[A.INT8_SPLIT]
    assert np.all(e8[1] == 0.0) and np.all(e64[1] == 0.0)
    assert np.isnan(e8[2, 1:]).all() and np.isnan(e64[2, 1:]).all()
    for w in (0, 3, 4, 5):
        scale = 1.0 + np.abs(e64[w]).sum()
        assert np.abs(e8[w] - e64[w]).max() < 1e-9 * scale, w


def test_int8_energy_gate_c3():
    """The int8 form at the full NMO=128 / 32+32 / Nchol=640 shape (Cholesky vectors decaying as
    exp(-n/64)) against the FP64 oracle on walker states of an FP64 run: the seeded trial state
    and the states after steps 0, 1, 2 and 4 (QR and population control after step 0; on this
    Hamiltonian the phaseless projection removes most walkers within a few steps, so the states
    are taken early). Every live walker's E_loc = E1 + EXX + EJ within 1e-5 Ha of the oracle's
    on the oracle's G; the FP64 form within 1e-10 of the energy scale."""
    s, _, _, ortho = synth.workload("c3")
    W = 48
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=0)
    wset = A.WalkerSet(ham, W)
    prop = A.Propagator(ham)
    prop.seed(wset)
    sums = prop.weighted_sums(wset).cpu().numpy()
    Eshift = float((sums[0] / sums[1]).real)
    xi = A.to_device(synth.fields(s.seed, 5, W, s.K))
    uniforms = np.random.default_rng([s.seed, 31]).random(5)
    states = [wset.host()]
    for t in range(5):
        prop.Propagate(wset, xi[t], Eshift)
        if t % ortho == 0:
            prop.Orthogonalize(wset)
        if t % 10 == 0:
            prop.population_control(wset, float(uniforms[t]))
        if t in (0, 1, 2, 4):
            states.append(wset.host())
    ham8 = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham8.reserve(W)
    ham8.set_precision(A.INT8_SPLIT)
    R = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt).tensors()
    P = s.M * s.N
    worst, nlive = 0.0, 0
    for k, st in enumerate(states):
        live = np.abs(st[:, P + A.P_WEIGHT]) > 0
        if not np.any(live):
            continue
        nlive += int(live.sum())
        Phi = st[:, :P].reshape(W, s.M, s.N)
        G, _ = O.mixed_density_matrix(s.PsiT, Phi)
        Eref = O.energy(G, R["haj"], R["Lnak"], E0=0.0)
        Gd = A.to_device(G)
        out = {}
        for key, h in (("fp64", ham), ("int8", ham8)):
            E = torch.zeros((W, 3), dtype=torch.complex128, device="cuda")
            h.energy(Gd, E)
            torch.cuda.synchronize()
            out[key] = E.cpu().numpy()
        scale = 1.0 + np.abs(Eref).sum(1)
        e64 = np.abs(out["fp64"] - Eref).max(1) / scale
        assert np.all(e64[live] < 1e-10), (k, float(e64[live].max()))
        dE = np.abs(out["int8"].sum(1) - Eref.sum(1))
        worst = max(worst, float(dE[live].max()))
        assert np.all(dE[live] < 1e-5), (k, float(dE[live].max()), float(scale[live].max()))
    assert nlive >= W
    print(f"int8 form at NMO=128/N=32/K=640: max |dE_loc| {worst:.3e} Ha over {nlive} live "
          f"walker states")
    ham8.close()
    ham.close()
