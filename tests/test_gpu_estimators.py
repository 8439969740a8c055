"""Estimator walker sums (afq_estimator_sums) and the run loop with the energy estimator on the
GPU against the oracle's restatement of EnergyEstimator / averagePseudoEnergy / run_afqmc."""
import numpy as np
import pytest

from oracle import oracle as O
from safire_amd import estimators as ES
from safire_amd import synth
from tests._helpers import rand_c, relerr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("W", [1, 7, 513])
@pytest.mark.parametrize("given", [False, True])
def test_estimator_sums_match_oracle(W, given):
    s = synth.make_system(6, 2, 4, seed=3)
    P, ldw = s.M * s.N, s.M * s.N + A.P_COUNT
    rng = np.random.default_rng(W + 10 * given)
    buf = rand_c(rng, W, ldw)
    buf[:, P + A.P_OVLP] *= 0.3
    buf[rng.random(W) < 0.2, P + A.P_WEIGHT] = 0.0
    E = rand_c(rng, W, 3) * 5 if given else None
    ov = rand_c(rng, W) * 0.3 if given else None
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    prop = A.Propagator(ham)
    wset = A.WalkerSet(ham, W, buf.copy())
    Ed, ovd = (dev(E), dev(ov)) if given else (None, None)
    got = prop.estimator_sums(wset, Ed, ovd).cpu().numpy()
    w = buf[:, P + A.P_WEIGHT]
    Eu = E if given else buf[:, P + A.P_E1:P + A.P_EJ + 1]
    ou = ov if given else buf[:, P + A.P_OVLP]
    ref = O.estimator_measure(w, Eu, ou, 0, s.dt)
    den = got[3]
    np.testing.assert_allclose(den, ref["TotalWeight"], rtol=1e-13)
    np.testing.assert_allclose(got[0:3] / den, [ref["OnebodyEnergy"], ref["ExchangeEnergy"],
                                                ref["CoulombEnergy"]], rtol=1e-12)
    np.testing.assert_allclose(got[4] / den, ref["Overlap"], rtol=1e-12)
    neff, phase, _ = ES.weight_statistics(got)
    assert neff == pytest.approx(ref["EffectiveNumWalkers"], rel=1e-13)
    assert phase == pytest.approx(ref["Phase"], rel=1e-13)
    pe = O.average_pseudo_energy(w, buf[:, P + A.P_PSEUDO_ELOC], s.dt)
    assert ES.average_pseudo_energy(got, s.dt) == pytest.approx(pe, rel=1e-12)
    e = O.average_energy(w, *[buf[:, P + c] for c in (A.P_E1, A.P_EXX, A.P_EJ)])
    assert ES.average_energy(got) == pytest.approx(e, rel=1e-12)
    ham.close()


@pytest.mark.parametrize("hybrid,neq", [(False, 8), (True, 0)])
def test_run_afqmc_with_energy_estimator_c1(hybrid, neq):
    """c1, 64 walkers, 24 steps, QR every 5, population control every 10, EnergyEstimator every
    2 steps in bins of 2. The local-energy form reads the energies the step left on the walkers
    and relaxes Eshift over 8 equilibration steps; the hybrid form recomputes the energies
    (EnergyEstimator.h:66-75) and keeps Eshift (its complex pseudo energy can make the growth
    estimator NaN, as std::log does in the reference). Growth energies, Eshift and every binned
    observable agree with the oracle's run of the same loop."""
    s, W, _, _ = synth.workload("c1")
    steps, every = 24, 2
    xi = synth.fields(s.seed, steps, W, s.K)
    u = np.random.default_rng(5).random(steps)
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=int(hybrid))
    prop = A.Propagator(ham)
    wset = A.WalkerSet(ham, W)
    prop.seed(wset)
    E0 = ES.average_energy(prop.estimator_sums(wset).cpu().numpy())
    meas = ES.Measurements(binsize=2)
    est = ES.EnergyEstimator(ham, measure_interval=every, walkers_carry_energy=not hybrid)
    fields = [dev(xi[t]) for t in range(steps)]
    Eshift, growth = ES.run_afqmc(prop, wset, fields, u, steps, E0, estimators=[est], meas=meas,
                                  equilibration_steps=neq)

    ref = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    buf = ref.new_walkers(W)
    ref.seed_properties(buf)
    P = s.M * s.N
    props = lambda p: buf[:, P + p]  # noqa: E731
    rE = O.average_energy(props(A.P_WEIGHT), props(A.P_E1), props(A.P_EXX), props(A.P_EJ))
    assert E0 == pytest.approx(rE, rel=1e-12)
    f = O.eshift_relaxation_factor(neq)
    rshift, rgrowth, samples = rE, [], {}
    for t in range(steps):
        ref.step(buf, xi[t], rshift, hybrid=hybrid)
        if t % 5 == 0:
            ref.orthogonalize(buf)
        if t % 10 == 0:
            O.population_control(buf, P, u[t])
        e = O.average_pseudo_energy(props(A.P_WEIGHT), props(A.P_PSEUDO_ELOC), s.dt)
        if t >= neq:
            if t % every == 0:
                if hybrid:
                    En, ov = ref.energy(buf)
                else:
                    En, ov = buf[:, P + A.P_E1:P + A.P_EJ + 1], props(A.P_OVLP)
                for k, v in O.estimator_measure(props(A.P_WEIGHT), En, ov, t, s.dt).items():
                    samples.setdefault(k, []).append(v)
        else:
            rshift += f * (e - rshift)
        rgrowth.append(e)
    assert Eshift == pytest.approx(rshift, rel=1e-10)
    rgrowth = np.array(rgrowth)
    assert np.array_equal(np.isnan(growth), np.isnan(rgrowth))
    ok = ~np.isnan(rgrowth)
    assert relerr(growth[ok], rgrowth[ok]) < 1e-10
    assert sorted(samples) == meas.names()
    for k, v in samples.items():
        rb = np.array(O.accumulator_bins(v, 2))
        gb = meas.bins(k)
        assert gb.shape == rb.shape, k
        assert relerr(gb, rb) < 1e-9, k
    ham.close()
