"""Host side of the estimators and the run loop (safire_amd/estimators.py) against the oracle's
restatement of EnergyEstimator / Accumulator / average_energy and closed forms (no GPU)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from safire_amd import estimators as ES
from tests._helpers import rand_c


def sums_of(w, E, ov, pseudo, dt):
    """The nine walker sums of afq_estimator_sums, formed on the host."""
    return np.array([(E[:, 0] * w).sum(), (E[:, 1] * w).sum(), (E[:, 2] * w).sum(), w.sum(),
                     (np.exp(ov) * w).sum(), np.abs(w).sum(), (np.abs(w) ** 2).sum(),
                     (w * np.exp(dt * pseudo)).sum(), (w * E.sum(1)).sum()], np.complex128)


def test_accumulator_bins_closed_form():
    acc = ES.Accumulator(4)
    for x in range(14):                  # 3 complete bins, 2 samples in the open one
        acc.add(float(x))
    np.testing.assert_array_equal(acc.full_bins(), [1.5, 5.5, 9.5])
    with pytest.raises(ValueError):
        ES.Accumulator(0)


@pytest.mark.parametrize("binsize", [1, 2, 5])
def test_accumulator_matches_oracle(binsize):
    rng = np.random.default_rng(binsize)
    scal = rand_c(rng, 23)
    arrs = [rng.standard_normal((2, 3)) for _ in range(12)]
    for samples in (scal, arrs):
        acc = ES.Accumulator(binsize)
        for x in samples:
            acc.add(x)
        ref = O.accumulator_bins(samples, binsize)
        got = acc.full_bins()
        assert len(got) == len(ref)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("W", [1, 6, 40])
def test_measure_matches_oracle(W):
    rng = np.random.default_rng(W)
    dt = 0.005
    w = rand_c(rng, W)
    E = rand_c(rng, W, 3) * 10
    ov = rand_c(rng, W) * 0.5
    pseudo = rand_c(rng, W)
    meas = ES.Measurements()
    est = ES.EnergyEstimator.__new__(ES.EnergyEstimator)
    est.measure_interval, est.timestep, est.walkers_carry_energy = 1, dt, True
    est.observe(7, meas, sums_of(w, E, ov, pseudo, dt))
    ref = O.estimator_measure(w, E, ov, 7, dt)
    for name, val in ref.items():
        got = meas.observables[name]._bins[0]      # binsize 1: the first complete bin
        np.testing.assert_allclose(got, val, rtol=1e-13, atol=0, err_msg=name)
    s = sums_of(w, E, ov, pseudo, dt)
    assert math.isclose(ES.average_pseudo_energy(s, dt), O.average_pseudo_energy(w, pseudo, dt),
                        rel_tol=1e-12)
    assert math.isclose(ES.average_energy(s), O.average_energy(w, E[:, 0], E[:, 1], E[:, 2]),
                        rel_tol=1e-12)


def test_measure_closed_forms():
    """Equal positive weights: N_eff = W, phase 1, total W, energies the plain means; constant
    pseudo energy e: the growth estimator returns e."""
    W, dt = 8, 0.01
    w = np.full(W, 2.0 + 0j)
    E = np.tile([1.0, -3.0, 0.5], (W, 1)).astype(np.complex128)
    ov = np.zeros(W, np.complex128)
    pseudo = np.full(W, -1.75 + 0j)
    meas = ES.Measurements()
    est = ES.EnergyEstimator.__new__(ES.EnergyEstimator)
    est.measure_interval, est.timestep, est.walkers_carry_energy = 1, dt, True
    s = sums_of(w, E, ov, pseudo, dt)
    est.observe(0, meas, s)
    b = {k: meas.observables[k]._bins[0] for k in meas.names()}
    assert b["EffectiveNumWalkers"] == pytest.approx(W)
    assert b["Phase"] == pytest.approx(1.0)
    assert b["TotalWeight"] == pytest.approx(2.0 * W)
    assert b["Energy"] == pytest.approx(-1.5)
    assert b["Overlap"] == pytest.approx(1.0)
    assert b["ProjectionTime"] == pytest.approx(dt)
    assert ES.average_pseudo_energy(s, dt) == pytest.approx(-1.75, rel=1e-12)
    assert ES.average_energy(s) == pytest.approx(-1.5)


def test_eshift_relaxation_factor():
    assert ES.eshift_relaxation_factor(20) == pytest.approx(1 - math.exp(-0.5))
    assert ES.eshift_relaxation_factor(1) == 1.0
    assert ES.eshift_relaxation_factor(0) == 1.0
    for n in (0, 1, 2, 50):
        assert ES.eshift_relaxation_factor(n) == O.eshift_relaxation_factor(n)


class _Host:
    def __init__(self, a):
        self.a = a

    def cpu(self):
        return self

    def numpy(self):
        return self.a


class _FakeProp:
    """Records the run loop's calls; the walker sums give a constant pseudo energy e."""

    class _Ham:
        dt = 0.01

    def __init__(self, e):
        self.ham = self._Ham()
        self.calls = []
        self.e = e

    def Propagate(self, wset, xi, Eshift):
        self.calls.append(("P", xi, Eshift))

    def Orthogonalize(self, wset):
        self.calls.append(("O",))

    def population_control(self, wset, u):
        self.calls.append(("C", u))

    def estimator_sums(self, wset):
        s = np.zeros(9, np.complex128)
        s[3] = 4.0
        s[7] = 4.0 * np.exp(self.ham.dt * self.e)
        return _Host(s)


class _FakeEstimator:
    def __init__(self):
        self.steps = []

    def measure(self, step, meas, wset):
        self.steps.append(step)


def test_run_loop_schedule_and_relaxation():
    """run_afqmc.cpp:69-91: Propagate every step, Orthogonalize when step % 5 == 0, population
    control when step % 10 == 0, Eshift relaxed towards the growth energy during equilibration
    (Eshift_t = e + (E0 - e)(1 - f)^t for a constant growth energy e), estimators afterwards."""
    e, E0, neq, steps = -2.5, 1.0, 8, 23
    prop = _FakeProp(e)
    est = _FakeEstimator()
    fields = list(range(steps))
    Eshift, growth = ES.run_afqmc(prop, None, fields, np.arange(steps) / 100.0, steps, E0,
                                  estimators=[est], meas=ES.Measurements(),
                                  equilibration_steps=neq)
    f = ES.eshift_relaxation_factor(neq)
    props = [c for c in prop.calls if c[0] == "P"]
    assert [c[1] for c in props] == fields
    for t, c in enumerate(props):
        assert c[2] == pytest.approx(e + (E0 - e) * (1 - f) ** min(t, neq), rel=1e-12)
    assert sum(c[0] == "O" for c in prop.calls) == len(range(0, steps, 5))
    assert [c[1] for c in prop.calls if c[0] == "C"] == [0.0, 0.1, 0.2]
    assert est.steps == list(range(neq, steps))
    assert Eshift == pytest.approx(e + (E0 - e) * (1 - f) ** neq, rel=1e-12)
    np.testing.assert_allclose(growth, e, rtol=1e-12)
