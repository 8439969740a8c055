"""Estimators and the run loop around the walker step, over the C ABI.

Mirrors the reference's measurement layer for the closed-shell single-determinant run:

* ``Accumulator`` / ``Measurements`` (Estimators/Accumulator.hpp:82-150, Measurements.hpp:53-90):
  samples averaged into bins of ``binsize``; complete bins are kept, the unfinished one is not
  reported. The reference appends complete bins to an HDF5 file (``results.h5``); h5py is not
  available here, so ``Measurements.bins(name)`` returns them as arrays instead.
* ``MeasurementOutput`` (Measurements.hpp:96-128): every observable is divided by the total
  weight sum_w w_w of the measurement.
* ``weight_statistics`` (Measurements.hpp:130-149): effective walker number
  (sum|w|)^2 / sum|w|^2, phase sum w / sum|w|, total weight sum w.
* ``EnergyEstimator.measure`` (EnergyEstimator.h:51-92): Energy, OnebodyEnergy, ExchangeEnergy,
  CoulombEnergy, Overlap, ProjectionTime, EffectiveNumWalkers, Phase, TotalWeight every
  ``measure_interval`` steps.
* ``average_pseudo_energy`` / ``average_energy`` (Drivers/average_energy.hpp:53-79): the growth
  estimator log(sum w exp(dt E_pseudo) / sum w) / dt and the weighted mixed energy.
* ``run_afqmc`` (Drivers/run_afqmc.cpp:69-91): Propagate, Orthogonalize every
  ``walker_ortho_interval`` steps, population control every ``population_control_interval``
  steps, then either the estimators (after equilibration) or the Eshift relaxation
  Eshift += f (averagePseudoEnergy - Eshift), with f = 1 - exp(-10 / equilibration_steps)
  (parameter_defaults.cpp:310-318).

Every walker sum is one reduction on the GPU (``afq_estimator_sums``); the host combines its nine
numbers, as the reference combines its reduced sums on the host.
"""
from __future__ import annotations

import math

import numpy as np

from . import afqmc as A

DEFAULT_MEASURE_INTERVAL = 1


class Accumulator:
    """Averages samples (scalars or arrays) into bins of ``binsize`` (Accumulator.hpp:82-102)."""

    def __init__(self, binsize: int = 1):
        if binsize <= 0:
            raise ValueError(f"binsize has to be positive, but it is {binsize}")
        self.binsize = int(binsize)
        self._bins: list = []
        self._filling = 0

    def add(self, sample):
        sample = np.asarray(sample)
        if not self._bins:
            self._bins.append(np.zeros_like(sample, dtype=np.result_type(sample, np.float64)))
        self._bins[-1] = self._bins[-1] + sample
        self._filling += 1
        if self._filling >= self.binsize:
            self._bins[-1] = self._bins[-1] / self.binsize
            self._bins.append(np.zeros_like(self._bins[-1]))
            self._filling = 0

    def full_bins(self) -> np.ndarray:
        """The complete bins, stacked along a new first axis (Accumulator.hpp:12-33)."""
        if not self._bins:
            return np.zeros((0,))
        full = self._bins[:-1]
        if not full:
            return np.zeros((0,) + self._bins[-1].shape, self._bins[-1].dtype)
        return np.stack(full)


class Measurements:
    """Named observables, each an Accumulator of one sample type (Measurements.hpp:53-90)."""

    def __init__(self, prefix: str = "", binsize: int = 1):
        if binsize <= 0:
            raise ValueError(f"binsize has to be positive, but it is {binsize}")
        self.prefix = prefix
        self.binsize = int(binsize)
        self.observables: dict[str, Accumulator] = {}

    def measure(self, name: str, sample):
        acc = self.observables.setdefault(name, Accumulator(self.binsize))
        acc.add(sample)

    def bins(self, name: str) -> np.ndarray:
        return self.observables[name].full_bins()

    def names(self):
        return sorted(self.observables)


class MeasurementOutput:
    """Divides every observable by the total weight of the measurement (Measurements.hpp:96-128)."""

    def __init__(self, meas: Measurements, prefix: str, total_weight: complex):
        self.meas = meas
        self.prefix = f"{prefix}/" if prefix else ""
        self.denominator = complex(total_weight)

    def measure(self, name: str, sample):
        self.meas.measure(self.prefix + name, np.asarray(sample) / self.denominator)


def weight_statistics(sums: np.ndarray):
    """(effective number of walkers, phase, total weight) from the estimator sums
    (Measurements.hpp:130-149)."""
    s_w, s_abs, s_abs2 = complex(sums[3]), float(sums[5].real), float(sums[6].real)
    return (s_abs * s_abs / s_abs2).real if s_abs2 else float("nan"), s_w / s_abs, s_w


def average_pseudo_energy(sums: np.ndarray, dt: float) -> float:
    """Growth estimator log(sum w exp(dt E_pseudo) / sum w) / dt (average_energy.hpp:53-63).
    The hybrid pseudo energy is complex, so the average can have a non-positive real part; the
    reference's std::log then gives NaN (-inf at 0), and so does this."""
    x = (complex(sums[7]) / complex(sums[3])).real
    if x > 0:
        return math.log(x) / dt
    return -math.inf if x == 0 else math.nan


def average_energy(sums: np.ndarray) -> float:
    """sum w (E1+EXX+EJ) / sum w over the walker slots (average_energy.hpp:67-79)."""
    return (complex(sums[8]) / complex(sums[3])).real


class EnergyEstimator:
    """EnergyEstimator::measure (EnergyEstimator.h:51-92) over a GPU walker set.

    walkers_carry_energy: the propagator evaluated the local energy of every walker in the step
    (the local-energy form), so E1/EXX/EJ/OVLP are read from the walker slots; otherwise the
    energy is recomputed from the walkers (hybrid form)."""

    def __init__(self, ham: "A.HamiltonianOperations", measure_interval: int = DEFAULT_MEASURE_INTERVAL,
                 walkers_carry_energy: bool = True, timestep: float | None = None):
        if measure_interval <= 0:
            raise ValueError(f"'measure_interval' must be positive, but it is {measure_interval}.")
        self.ham = ham
        self.prop = A.Propagator(ham)
        self.measure_interval = int(measure_interval)
        self.walkers_carry_energy = bool(walkers_carry_energy)
        self.timestep = float(ham.dt if timestep is None else timestep)

    def sums(self, wset: "A.WalkerSet") -> np.ndarray:
        torch = A._torch()
        E = ov = None
        if not self.walkers_carry_energy:
            E = torch.zeros((wset.W, 3), dtype=torch.complex128, device="cuda")
            ov = torch.zeros(wset.W, dtype=torch.complex128, device="cuda")
            A.check(A.lib().afq_wfn_energy(self.ham.handle, wset.W, A.ptr(wset.buf), wset.ldw,
                                           A.ptr(E), A.ptr(ov), A.stream_ptr()))
        return self.prop.estimator_sums(wset, E, ov).cpu().numpy()

    def measure(self, step: int, meas: Measurements, wset: "A.WalkerSet"):
        if step % self.measure_interval != 0:
            return
        self.observe(step, meas, self.sums(wset))

    def observe(self, step: int, meas: Measurements, s: np.ndarray):
        """The host side of one measurement, from the nine walker sums."""
        out = MeasurementOutput(meas, "", s[3])
        avg = s[0:3]
        out.measure("Energy", avg.sum())
        out.measure("OnebodyEnergy", avg[0])
        out.measure("ExchangeEnergy", avg[1])
        out.measure("CoulombEnergy", avg[2])
        out.measure("Overlap", s[4])
        neff, phase, total = weight_statistics(s)
        meas.measure("ProjectionTime", (step + 1) * self.timestep)
        meas.measure("EffectiveNumWalkers", neff)
        meas.measure("Phase", phase)
        meas.measure("TotalWeight", total)


def eshift_relaxation_factor(equilibration_steps: int) -> float:
    """parameter_defaults.cpp:310-318."""
    if equilibration_steps > 1:
        return 1.0 - math.exp(-10.0 / equilibration_steps)
    return 1.0


def run_afqmc(prop: "A.Propagator", wset: "A.WalkerSet", fields, uniforms, steps: int,
              Eshift: float, estimators=(), meas: Measurements | None = None,
              walker_ortho_interval: int = 5, population_control_interval: int = 10,
              equilibration_steps: int = 0, relaxation_factor: float | None = None):
    """The loop of run_afqmc (run_afqmc.cpp:69-91) on one GPU. fields[t] (or fields(t) when
    fields is callable) is step t's [W, K] field array on the device, uniforms[t] (or
    uniforms(t)) the comb's uniform for step t. Returns (Eshift, growth energies per step)."""
    f = eshift_relaxation_factor(equilibration_steps) if relaxation_factor is None else relaxation_factor
    dt = prop.ham.dt
    growth = []
    for step in range(steps):
        prop.Propagate(wset, fields(step) if callable(fields) else fields[step], Eshift)
        if step % walker_ortho_interval == 0:
            prop.Orthogonalize(wset)
        if step % population_control_interval == 0:
            # a callable draws its uniform only here, one per population control, as the
            # reference's walker generator does
            prop.population_control(wset, float(uniforms(step) if callable(uniforms)
                                                 else uniforms[step]))
        sums = prop.estimator_sums(wset).cpu().numpy()
        e = average_pseudo_energy(sums, dt)
        if step >= equilibration_steps:
            for est in estimators:
                est.measure(step, meas, wset)
        else:
            Eshift += f * (e - Eshift)
        growth.append(e)
    return Eshift, np.array(growth)
