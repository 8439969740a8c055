"""Whole-run check against the reference's statistical result for BH (closed-shell RHF trial,
dense Cholesky Hamiltonian, NOMSD): the reference's functional test
(tests/functional/statistical_references/BH/dense_rhf_basis_closed/rhf_nomsd/closed) measured
E = -25.165509 +- 0.000322 Ha with 96 ranks x 16 walkers, dt = 0.01, 48000 steps, 3000
equilibration steps, QR and population control every 10 steps, energy every 10 steps (hybrid
propagation, energies recomputed by the estimator). The inputs and the reference's means and
errors are fixtures converted from its HDF5 files (tests/golden/make_bh_fixtures.py).

The same run here: 2048 walkers on one GPU, in the FP64 form and in the int8-split form (FP32
vbias/vHS, int8-split exchange in the estimator's energies, FP64 elsewhere), Gaussian fields
from a seeded device
generator (the run is deterministic for a given seed; field-for-field parity with the
reference's generator is a separate mode). The mean of every energy observable must lie within
4 combined standard errors of the reference's; this run's error comes from 30 blocks of
measurements (1500 steps = 15 Ha^-1 of imaginary time each)."""
import json
import os

import numpy as np
import pytest

from safire_amd import estimators as ES

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bh_rhf_closed.npz")


@pytest.mark.parametrize("precision", [A.FP64, A.INT8_SPLIT])
def test_bh_rhf_closed_energy_matches_reference_statistics(precision):
    d = np.load(FIX)
    inp = json.loads(str(d["run"]))
    run = inp["execute"]
    ref = json.loads(str(d["stats"]))
    dt, steps, neq = run["timestep"], run["steps"], run["equilibration_steps"]
    W = 2048
    ham = A.HamiltonianOperations(d["hcore"], d["Likn"], d["PsiT"], float(d["E0"]), dt)
    ham.set_params(hybrid=1)
    ham.set_precision(precision)
    prop = A.Propagator(ham)
    wset = A.WalkerSet(ham, W)
    prop.seed(wset)
    E_trial = ES.average_energy(prop.estimator_sums(wset).cpu().numpy())
    # RHF energy (closed form); the int8 split's exchange is exact to ~2^-39 of the row maxima
    tol = 1e-10 if precision == A.FP64 else 1e-7
    assert E_trial == pytest.approx(-25.06369204663835, abs=tol)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(inp["seed"])
    xi = torch.empty((W, ham.K), dtype=torch.float64, device="cuda")

    def fields(t):
        return torch.randn((W, ham.K), dtype=torch.float64, device="cuda", generator=gen, out=xi)

    u = np.random.default_rng(inp["seed"]).random(steps)
    meas = ES.Measurements()
    est = ES.EnergyEstimator(ham, measure_interval=run["measure_interval"],
                             walkers_carry_energy=False)
    Eshift, growth = ES.run_afqmc(prop, wset, fields, u, steps, E_trial, estimators=[est],
                                  meas=meas, walker_ortho_interval=run["walker_ortho_interval"],
                                  population_control_interval=run["population_control_interval"],
                                  equilibration_steps=neq)
    assert np.all(np.isfinite(growth))
    nb = 30
    for obs in ("Energy", "OnebodyEnergy", "ExchangeEnergy", "CoulombEnergy"):
        x = meas.bins(obs).real
        blocks = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
        mean, err = blocks.mean(), blocks.std(ddof=1) / np.sqrt(nb)
        rmean, rerr = ref[f"{obs}_mean"][0], ref[f"{obs}_error"][0]
        assert err < 5e-3, (obs, err)
        assert abs(mean - rmean) < 4 * np.hypot(err, rerr), (obs, mean, err, rmean, rerr)
    ham.close()
