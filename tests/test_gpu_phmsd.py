"""PHMSD excited-determinant energies on the GPU against the FP64 oracle (which is pinned by the
reference's literal values, tests/test_oracle_phmsd.py), through the C ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._helpers import load_json, rand_c
from tests.test_oracle_phmsd import kat_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402


def dev(a, dtype=None):
    return A.to_device(a, dtype)


def gpu_2body(iexcit, refc, Twina, R, wgt, EX0=None, EJ0=None):
    nw, ndet, nex, nact = R.shape
    nelec, nchol = Twina.shape[1], Twina.shape[2]
    EX = dev(np.zeros(nw, complex) if EX0 is None else EX0)
    EJ = dev(np.zeros(nw, complex) if EJ0 is None else EJ0)
    KE = torch.zeros((ndet, nw, nchol), dtype=torch.complex128, device="cuda")
    ie = torch.from_numpy(np.ascontiguousarray(iexcit, np.int32)).to("cuda")
    rc = torch.from_numpy(np.ascontiguousarray(refc, np.int32)).to("cuda")
    # every device array stays referenced until the launch: a temporary freed while the
    # arguments are still being built could be handed to the next upload
    Td, Rd, Wd = dev(Twina), dev(R), dev(wgt)
    A.check(A.lib().afq_ph_excited_2body_energy(
        nw, ndet, nex, nact, nelec, nchol, A.ptr(ie), A.ptr(rc), A.ptr(Td), A.ptr(Rd),
        A.ptr(Wd), A.ptr(EX), A.ptr(EJ), A.ptr(KE), A.stream_ptr()))
    torch.cuda.synchronize()
    return EX.cpu().numpy(), EJ.cpu().numpy(), KE.cpu().numpy()


@pytest.mark.parametrize("nex", [1, 2, 3, 4, 5])
def test_ph_2body_literal_values_on_gpu(nex):
    ref = np.array(load_json("phmsd_2body_kat.json")["EX_EJ_by_nex"][nex - 1])
    inputs = kat_inputs(nex)
    EX, EJ, KE = gpu_2body(*inputs)
    np.testing.assert_allclose(EX[0], ref[0, 0] + 1j * ref[0, 1], rtol=1e-11)
    np.testing.assert_allclose(EJ[0], ref[1, 0] + 1j * ref[1, 1], rtol=1e-11)
    _, _, KEo = O.ph_excited_2body(*inputs)
    np.testing.assert_allclose(KE, KEo, rtol=1e-12, atol=1e-12 * np.abs(KEo).max())


@pytest.mark.parametrize("nw,ndet,nex,nelec,nact,nchol", [(3, 7, 1, 4, 9, 6), (5, 40, 3, 10, 24, 70),
                                                           (2, 300, 4, 10, 64, 300), (4, 20, 8, 12, 40, 33)])
def test_ph_2body_matches_oracle(nw, ndet, nex, nelec, nact, nchol):
    rng = np.random.default_rng(nex * 100 + ndet)
    iexcit = np.stack([np.concatenate([rng.choice(nelec, nex, replace=False),
                                       nelec + rng.choice(nact - nelec, nex, replace=False)])
                       for _ in range(ndet)]).astype(np.int32)
    refc = rng.permutation(nelec).astype(np.int32)              # any reference order
    Twina, R, wgt = rand_c(rng, nw, nelec, nchol, nact), rand_c(rng, nw, ndet, nex, nact), \
        rand_c(rng, ndet, nw)
    EX0, EJ0 = rand_c(rng, nw), rand_c(rng, nw)                  # both accumulate
    EXo, EJo, KEo = O.ph_excited_2body(iexcit, refc, Twina, R, wgt)
    EX, EJ, KE = gpu_2body(iexcit, refc, Twina, R, wgt, EX0, EJ0)
    scale = np.abs(EXo).max() + np.abs(EJo).max()
    np.testing.assert_allclose(EX, EX0 + EXo, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(EJ, EJ0 + EJo, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(KE, KEo, rtol=0, atol=1e-12 * np.abs(KEo).max())


def test_ph_2body_bad_index_gives_nan():
    iexcit, refc, Twina, R, wgt = kat_inputs(2, ndet=6)
    iexcit = iexcit.copy()
    iexcit[3, 2] = 999                                           # orbital beyond nact
    EX, EJ, _ = gpu_2body(iexcit, refc, Twina, R, wgt)
    assert np.isnan(EX[0])


def test_ph_1body_matches_oracle():
    rng = np.random.default_rng(4)
    nw, ndet, nex, nelec, nact = 3, 500, 3, 10, 40
    iexcit = np.stack([np.concatenate([rng.choice(nelec, nex, replace=False),
                                       nelec + rng.choice(nact - nelec, nex, replace=False)])
                       for _ in range(ndet)]).astype(np.int32)
    refc = np.arange(nelec, dtype=np.int32)
    S, R, wgt = rand_c(rng, nw, nelec, nact), rand_c(rng, nw, ndet, nex, nact), rand_c(rng, ndet, nw)
    Eo = O.ph_excited_1body(iexcit, refc, S, R, wgt)
    E = dev(np.zeros(nw, complex))
    ie = torch.from_numpy(iexcit).to("cuda")
    rc = torch.from_numpy(refc).to("cuda")
    Sd, Rd, Wd = dev(S), dev(R), dev(wgt)
    A.check(A.lib().afq_ph_excited_1body_energy(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(rc),
                                                A.ptr(Sd), A.ptr(Rd), A.ptr(Wd), A.ptr(E),
                                                A.stream_ptr()))
    torch.cuda.synchronize()
    np.testing.assert_allclose(E.cpu().numpy(), Eo, rtol=1e-12)
