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


# ---- excitation overlaps and R matrices (calculate_overlaps, compact R, calculate_R) ----
from tests.test_oracle_phmsd_R import shells_for  # noqa: E402


def i32(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to("cuda")


@pytest.mark.parametrize("nex", [1, 2, 3, 4, 5, 8])
def test_ph_overlaps_match_oracle(nex):
    rng = np.random.default_rng(40 + nex)
    nw, nelec, nact, ndet = 33, 10, 24, 70
    T = rand_c(rng, nw, nact, nelec)
    ex = shells_for(rng, nelec, nact, ndet, nex)
    # a target orbital with an all-zero row of T: every excitation into it is exactly singular
    # (a zero pivot in any arithmetic)
    T[:, ex[3, nex], :] = 0
    ie, Td = i32(ex), dev(T)
    ov = torch.zeros((ndet, nw), dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_ph_overlaps(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(Td), A.ptr(ov),
                                    A.stream_ptr()))
    ref = O.ph_overlaps(ex, T)
    got = ov.cpu().numpy()
    assert np.all(ref[3] == 0) and np.all(got[3] == 0)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("nex", [1, 2, 3, 4, 5])
def test_ph_compact_R_matches_oracle(nex):
    rng = np.random.default_rng(50 + nex)
    nw, nelec, nact, ndet = 17, 10, 24, 40
    T = rand_c(rng, nw, nact, nelec)
    refc = rng.permutation(nact)[:nelec]
    ex = shells_for(rng, nelec, nact, ndet, nex)
    others = np.setdiff1d(np.arange(nact), refc)
    ex[:, nex:] = np.array([rng.choice(others, nex, replace=False) for _ in range(ndet)])
    ie, rc, Td = i32(ex), i32(refc), dev(T)
    R = torch.empty((nw, ndet, nex, nact), dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_ph_compact_R(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(rc), A.ptr(Td),
                                     A.ptr(R), A.stream_ptr()))
    ref = O.ph_compact_R(ex, refc, T)
    np.testing.assert_allclose(R.cpu().numpy(), ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_ph_calculate_R_matches_oracle():
    rng = np.random.default_rng(7)
    nw, nelec, nact = 19, 10, 30
    T = rand_c(rng, nw, nact, nelec)
    refc = np.arange(nelec)
    shells = [shells_for(rng, nelec, nact, n, k + 1) for k, n in enumerate([12, 30, 20, 9])]
    T[:, shells[1][4, 2], :] = 0                # determinants into this orbital are singular
    nd = 1 + sum(len(s) for s in shells)
    wgt = rand_c(rng, nd, nw)
    Td, rc, Wd = dev(T), i32(refc), dev(wgt)
    R = torch.zeros((nw, nelec, nact), dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_ph_accumulate_R(nw, 1, 0, nact, nelec, None, A.ptr(rc), None, None,
                                        A.ptr(Wd), A.ptr(R), A.stream_ptr()))
    row = 1
    keep = []
    for sh in shells:
        nex, ndet = sh.shape[1] // 2, sh.shape[0]
        ie = i32(sh)
        ov = torch.zeros((ndet, nw), dtype=torch.complex128, device="cuda")
        Rc = torch.empty((nw, ndet, nex, nact), dtype=torch.complex128, device="cuda")
        A.check(A.lib().afq_ph_overlaps(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(Td),
                                        A.ptr(ov), A.stream_ptr()))
        A.check(A.lib().afq_ph_compact_R(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(rc),
                                         A.ptr(Td), A.ptr(Rc), A.stream_ptr()))
        wsh = Wd[row:row + ndet]
        A.check(A.lib().afq_ph_accumulate_R(nw, ndet, nex, nact, nelec, A.ptr(ie), A.ptr(rc),
                                            A.ptr(Rc), A.ptr(ov), A.ptr(wsh), A.ptr(R),
                                            A.stream_ptr()))
        keep += [ie, ov, Rc, wsh]
        row += ndet
    torch.cuda.synchronize()
    ref = O.ph_calculate_R(shells, refc, T, wgt)
    np.testing.assert_allclose(R.cpu().numpy(), ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("nmo,nel,nact,nb", [(20, 5, 12, 7), (64, 10, 64, 33), (48, 16, 30, 5)])
def test_ph_woodbury_matches_oracle(nmo, nel, nact, nb):
    rng = np.random.default_rng(nmo + nel)
    At = rand_c(rng, nact, nmo)
    ref = np.sort(rng.choice(nact, nel, replace=False))
    Phi = rand_c(rng, nb, nmo, nel)
    ov0 = rand_c(rng, nb) * 0.1
    Ad, rd, Pd = dev(At), i32(ref), dev(Phi)
    G = torch.zeros((nb, nel, nmo), dtype=torch.complex128, device="cuda")
    ov = dev(ov0)
    QQ0 = torch.zeros((nb, nact, nel), dtype=torch.complex128, device="cuda")
    work = torch.zeros(nel * nmo, dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_ph_mixed_density_matrix_woodbury(
        nb, nmo, nel, nact, A.ptr(Ad), A.ptr(rd), A.ptr(Pd), nmo * nel, A.ptr(G), A.ptr(ov),
        A.ptr(QQ0), A.ptr(work), A.stream_ptr()))
    Gr, ovr, Qr = O.ph_woodbury(At, ref, Phi, ov0)
    np.testing.assert_allclose(G.cpu().numpy(), Gr, rtol=1e-10, atol=1e-12 * np.abs(Gr).max())
    np.testing.assert_allclose(QQ0.cpu().numpy(), Qr, rtol=1e-10, atol=1e-12 * np.abs(Qr).max())
    np.testing.assert_allclose(ov.cpu().numpy(), ovr, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nu,nd,nw", [(1, 1, 1), (23, 17, 9), (300, 260, 70)])
def test_ph_spin_weights_match_oracle(nu, nd, nw):
    from tests.test_oracle_phmsd_R import random_csr
    rng = np.random.default_rng(nu + nw)
    C, cp, ci, cv = random_csr(rng, nu, nd)
    rup, rdn = rand_c(rng, nu, nw), rand_c(rng, nd, nw)
    cpd, cid, cvd, rud, rdd = i32(cp), i32(ci), dev(cv), dev(rup), dev(rdn)
    wgt = torch.empty((nu, nw), dtype=torch.complex128, device="cuda")
    ov = torch.empty(nw, dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_ph_spin_weights(nw, nu, nd, A.ptr(cpd), A.ptr(cid), A.ptr(cvd),
                                        A.ptr(rud), A.ptr(rdd), A.ptr(wgt), A.ptr(ov),
                                        A.stream_ptr()))
    wr, ovr = O.ph_spin_weights(cp, ci, cv, rup, rdn)
    np.testing.assert_allclose(wgt.cpu().numpy(), wr, rtol=1e-12, atol=1e-14 * np.abs(wr).max())
    np.testing.assert_allclose(ov.cpu().numpy(), ovr, rtol=1e-11, atol=1e-13 * np.abs(ovr).max())
