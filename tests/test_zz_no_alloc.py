"""afq_ph_excited_2body_energy_noalloc (static scratch, no allocation) against
afq_ph_excited_2body_energy (per-call temporary) through the C ABI: two calls in a row on one
GPU, the second smaller, each bit-identical to the allocating entry point; an oversized call is
refused before any launch."""
import numpy as np
import pytest

from tests._helpers import rand_c

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402


def inputs(rng, nw, ndet, nex, nelec, nact, nchol):
    iexcit = np.stack([np.concatenate([rng.choice(nelec, nex, replace=False),
                                       nelec + rng.choice(nact - nelec, nex, replace=False)])
                       for _ in range(ndet)]).astype(np.int32)
    refc = rng.permutation(nelec).astype(np.int32)
    return dict(iexcit=iexcit, refc=refc, Twina=rand_c(rng, nw, nelec, nchol, nact),
                R=rand_c(rng, nw, ndet, nex, nact), wgt=rand_c(rng, ndet, nw),
                EX=rand_c(rng, nw), EJ=rand_c(rng, nw))


def call(fn, d, nex):
    nw, ndet = d["R"].shape[0], d["R"].shape[1]
    nelec, nchol, nact = d["Twina"].shape[1:]
    dev = {k: A.to_device(v) for k, v in d.items()}
    KE = torch.zeros((ndet, nw, nchol), dtype=torch.complex128, device="cuda")
    A.check(fn(nw, ndet, nex, nact, nelec, nchol, A.ptr(dev["iexcit"]), A.ptr(dev["refc"]),
               A.ptr(dev["Twina"]), A.ptr(dev["R"]), A.ptr(dev["wgt"]), A.ptr(dev["EX"]),
               A.ptr(dev["EJ"]), A.ptr(KE), A.stream_ptr()))
    torch.cuda.synchronize()
    return dev["EX"].cpu().numpy(), dev["EJ"].cpu().numpy(), KE.cpu().numpy()


def test_noalloc_matches_allocating_entry_point():
    L = A.lib()
    assert L.afq_ph_noalloc_scratch_doubles() == 1 << 20
    rng = np.random.default_rng(2024)
    # the c5 PH shape (64 active orbitals, 10 electrons, 320 vectors), then a smaller call that
    # leaves part of the first call's partials in the shared buffer
    for nw, ndet, nex, nelec, nact, nchol in [(16, 300, 4, 10, 64, 320), (3, 40, 3, 10, 24, 70)]:
        d = inputs(rng, nw, ndet, nex, nelec, nact, nchol)
        print(f"noalloc 2-body: nw={nw} ndet={ndet} nex={nex} nelec={nelec} nact={nact} "
              f"nchol={nchol}", flush=True)
        ref = call(L.afq_ph_excited_2body_energy, d, nex)
        got = call(L.afq_ph_excited_2body_energy_noalloc, d, nex)
        for r, g, name in zip(ref, got, ("EX", "EJ", "KE")):
            assert np.array_equal(r.view(np.float64), g.view(np.float64)), name
        assert np.all(np.isfinite(got[0])) and np.abs(got[0] - d["EX"]).max() > 0


def test_noalloc_refuses_oversized_call():
    L = A.lib()
    # 4 * 1025 * ceil(1024 / 4) doubles of partials > 2^20: refused before any launch
    rc = L.afq_ph_excited_2body_energy_noalloc(1025, 10, 2, 64, 10, 1024, None, None, None, None,
                                               None, None, None, None, A.stream_ptr())
    assert rc == 6   # AFQ_ECAPACITY
    assert "static scratch" in L.afq_last_error().decode()
