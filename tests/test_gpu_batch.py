"""Uneven batch of independent systems (safire_amd/batch.py).

* The captured multi-stream graph of all systems' steps leaves every system's walkers
  bit-identical to running that system alone, eagerly, from the same start (the systems share
  nothing).
* The full 32-system batch (NMO in {16, ..., 96}, N = NMO/4, Nchol = 5 NMO, 32-256 walkers each)
  against the FP64 oracle: one graph per step over all systems on 4 capture streams, every
  system's walker buffer set to the oracle's before each step, each step within
  1e-10 + 1e-13 kappa_w per walker (kappa_w: the condition number of the walker's overlap
  matrix, which amplifies rounding through G)."""
import numpy as np
import pytest

from oracle import oracle as O
from safire_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402
from safire_amd.batch import SystemBatch  # noqa: E402


@pytest.mark.parametrize("precision", [A.FP64, A.INT8_SPLIT])
def test_batched_graph_matches_systems_alone(precision):
    systems = [(synth.make_system(M, N, 5 * M, 77 + i, name=f"b{i}"), W)
               for i, (M, N, W) in enumerate([(16, 4, 40), (40, 10, 33), (24, 6, 96),
                                              (72, 18, 50), (32, 8, 64)])]
    b = SystemBatch(systems, steps=12, ortho=5, precision=precision)
    b.run_sequential()
    torch.cuda.synchronize()
    alone = [it["wset"].buf.clone() for it in b.items]
    b.reset()
    g = b.capture(streams=3)
    g.replay()
    torch.cuda.synchronize()
    for it, ref in zip(b.items, alone):
        assert torch.equal(it["wset"].buf, ref), it["sys"].name
    del g
    b.close()


def test_c4_batch_parity_with_oracle():
    systems, _, ortho = synth.batch_workload()
    steps = 5
    b = SystemBatch(systems, steps=steps, ortho=ortho, precision=A.FP64)
    # every system alone, eagerly (this also sets each kernel's attributes before capture)
    b.run_sequential()
    torch.cuda.synchronize()
    alone = [it["wset"].buf.clone() for it in b.items]
    b.reset()
    g = b.capture(streams=4)
    g.replay()
    torch.cuda.synchronize()
    for it, ref in zip(b.items, alone):
        assert torch.equal(it["wset"].buf, ref), it["sys"].name
    del g, alone
    b.reset()
    graphs = [b.capture(streams=4, t0=t, t1=t + 1) for t in range(steps)]
    refs = []
    for it in b.items:
        s, W = it["sys"], it["W"]
        ref = O.System(s.h, s.Likn, s.PsiT, s.E0, s.dt)
        buf = ref.new_walkers(W)
        ref.seed_properties(buf)
        refs.append((ref, buf, it["xi"].cpu().numpy()))
    nwalk = 0
    for t in range(steps):
        kappas = []
        for it, (ref, buf, xi) in zip(b.items, refs):
            s = it["sys"]
            Phi = buf[:, :s.M * s.N].reshape(it["W"], s.M, s.N)
            kappas.append(np.linalg.cond(s.PsiT[None] @ Phi))
            it["wset"].buf.copy_(torch.from_numpy(buf))
        graphs[t].replay()
        torch.cuda.synchronize()
        for it, (ref, buf, xi), kappa in zip(b.items, refs, kappas):
            s, W = it["sys"], it["W"]
            P = s.M * s.N
            ref.step(buf, xi[t], it["Eshift"], hybrid=False)
            if t % ortho == 0:
                ref.orthogonalize(buf)
            if t % 10 == 0:
                off, rc = O.population_control(buf, P, float(it["u"][t]))
                assert rc == 0
            g = it["wset"].host()
            dead = buf[:, P + A.P_WEIGHT] == 0
            assert np.array_equal(dead, g[:, P + A.P_WEIGHT] == 0), (s.name, t)
            e_sm = np.abs(g[:, :P] - buf[:, :P]).max(1) / np.abs(buf[:, :P]).max(1)
            S = (1 + np.abs(buf[:, P + A.P_E1]) + np.abs(buf[:, P + A.P_EXX])
                 + np.abs(buf[:, P + A.P_EJ]))
            scale = 1 + np.abs(buf[:, P:])
            for slot in (A.P_PSEUDO_ELOC, A.P_E1, A.P_EXX, A.P_EJ):
                scale[:, slot] = np.maximum(scale[:, slot], S)
            for slot in (A.P_PHASE, A.P_WEIGHT):
                scale[:, slot] = np.maximum(scale[:, slot], 1 + s.dt * S)
            e_pr = (np.abs(g[:, P:] - buf[:, P:]) / scale).max(1)
            e_sm[dead] = 0.0
            e_pr[dead] = 0.0
            bound = 1e-10 + 1e-13 * kappa
            assert np.all(e_sm < bound), (s.name, t, float((e_sm / bound).max()))
            assert np.all(e_pr < bound), (s.name, t, float((e_pr / bound).max()))
            nwalk += W
    print(f"c4: {len(b.items)} systems, {steps} steps, {nwalk} walker-steps checked")
    del graphs
    b.close()
