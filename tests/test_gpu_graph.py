"""The walker step, Orthogonalize and population control captured in a HIP graph: every replay,
from the same starting walkers and with eager work and host synchronisation between replays,
gives the eager run's walker buffer bit for bit (the kernels are deterministic and nothing in
the captured work is a memset or a host-dependent decision)."""
import numpy as np
import pytest

from safire_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402


def run_steps(prop, wset, xi, uniforms, Eshift, t0, n, ortho=5, popctl=10):
    for t in range(t0, t0 + n):
        prop.Propagate(wset, xi[t], Eshift)
        if t % ortho == 0:
            prop.Orthogonalize(wset)
        if t % popctl == 0:
            prop.population_control(wset, float(uniforms[t]))


@pytest.mark.parametrize("prec", [A.FP64, A.INT8_SPLIT])
def test_graph_replays_match_eager(prec):
    s = synth.make_system(24, 6, 96, seed=21)
    W, n = 40, 12
    xi = torch.from_numpy(synth.fields(s.seed, n + 2, W, s.K)).to("cuda")
    uniforms = np.random.default_rng(5).random(n + 2)
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=0)
    ham.set_precision(prec)
    prop = A.Propagator(ham)
    wset = A.WalkerSet(ham, W)
    prop.seed(wset)
    sums = prop.weighted_sums(wset).cpu().numpy()
    Eshift = float((sums[0] / sums[1]).real)
    run_steps(prop, wset, xi, uniforms, Eshift, 0, 2)        # eager warm-up: steps 0, 1
    torch.cuda.synchronize()
    start = wset.buf.clone()
    run_steps(prop, wset, xi, uniforms, Eshift, 2, n)         # eager reference: steps 2 .. n+1
    torch.cuda.synchronize()
    eager = wset.buf.clone()
    assert torch.isfinite(torch.view_as_real(eager)).all()

    wset.buf.copy_(start)
    g = A.capture(lambda: run_steps(prop, wset, xi, uniforms, Eshift, 2, n))
    torch.cuda.synchronize()
    assert torch.equal(wset.buf, start)                       # capture runs nothing
    for rep in range(3):
        wset.buf.copy_(start)
        prop.weighted_sums(wset)                              # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(wset.buf, eager), rep
    ham.close()
