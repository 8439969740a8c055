"""The GEMM autotuner (afq_autotune): every tiling of the FP32 vbias and vHS GEMMs gives the
same bits, the choice and its JSON cache round-trip, and a cached shape is not timed again."""
import json

import numpy as np
import pytest

from tests._helpers import tiny_system

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from safire_amd import afqmc as A  # noqa: E402


def contractions(ham, G, X):
    v = torch.zeros((G.shape[0], ham.K), dtype=torch.complex128, device="cuda")
    V = torch.zeros((X.shape[0], ham.M * ham.M), dtype=torch.complex128, device="cuda")
    ham.vbias(G, v)
    ham.vHS(X, V)
    torch.cuda.synchronize()
    return v.clone(), V.clone()


@pytest.mark.parametrize("precision", [A.FP32, A.FP64])
def test_every_tiling_gives_the_same_bits(tmp_path, precision):
    rng = np.random.default_rng(3)
    h, Likn, PsiT = tiny_system(rng, M=40, N=10, K=150)
    ham = A.HamiltonianOperations(h, Likn, PsiT, 0.0, 0.01)
    ham.set_precision(precision)
    W = 70
    ham.reserve(W)
    G = A.to_device(rng.standard_normal((W, 10 * 40)) + 1j * rng.standard_normal((W, 10 * 40)))
    X = A.to_device(rng.standard_normal((W, 150)) + 1j * rng.standard_normal((W, 150)))
    v0, V0 = contractions(ham, G, X)
    cache = tmp_path / "tiles.json"
    cfg = ham.autotune(W, str(cache))
    entries = json.loads(cache.read_text())
    assert len(entries) == 2 and sorted(entries.values()) == sorted(cfg)
    v1, V1 = contractions(ham, G, X)
    assert torch.equal(v0, v1) and torch.equal(V0, V1)
    # force every tiling through the cache: all give identical results
    n_vb, n_vh = (2, 4) if precision == A.FP32 else (1, 1)
    for cb in range(n_vb):
        for ch in range(n_vh):
            forced = {k: (cb if k.startswith("vbias") else ch) for k in entries}
            cache.write_text(json.dumps(forced))
            assert ham.autotune(W, str(cache)) == (cb, ch)
            v2, V2 = contractions(ham, G, X)
            assert torch.equal(v0, v2) and torch.equal(V0, V2), (cb, ch)
    # a cached shape is read, not timed: the file comes back unchanged
    before = cache.read_text()
    ham.autotune(W, str(cache))
    assert json.loads(cache.read_text()) == json.loads(before)
    ham.close()
