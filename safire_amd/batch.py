"""An uneven batch of independent systems stepped together on one GPU (the c4 workload).

Each system keeps its own handle (Hamiltonian tensors, workspace), walker set and fields; a
step of the batch is one step of every system. ``SystemBatch.capture`` records `steps` steps of
all systems into one HIP graph, the systems spread over `streams` capture streams forked from
and joined back into the launch stream, so the replay runs their kernels concurrently (small
systems leave most of the GPU idle one at a time) and pays one launch for the whole run.
``SystemBatch.run_sequential`` is the reference: the systems one at a time, eager launches.

Systems never interact (the reference runs them as separate jobs), so every system's walkers
after a batched run are bit-identical to a run of that system alone from the same start.
"""
from __future__ import annotations

import numpy as np

from . import afqmc as A
from . import synth

POPCTL_INTERVAL = 10


class SystemBatch:
    def __init__(self, systems, steps: int, ortho: int = 5, precision: int = A.FP64,
                 hybrid: bool = False, seed: int = 0, tile_cache: str | None = None):
        """tile_cache: the autotuner's JSON cache; when given, every system's GEMM tilings are
        tuned for its shape and walker count (cached shapes are not timed again)."""
        torch = A._torch()
        self.steps, self.ortho = steps, ortho
        self.items = []
        self.tilings = []
        for i, (s, W) in enumerate(systems):
            ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, s.E0, s.dt)
            ham.set_params(hybrid=int(hybrid))
            ham.set_precision(precision)
            if tile_cache is not None:
                self.tilings.append(ham.autotune(W, tile_cache))
            prop = A.Propagator(ham)
            xi = torch.from_numpy(synth.fields(s.seed, steps, W, s.K)).to("cuda")
            u = np.random.default_rng([s.seed, 31]).random(steps)
            self.items.append(dict(sys=s, W=W, ham=ham, prop=prop, xi=xi, u=u))
        self.reset()

    def reset(self):
        """Every system back at its trial (fresh walker sets, Eshift = trial energy)."""
        for it in self.items:
            it["wset"] = A.WalkerSet(it["ham"], it["W"])
            it["prop"].seed(it["wset"])
            sums = it["prop"].weighted_sums(it["wset"]).cpu().numpy()
            it["Eshift"] = float((sums[0] / sums[1]).real)

    def walker_steps(self) -> int:
        return sum(it["W"] for it in self.items) * self.steps

    def _system_steps(self, it, t0, t1):
        for t in range(t0, t1):
            it["prop"].Propagate(it["wset"], it["xi"][t], it["Eshift"])
            if t % self.ortho == 0:
                it["prop"].Orthogonalize(it["wset"])
            if t % POPCTL_INTERVAL == 0:
                it["prop"].population_control(it["wset"], float(it["u"][t]))

    def run_sequential(self):
        """All steps of each system in turn, eager launches on the current stream."""
        for it in self.items:
            self._system_steps(it, 0, self.steps)

    def capture(self, streams: int = 4, t0: int = 0, t1: int | None = None):
        """One HIP graph of steps t0 .. t1-1 (all steps by default) of all systems, the systems
        spread over `streams` capture streams: largest first, each to the least loaded stream,
        so the streams carry similar work."""
        t1 = self.steps if t1 is None else t1
        torch = A._torch()
        order = sorted(range(len(self.items)),
                       key=lambda i: -self.items[i]["W"] * self.items[i]["sys"].M ** 2 *
                       self.items[i]["sys"].K)
        lanes = [[] for _ in range(streams)]
        load = [0.0] * streams
        for i in order:
            it = self.items[i]
            j = int(np.argmin(load))
            lanes[j].append(it)
            load[j] += it["W"] * it["sys"].M ** 2 * it["sys"].K
        side = [torch.cuda.Stream() for _ in range(streams)]

        def body():
            main = torch.cuda.current_stream()
            for j, st in enumerate(side):
                st.wait_stream(main)
                with torch.cuda.stream(st):
                    for it in lanes[j]:
                        self._system_steps(it, t0, t1)
            for st in side:
                main.wait_stream(st)

        return A.capture(body)

    def close(self):
        for it in self.items:
            it["ham"].close()
