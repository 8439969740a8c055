#!/usr/bin/env python3
"""Walker-step throughput on one MI355X: the phaseless-AFQMC step with a Cholesky Hamiltonian.

Workload (default): NMO=128, 32 up + 32 down electrons, Nchol=640 (Cholesky vectors decaying as
exp(-n/64)), 1024 walkers, dt=0.005, the local-energy step (G -> vbias -> X -> vHS ->
P1 exp(iV) P1 -> G, E1, EXX, EJ -> weights) with QR re-orthogonalisation every 5 steps: the
largest single-system workload. The NMO=64 / 16+16 / Nchol=320 / 512-walker workload runs after
it (--extra) and its results are added under its name. Synthetic seeded inputs; everything
resident in HBM before the timed region.

Every precision form of the same step is timed (--forms): the plain FP64 form and the mixed
forms, which change only the Cholesky contractions (vbias and vHS in FP32; the exchange in FP32,
three-slice BF16 or the int8 Ozaki split); G, overlaps, LU, QR, propagation, weights and energy
sums stay FP64. Each form runs its timed steps twice from the same start: eagerly with region
timers (HIP events on the launch stream), then as one captured HIP graph replayed once; the
graph's final walker buffer must equal the eager run's bit for bit (graph_matches_eager).
`value` is the graph throughput of the fastest form whose largest per-walker
|E_loc(form) - E_loc(FP64)| on identical walker states is below 1e-5 Ha; the line carries every
form's throughput and gate error. A step = one pass over all walkers.

roofline: the exchange-energy kernel (the dominant one), algorithmic FLOP per launch
W*(8 N^2 M K + 8 N^2 K) divided by its mean duration measured with HIP events on the launch
stream over the eager timed run. cpu_baseline: the C++ FP64 oracle (OpenMP over walkers) timed
on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "walker-steps/s mixed vs FP64 at |ΔE_loc|<1e-5 Ha; % MFMA/HBM peak"
# dense MFMA peak of the arithmetic each form's exchange kernel runs on
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "int8": 5000.0}
# executed MFMA flops per algorithmic flop of the exchange contraction
# (BF16X3: 3 real products per complex product x 3 hi/lo terms = 18 flops per complex
# multiply-add against the 8 counted; int8: the slice products per Cholesky vector, 15 for the
# five-slice kernel of N <= 16, per-vector orders for N > 16 with their mean on the NMO=128 /
# 32+32 / Nchol=640 workload read from the SQ_INSTS_VALU_MFMA_MOPS_I8 counter)
EXECUTED_PER_ALGO = {"fp64": 1.0, "fp32": 1.0, "bf16x3": 18.0 / 8.0, "int8": 15.0}
INT8_PRODUCTS = {"c3": 13.37}


def executed_per_algo(form, workload):
    if form == "int8":
        return INT8_PRODUCTS.get(workload, EXECUTED_PER_ALGO["int8"])
    return EXECUTED_PER_ALGO[form]
GATE_HA = 1e-5
# HBM bytes per launch of each form's exchange kernel (FETCH_SIZE and WRITE_SIZE counter runs
# scaled by a copy kernel of known size), as recorded in profiles/walker_step_c2_hbm_traffic.md;
# null for a workload or form without such a measurement
TRAFFIC_BYTES = {("c2", "int8"): 38.54e6, ("c2", "fp64"): 53.29e6,
                 ("c3", "int8"): 586.57e6, ("c3", "fp64"): 21822.38e6,
                 ("c3", "bf16x3"): 1739.17e6}
TRAFFIC_SOURCE = {"c2": "profiles/walker_step_c2_hbm_traffic.md",
                  "c3": "profiles/walker_step_c3_hbm_traffic.md"}


def exchange_flops(W, M, N, K):
    return W * (8.0 * N * N * M * K + 8.0 * N * N * K)


def step_flops(M, N, K):
    """Algorithmic FLOP of one local-energy walker-step (complex MAC = 8, real x complex = 4)."""
    return (8 * M * N * N + 8 * N ** 3 + 8 * M * N * N      # T, LU + inverse, G
            + 8 * N * M * K                                  # vbias
            + 4 * M * M * K                                  # vHS
            + 16 * M * M * N + 48 * M * M * N                # 2 P1 products, Taylor-6
            + 16 * M * N * N + 8 * N ** 3 + 8 * N * M        # new G, LU/inverse, E1
            + 8 * N * N * M * K + 8 * N * N * K              # exchange build + reduce
            + (16 * M * N * N - 16 * N ** 3 / 3) / 5)        # QR every 5 steps


POPCTL_INTERVAL = 10   # population_control_interval default (config.h:44)


def run_loop_step(prop, wset, xi_t, Eshift, t, ortho, uniforms):
    """One iteration of run_afqmc's loop (run_afqmc.cpp:69-84): Propagate, Orthogonalize every
    `ortho` steps, population control every POPCTL_INTERVAL steps."""
    prop.Propagate(wset, xi_t, Eshift)
    if t % ortho == 0:
        prop.Orthogonalize(wset)
    if t % POPCTL_INTERVAL == 0:
        prop.population_control(wset, uniforms[t])


def same_bits(torch, a, b):
    """Bit-for-bit equality of two complex walker buffers (NaN payloads included)."""
    return bool(torch.equal(torch.view_as_real(a).view(torch.int64),
                            torch.view_as_real(b).view(torch.int64)))


def run_form(A, torch, s, W, xi_dev, uniforms, warmup, steps, ortho, precision, hybrid, graph,
             tile_cache=None):
    ham = A.HamiltonianOperations(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    ham.set_params(hybrid=int(hybrid))
    ham.set_precision(precision)
    if tile_cache is not None:
        # the vbias and vHS GEMM tilings of this precision, timed for this shape and walker
        # count (every tiling sums k in the same order: the choice changes no result)
        ham.autotune(W, tile_cache)
    wset = A.WalkerSet(ham, W)
    prop = A.Propagator(ham)
    prop.seed(wset)
    sums = prop.weighted_sums(wset).cpu().numpy()
    Eshift = float((sums[0] / sums[1]).real)
    for t in range(warmup):
        run_loop_step(prop, wset, xi_dev[t], Eshift, t, ortho, uniforms)
    torch.cuda.synchronize()
    ham.enable_timers(16 * steps + 16)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for t in range(warmup, warmup + steps):
        run_loop_step(prop, wset, xi_dev[t], Eshift, t, ortho, uniforms)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    timers = ham.read_timers()
    ham.enable_timers(0)
    out = dict(ham=ham, wset=wset, ms=ms, eager_ms=ms, wall_s=wall, timers=timers, Eshift=Eshift,
               graph=None)
    if graph:
        # the same steps from the same start, captured once (untimed) and replayed once (timed)
        gset = A.WalkerSet(ham, W)
        prop.seed(gset)
        for t in range(warmup):
            run_loop_step(prop, gset, xi_dev[t], Eshift, t, ortho, uniforms)
        torch.cuda.synchronize()

        def steps_fn():
            for t in range(warmup, warmup + steps):
                run_loop_step(prop, gset, xi_dev[t], Eshift, t, ortho, uniforms)

        g = A.capture(steps_fn)
        torch.cuda.synchronize()
        start.record()
        g.replay()
        stop.record()
        torch.cuda.synchronize()
        out["ms"] = start.elapsed_time(stop)
        out["graph"] = same_bits(torch, gset.buf, wset.buf)   # bit-identical to the eager run
        del g, gset
    return out


def eloc(A, torch, ham, wset):
    W = wset.W
    E = torch.zeros((W, 3), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(W, dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_wfn_energy(ham.handle, W, A.ptr(wset.buf), wset.ldw, A.ptr(E), A.ptr(ov),
                                   A.stream_ptr()))
    torch.cuda.synchronize()
    return E.cpu().numpy()


def contraction_errors(A, torch, ham64, hamf, wset):
    """Relative error (max |x - x64| / max |x64|) of a form's vbias and vHS against FP64 on the
    same walker state: vbias from the state's G, vHS from that vbias as X."""
    W = wset.W
    G = torch.zeros((W, ham64.N * ham64.M), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(W, dtype=torch.complex128, device="cuda")
    A.check(A.lib().afq_mixed_density_matrix(ham64.handle, W, A.ptr(wset.buf), wset.ldw,
                                             A.ptr(G), A.ptr(ov), A.stream_ptr()))
    out = {}
    v = {}
    for key, h in (("fp64", ham64), ("form", hamf)):
        h.reserve(W)
        v[key] = torch.zeros((W, h.K), dtype=torch.complex128, device="cuda")
        h.vbias(G, v[key])
    V = {}
    for key, h in (("fp64", ham64), ("form", hamf)):
        V[key] = torch.zeros((W, h.M * h.M), dtype=torch.complex128, device="cuda")
        h.vHS(v["fp64"], V[key])
    torch.cuda.synchronize()
    for name, d in (("vbias", v), ("vHS", V)):
        ref = d["fp64"]
        out[name] = float((d["form"] - ref).abs().max() / ref.abs().max())
    return out


def cpu_baseline(s, W, ortho, seconds):
    """The FP64 oracle (C++, OpenMP over walkers) on a bounded sample of the same workload."""
    import numpy as np

    from oracle import oracle as O
    from safire_amd import synth
    Ws = W
    ref = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt)
    buf = ref.new_walkers(Ws)
    ref.seed_properties(buf)
    Eshift = float(ref.energy(buf)[0].sum(1).real.mean())
    max_steps = 200
    uniforms = np.random.default_rng([s.seed, 31]).random(max_steps)
    done, el = 0, 0.0
    while done < max_steps and el < seconds:
        xi = synth.fields(s.seed, 1, W, s.K, first_step=done)[0]   # input, not timed
        t0 = time.perf_counter()
        ref.step(buf, xi[:Ws], Eshift, hybrid=False)
        if done % ortho == 0:
            ref.orthogonalize(buf)
        if done % POPCTL_INTERVAL == 0:
            O.population_control(buf, s.M * s.N, uniforms[done])
        el += time.perf_counter() - t0
        done += 1
    return {"value": round(Ws * done / el, 2), "unit": "walker-steps/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": (f"{s.name} system, {Ws} of {W} walkers, {done} local-energy steps "
                       f"(QR every {ortho}), C++ FP64 oracle, OpenMP over walkers, {el:.1f} s")}


def bench_phmsd(args, A, torch, synth):
    """c5: the PHMSD walker step (COLLINEAR walkers, 1000 determinants) in the local-energy
    form, FP64 (the PHMSD energy has no reduced-precision form). A step = one pass of
    Propagate over all walkers; QR every 5 steps and population control every 10, as the c2 leg.
    Eager launches, timed with HIP events per step (a progress line per step)."""
    import numpy as np
    s, W, steps, ortho = synth.workload("c5")
    steps = args.steps if args.steps != 100 else 5
    warmup = min(args.warmup, 1)
    occs, ci = synth.phmsd_expansion(s.M, s.N, 1000, s.seed)
    # all M orbitals of the trial's QR as rows (its first N are the c1-c3 trial's)
    orb = np.ascontiguousarray(s.meta["orbitals"].conj().T).astype(np.complex128)[None]
    ph = A.PHMSD(occs, ci, orb, s.h, s.Likn, 0.0, walker_block=64)
    ph.prepare(s.dt, W)
    ph.set_params(hybrid=0)
    ph.set_energy_algorithm(args.ph_energy)
    xi = torch.from_numpy(synth.fields(s.seed, warmup + steps, W, s.K)).to("cuda")
    uniforms = np.random.default_rng(s.seed).random(warmup + steps)
    print(f"[bench] c5: M={s.M} N={s.N}+{s.N} K={s.K} W={W}, {len(ci)} determinants "
          f"({ph.nunique[0]} + {ph.nunique[1]} unique excitations)", file=sys.stderr, flush=True)
    b = A.to_device(ph.new_walkers(W))
    ph.seed(b)
    E, _ = ph.energy(b)
    Eshift = float(E.sum(1).real.mean())
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    total = 0.0
    for t in range(warmup + steps):
        start.record()
        ph.step(b, xi[t], Eshift)
        if t % ortho == 0:
            ph.Orthogonalize(b)
        if t % POPCTL_INTERVAL == 0:
            ph.population_control(b, float(uniforms[t]))
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop)
        if t >= warmup:
            total += ms
        print(f"[bench] c5 step {t}: {ms:.1f} ms", file=sys.stderr, flush=True)
    ms = total / steps
    value = W / (ms / 1e3)
    print(f"[bench] c5 fp64 ({args.ph_energy} energy): {ms:.2f} ms/step, {value:.0f} "
          f"walker-steps/s", file=sys.stderr, flush=True)
    # the dominant part of the step: the PHMSD local energy, timed alone (HIP events on the
    # launch stream, 3 calls) on the walkers of the last step
    eflops = ph_energy_flops(occs, s.N, s.M, s.K, W, args.ph_energy)
    ph.energy(b)
    torch.cuda.synchronize()
    start.record()
    for _ in range(3):
        ph.energy(b)
    stop.record()
    torch.cuda.synchronize()
    e_ms = start.elapsed_time(stop) / 3
    e_tf = eflops / (e_ms / 1e3) / 1e12
    print(f"[bench] c5 energy ({args.ph_energy}): {e_ms:.2f} ms per call, {e_tf:.2f} TFLOP/s",
          file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = ph_cpu_baseline(s, occs, ci, orb, b, args.ph_energy, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "walker-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": f"c5: PHMSD trial, NMO={s.M}, {s.N}+{s.N} electrons, "
                               f"Nchol={s.K}, {len(ci)} determinants of rank 1-4, {W} walkers, "
                               f"dt={s.dt}, local-energy step, QR every {ortho}",
                   "form": "FP64, eager launches", "ph_energy_algorithm": args.ph_energy,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "fp64 (vector ALU; no MFMA form)", "achieved": round(e_tf, 3),
                     "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                     "frac": round(e_tf / PEAK_TFLOPS["fp64"], 4), "traffic": None,
                     "kernel": f"PHMSD local energy ({args.ph_energy}: k_ph_2body_lds "
                               f"and companions, one call)" if args.ph_energy == "woodbury"
                               else "PHMSD local energy (reference: k_ph0_inverse + k_ph0_chol, "
                                    "one call)",
                     "flop_per_launch": eflops, "mean_launch_ms": round(e_ms, 4)},
        "cpu_baseline": cpu,
    }
    ph.close()
    print(json.dumps(out))
    return 0


def ph_energy_flops(occs, N, M, K, W, algorithm):
    """Algorithmic FLOP of one PHMSD local energy over W walkers (complex MAC = 8), counted from
    the expansion: per spin, every unique excited configuration of rank nex contributes
    K (nelec nex nact + nelec^2) complex MACs to the excited-determinant 2-body energy (Woodbury,
    phmsd_impl.hpp:29-147), or, for the reference algorithm (PHMSD.cpp:122-370), K nelec^3 for
    its Y = P T^-1 rows and tr(Y Y) (nelec = N, nact = M)."""
    import numpy as np
    total = 0.0
    for spin in range(2):
        part = occs[:, spin * N:(spin + 1) * N] - spin * M
        uniq = np.unique(np.sort(part, axis=1), axis=0)
        ranks = (uniq >= N).sum(1)
        ranks = ranks[ranks > 0]
        if algorithm == "woodbury":
            total += float(np.sum(K * (N * ranks * M + N * N)))
        else:
            total += float(len(ranks) * K * (N ** 3 + N * N))
    return 8.0 * W * total


def ph_cpu_baseline(s, occs, ci, orb, b, algorithm, seconds):
    """The oracle's PHMSD local energy (numpy restatement) on a bounded sample of the c5
    walkers: the step's dominant part; walker-energies/s reported as walker-steps/s, the rest of
    the step (propagation, vbias, QR) excluded, so an upper bound on the CPU step rate."""
    import numpy as np

    from oracle import ph_energy as P
    buf = b.cpu().numpy()
    M, N = s.M, s.N
    Pu = buf[:, :M * N].reshape(-1, M, N)
    Pd = buf[:, M * N:2 * M * N].reshape(-1, M, N)
    # orbital rows of both spins, half-rotated Cholesky vectors and one-body term, as the
    # oracle's tests build them
    A = [np.ascontiguousarray(orb[0]), np.ascontiguousarray(orb[-1])]
    L = np.ascontiguousarray(s.Likn.transpose(2, 0, 1))
    Lna = [np.ascontiguousarray(np.einsum("al,nlk->nak", A[sp], L)) for sp in range(2)]
    haj = [np.ascontiguousarray(A[sp] @ s.h) for sp in range(2)]
    struct = P.build_ph_struct(ci, occs, M, N, N)
    fn = P.energy_alg1 if algorithm == "woodbury" else P.energy_alg0
    done, el = 0, 0.0
    while done < len(Pu) and el < seconds:
        t0 = time.perf_counter()
        fn(struct, A, Lna, haj, 0.0, Pu[done:done + 1], Pd[done:done + 1])
        el += time.perf_counter() - t0
        done += 1
    return {"value": round(done / el, 3), "unit": "walker-steps/s", "cores": 1, "kind": "port",
            "sample": (f"c5 system, PHMSD local energy ({algorithm}) of {done} walkers, numpy "
                       f"restatement (oracle/ph_energy.py), {el:.1f} s; the rest of the step is "
                       f"not included (an upper bound on the CPU step rate)")}


def batch_cpu_baseline(systems, ortho, seconds):
    """The FP64 oracle (C++, OpenMP over walkers) stepping the systems one after another, each
    with all its walkers, until the time bound; walker-steps/s over the systems reached."""
    import numpy as np

    from oracle import oracle as O
    from safire_amd import synth
    done, ws, el = 0, 0, 0.0
    for s, W in systems:
        if el >= seconds:
            break
        ref = O.System(s.h, s.Likn, s.PsiT, 0.0, s.dt)
        buf = ref.new_walkers(W)
        ref.seed_properties(buf)
        Eshift = float(ref.energy(buf)[0].sum(1).real.mean())
        xi = synth.fields(s.seed, 1, W, s.K)[0]
        t0 = time.perf_counter()
        ref.step(buf, xi, Eshift, hybrid=False)
        ref.orthogonalize(buf)
        O.population_control(buf, s.M * s.N, float(np.random.default_rng(s.seed).random()))
        el += time.perf_counter() - t0
        done += 1
        ws += W
    return {"value": round(ws / el, 2), "unit": "walker-steps/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": (f"first {done} of {len(systems)} systems, one local-energy step each with "
                       f"QR and population control, all their walkers ({ws}), C++ FP64 oracle, "
                       f"OpenMP over walkers, {el:.1f} s")}


def bench_batch(args, A, torch, synth):
    """c4: 32 independent systems of uneven size. Each form runs every system's steps one system
    at a time (eager launches: the reference's way, one job per system), then as one captured
    graph with the systems spread over --streams concurrent capture streams (the walkers must
    match the one-at-a-time run bit for bit). value = total walker-steps/s of the batched graph
    of the fastest form within the accuracy gate (the int8 exchange's gate is checked per system
    by the c2/c3 legs and the tests; here both forms are timed)."""
    from safire_amd.batch import SystemBatch
    systems, steps, ortho = synth.batch_workload()
    steps = args.steps if args.steps != 100 else steps
    res = {}
    for name, prec in (("fp64", A.FP64), ("int8", A.INT8_SPLIT)):
        b = SystemBatch(systems, steps=steps, ortho=ortho, precision=prec,
                        tile_cache=args.tile_cache)
        b.run_sequential()                                  # warm-up (kernel attributes)
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run_sequential()
        torch.cuda.synchronize()
        seq = time.perf_counter() - t0
        # exchange kernels of every system, timed in a second one-at-a-time run (region timers)
        b.reset()
        for it in b.items:
            it["ham"].enable_timers(16 * steps + 16)
        b.run_sequential()
        torch.cuda.synchronize()
        ex_flop, ex_ms = 0.0, 0.0
        for it in b.items:
            tm = it["ham"].read_timers()
            it["ham"].enable_timers(0)
            sy = it["sys"]
            ex_flop += tm["exchange"][1] * exchange_flops(it["W"], sy.M, sy.N, sy.K)
            ex_ms += tm["exchange"][0]
        ref = [it["wset"].buf.clone() for it in b.items]
        b.reset()
        g = b.capture(streams=args.streams)
        b.reset()
        g = b.capture(streams=args.streams)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        g.replay()
        stop.record()
        torch.cuda.synchronize()
        bat = start.elapsed_time(stop) / 1e3
        same = all(same_bits(torch, it["wset"].buf, r) for it, r in zip(b.items, ref))
        ws = b.walker_steps()
        ex_tf = ex_flop / (ex_ms / 1e3) / 1e12 if ex_ms > 0 else 0.0
        res[name] = {"exchange_tflops": round(ex_tf, 3),
                     "exchange_ms_per_step": round(ex_ms / steps, 4),
                     "tilings_autotuned": len(b.tilings),
                     "batched_walker_steps_per_s": round(ws / bat, 1),
                     "one_at_a_time_walker_steps_per_s": round(ws / seq, 1),
                     "batched_ms_per_step": round(bat * 1e3 / steps, 4),
                     "one_at_a_time_ms_per_step": round(seq * 1e3 / steps, 4),
                     "batched_matches_one_at_a_time": same}
        print(f"[bench] c4 {name}: batched graph {bat * 1e3 / steps:.3f} ms/step, one at a time "
              f"{seq * 1e3 / steps:.3f} ms/step, identical walkers: {same}", file=sys.stderr,
              flush=True)
        del g
        b.close()
    head = "int8" if res["int8"]["batched_walker_steps_per_s"] >= res["fp64"]["batched_walker_steps_per_s"] else "fp64"
    Ms = sorted({s.M for s, _ in systems})
    out = {
        "metric": METRIC, "value": res[head]["batched_walker_steps_per_s"],
        "unit": "walker-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": res[head]["batched_ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": f"{head} contractions / fp64",
        "data": "synthetic",
        "config": {"workload": f"c4: {len(systems)} independent systems, NMO in {Ms}, "
                               f"N = NMO/4, Nchol = 5 NMO, walkers 32-256 "
                               f"({sum(W for _, W in systems)} in all), QR every {ortho}",
                   "form": "one HIP graph per run, systems on "
                           f"{args.streams} concurrent capture streams",
                   "parallelism": "single GPU"},
        "forms": res,
        "batched_over_one_at_a_time": round(res[head]["batched_walker_steps_per_s"] /
                                            res[head]["one_at_a_time_walker_steps_per_s"], 3),
        "mixed_over_fp64": round(res["int8"]["batched_walker_steps_per_s"] /
                                 res["fp64"]["batched_walker_steps_per_s"], 3),
        "roofline": {"bound": "mfma", "achieved": res[head]["exchange_tflops"],
                     "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                     "frac": round(res[head]["exchange_tflops"] / PEAK_TFLOPS[head], 4),
                     "traffic": None,
                     "kernel": f"k_exchange ({head}) over all systems: total algorithmic FLOP / "
                               f"total launch time, one-at-a-time timed run",
                     "executed_frac": round(res[head]["exchange_tflops"] *
                                            executed_per_algo(head, "c4") / PEAK_TFLOPS[head], 4)},
        "cpu_baseline": None if args.no_cpu else batch_cpu_baseline(systems, ortho,
                                                                    args.cpu_seconds),
    }
    print(json.dumps(out))
    return 0


FORM_PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "int8": 3}


def bench_single(args, A, torch, synth, workload, with_cpu, with_hybrid):
    """One single-system workload (c1, c2 or c3): every requested form timed, the accuracy gate,
    the roofline of the exchange kernel of the headline form; returns the JSON line as a dict.
    A form the shape does not support (AFQ_EUNSUPPORTED) is reported as such and skipped."""
    import numpy as np
    s, W, _, ortho = synth.workload(workload)
    M, N, K = s.M, s.N, s.K
    nst = max(args.warmup + args.steps, 31)   # the accuracy gate replays 31 steps
    t0 = time.perf_counter()
    xi = synth.fields(s.seed, nst, W, K)
    xi_dev = torch.from_numpy(xi).to("cuda")
    del xi
    print(f"[bench] {s.name}: M={M} N={N} K={K} W={W}, fields ready in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    uniforms = np.random.default_rng([s.seed, 31]).random(nst)
    forms, unsupported = {}, {}
    for f in args.forms.split(","):
        try:
            forms[f] = run_form(A, torch, s, W, xi_dev, uniforms, args.warmup, args.steps, ortho,
                                FORM_PRECISION[f], hybrid=False, graph=not args.no_graph,
                                tile_cache=args.tile_cache)
        except A.AfqError as e:
            if e.code != 5:   # AFQ_EUNSUPPORTED: the form has no kernel for this shape
                raise
            unsupported[f] = str(e)
            print(f"[bench] {f}: not supported at this shape ({e})", file=sys.stderr, flush=True)
            continue
        r = forms[f]
        print(f"[bench] {f}: {r['ms'] / args.steps:.3f} ms/step "
              f"({'graph' if r['graph'] is not None else 'eager'}; eager "
              f"{r['eager_ms'] / args.steps:.3f}; graph == eager: {r['graph']}), "
              f"{W * args.steps / (r['ms'] / 1e3):.0f} walker-steps/s; regions (eager, ms total): "
              + ", ".join(f"{k}={v[0]:.2f}" for k, v in r["timers"].items() if v[1]),
              file=sys.stderr, flush=True)

    # accuracy of each reduced-precision form on identical walker states: the seeded trial state,
    # the states after steps 1 and 3 and those right after population control at steps 10, 20
    # and 30 of an (untimed) FP64 run, over the walkers whose weight is not zero (a walker killed
    # by the phaseless projection keeps weight 0 and enters no estimator; on the NMO=128
    # Hamiltonian most walkers are killed within a few steps, hence the early states)
    ref_form = forms.get("fp64")
    gate = {}
    cerr = {}
    if ref_form is not None:
        P = M * N
        ham64 = ref_form["ham"]
        states = [A.WalkerSet(ham64, W)]
        run = A.WalkerSet(ham64, W)
        prop = A.Propagator(ham64)
        prop.seed(run)
        for t in range(31):
            run_loop_step(prop, run, xi_dev[t], ref_form["Eshift"], t, ortho, uniforms)
            if t in (1, 3, 10, 20, 30):
                states.append(A.WalkerSet(ham64, W, run.host()))
        for f in forms:
            if f == "fp64":
                continue
            worst, nlive = 0.0, 0
            for st in states:
                live = np.abs(st.host()[:, P + A.P_WEIGHT]) > 0
                if not np.any(live):
                    continue
                E64 = eloc(A, torch, ham64, st).sum(1)
                Ef = eloc(A, torch, forms[f]["ham"], st).sum(1)
                worst = max(worst, float(np.max(np.abs(E64 - Ef)[live])))
                nlive += int(live.sum())
            gate[f] = worst
            cerr[f] = {}
            for st in states:
                for k, e in contraction_errors(A, torch, ham64, forms[f]["ham"], st).items():
                    cerr[f][k] = max(cerr[f].get(k, 0.0), e)
            print(f"[bench] {f}: max |E_loc({f}) - E_loc(fp64)| = {worst:.3e} Ha over {nlive} "
                  f"live walker states; relative error of vbias {cerr[f]['vbias']:.2e}, "
                  f"vHS {cerr[f]['vHS']:.2e} (G is FP64 in every form)", file=sys.stderr,
                  flush=True)
        del states, run
    passing = [f for f in forms if f == "fp64" or gate.get(f, 1.0) < GATE_HA]
    # the headline: the fastest form that meets the accuracy gate
    mixed = min(passing, key=lambda f: forms[f]["ms"])
    fastest = min(forms, key=lambda f: forms[f]["ms"])

    head = forms[mixed]
    ms_step = head["ms"] / args.steps
    value = W * args.steps / (head["ms"] / 1e3)
    ex_ms, ex_n = head["timers"]["exchange"]
    ex_avg = ex_ms / max(ex_n, 1)
    achieved = exchange_flops(W, M, N, K) / (ex_avg / 1e3) / 1e12
    peak = PEAK_TFLOPS[mixed]
    traffic = TRAFFIC_BYTES.get((workload, mixed))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{mixed} contractions / fp64",
        "data": "synthetic",
        "config": {"workload": f"{s.name}: NMO={M}, {N}+{N} electrons, Nchol={K}, {W} walkers, "
                               f"dt={s.dt}, local-energy step, QR every {ortho}",
                   "form": f"{mixed} Cholesky contractions, fp64 G/LU/QR/propagation/weights",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": traffic,
                     "traffic_unit": "bytes per launch",
                     "traffic_source": TRAFFIC_SOURCE.get(workload) if traffic else None,
                     "kernel": f"k_exchange ({mixed} MFMA)",
                     "flop_per_launch": exchange_flops(W, M, N, K),
                     "mean_launch_ms": round(ex_avg, 5),
                     "executed_frac": round(achieved * executed_per_algo(mixed, workload) / peak,
                                            4)},
    }
    if ref_form is not None:
        v64 = W * args.steps / (ref_form["ms"] / 1e3)
        ex64 = ref_form["timers"]["exchange"]
        ex64_avg = ex64[0] / max(ex64[1], 1)
        tf64 = exchange_flops(W, M, N, K) / (ex64_avg / 1e3) / 1e12
        out["fp64"] = {"value": round(v64, 1),
                       "ms_per_step": round(ref_form["ms"] / args.steps, 4),
                       "exchange_tflops": round(tf64, 3),
                       "exchange_frac_fp64_peak": round(tf64 / PEAK_TFLOPS["fp64"], 4)}
        out["mixed_over_fp64"] = round(value / v64, 3)
    out["max_abs_dEloc_Ha"] = {f: gate[f] for f in gate}
    out["contraction_rel_err_vs_fp64"] = {f: {k: float(f"{e:.3g}") for k, e in c.items()}
                                          for f, c in cerr.items()}
    out["headline_form"] = mixed
    out["forms_walker_steps_per_s"] = {f: round(W * args.steps / (r["ms"] / 1e3), 1)
                                       for f, r in forms.items()}
    # MFMA utilisation of each form's exchange kernel: executed MFMA operations per launch (the
    # algorithmic count times the form's executed-per-algorithmic ratio) over its mean launch
    # time, against the dense peak of the MFMA type it runs on
    util = {}
    for f, r in forms.items():
        ms_, n_ = r["timers"]["exchange"]
        if n_:
            ex_rate = exchange_flops(W, M, N, K) / (ms_ / n_ / 1e3) / 1e12
            util[f] = round(ex_rate * executed_per_algo(f, workload) / PEAK_TFLOPS[f], 4)
    out["exchange_mfma_utilisation"] = util
    if unsupported:
        out["forms_unsupported"] = unsupported
    if fastest != mixed:
        out["note"] = (f"{fastest} is faster but misses |dE_loc| < {GATE_HA} Ha "
                       f"({gate[fastest]:.2e} Ha); value is the fastest form within the gate")
    out["step_tflops"] = round(W * step_flops(M, N, K) / (ms_step / 1e3) / 1e12, 3)
    out["regions_ms_per_step"] = {k: round(v[0] / args.steps, 4) for k, v in head["timers"].items()
                                  if v[1]}
    out["timing"] = ("one HIP graph of the timed steps, replayed once" if head["graph"] is not None
                     else "eager launches")
    out["eager_ms_per_step"] = round(head["eager_ms"] / args.steps, 4)
    if head["graph"] is not None:
        out["graph_matches_eager"] = {f: r["graph"] for f, r in forms.items()}
    if with_hybrid:
        # the reference's default step: hybrid weights, no local energy per step
        hyb = {}
        for f in sorted({"fp64", mixed}):
            r = run_form(A, torch, s, W, xi_dev, uniforms, args.warmup, args.steps, ortho,
                         FORM_PRECISION[f], hybrid=True, graph=not args.no_graph,
                         tile_cache=args.tile_cache)
            hyb[f] = {"walker_steps_per_s": round(W * args.steps / (r["ms"] / 1e3), 1),
                      "ms_per_step": round(r["ms"] / args.steps, 4)}
            r["ham"].close()
            print(f"[bench] hybrid {f}: {r['ms'] / args.steps:.3f} ms/step", file=sys.stderr,
                  flush=True)
        out["hybrid_form"] = hyb
    for f in forms.values():
        f["ham"].close()
    forms.clear()
    del xi_dev
    torch.cuda.empty_cache()
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline(s, W, ortho, args.cpu_seconds)
    return out


# keys of a secondary workload's line kept under its name in the headline line
EXTRA_KEYS = ("value", "ms_per_step", "config", "headline_form", "fp64", "mixed_over_fp64",
              "max_abs_dEloc_Ha", "forms_walker_steps_per_s", "forms_unsupported", "roofline",
              "exchange_mfma_utilisation", "contraction_rel_err_vs_fp64",
              "regions_ms_per_step", "graph_matches_eager")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--extra", default="c2", choices=["none", "c1", "c2", "c3"],
                    help="a second single-system workload whose results are added under its "
                         "name (not timed for the CPU baseline or the hybrid form)")
    ap.add_argument("--streams", type=int, default=4,
                    help="c4: capture streams the systems are spread over")
    ap.add_argument("--tile-cache", default=os.path.join(tempfile.gettempdir(),
                                                         "safire_amd_tiles.json"),
                    help="the GEMM autotuner's JSON cache (every form's vbias and vHS tilings "
                         "are tuned for the workload's shape)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--forms", default="fp64,fp32,bf16x3,int8")
    ap.add_argument("--ph-energy", default="woodbury", choices=["woodbury", "reference"],
                    help="c5: the PHMSD energy algorithm (the reference's `algorithm` option)")
    ap.add_argument("--no-hybrid", action="store_true",
                    help="skip the hybrid-weight (no per-step local energy) leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the eager launches instead of one captured graph of the steps")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the walker step runs on a single MI355X here and "
              f"--gpus {args.gpus} is not supported (walkers would shard across GPUs with no "
              f"communication inside a step).", file=sys.stderr)
        return 2

    import torch

    from safire_amd import afqmc as A
    from safire_amd import synth

    if args.workload == "c4":
        return bench_batch(args, A, torch, synth)
    if args.workload == "c5":
        return bench_phmsd(args, A, torch, synth)
    out = bench_single(args, A, torch, synth, args.workload, with_cpu=not args.no_cpu,
                       with_hybrid=not args.no_hybrid)
    if args.extra not in ("none", args.workload):
        ex = bench_single(args, A, torch, synth, args.extra, with_cpu=False, with_hybrid=False)
        out[args.extra] = {k: ex[k] for k in EXTRA_KEYS if k in ex}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
