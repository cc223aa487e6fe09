#!/usr/bin/env python3
"""Benchmark of the ACO hot path on one MI355X.

Default run (what `python bench.py --gpus 1 --steps K --warmup W` prints, one
JSON line): tours/s of rank-based Ant System with ant aging, n = 1000 cities
from 10 Gaussian clusters (centres uniform in [0,10000]^2, sigma 400,
clipped, rounded), EUC_2D integer distances, m = 1000 ants, candidate lists
k = 20, w = 6 ranked ants, alpha 1, beta 3, rho 0.1, fp32 pheromone /
choice_info.  One step = one ACO iteration (sweep, construction,
statistics, deposit), all on the device; the timed loop replays hipGraphs of
the iteration with the inputs resident in HBM.  Before the W warm-up
iterations the colony runs 500 iterations, well past the aging retirement
age (50), so the K timed iterations are steady state: retired ants build no
tour, recruited ants are young, and tours/s counts the tours actually
constructed.  A whole 2000-iteration run of the configuration is timed as
well ("full_run").

The same line carries the other configurations as sub-objects, each with its
own fixed iteration count (independent of --steps):
  c3  n = 20000 uniform, m = 1024, k = 32, Ant System with fp32 atomic
      deposits, 300 iterations: ms per iteration;
  c4  4096 uneven instances (n log-uniform in [50, 500]), 300 iterations in
      one batch: instances/s vs the same instances solved one after another;
  c5  learned edge prior: wall time to a 2% gap on 256 uniform n = 1000
      instances, 1000 iterations, prior vs the best plain Ant System.

`--gpus N` with N > 1 is refused: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from acogpu_amd import capi, instances  # noqa: E402

PEAK_HBM_GBS = 8000.0      # MI355X HBM3E, vendor figure
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA, vendor figure
KERNEL_NAMES = {0: "k_construct_cand", 1: "k_construct_full", 2: "k_construct_staged"}
CPU_FLAGS = "gcc -O3 -ffp-contract=off -fno-fast-math"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def c2_params(n=1000, m=1000):
    return capi.default_params(
        variant=capi.VARIANT_RBAS, mode=capi.MODE_CANDIDATE, n=n, m=m, k=20, w=6,
        alpha=1.0, beta=3.0, rho=0.1, tau0_policy=capi.TAU0_RBAS, aging=1,
        min_age=30, max_age=50, local_perf=10, global_perf=10, min_pop=40, max_pop=100,
        seed=20241203)


def construct_bytes(n, k, tours, scanned):
    """Algorithmic bytes of one construction launch: per tour n steps of
    k candidate indices + k weights (8k B) and one tour write (4 B), one
    distance read per edge (4 B), plus 4 B per unvisited city scanned by
    fallback steps.  n, tours may be per-instance arrays (batches)."""
    return float(np.sum(np.asarray(tours, dtype=np.float64) * (np.asarray(n) * (8.0 * k + 4) + 4.0 * np.asarray(n)))
                 + 4.0 * scanned)


def hbm_roofline(kernel, bytes_launch, kernel_ms, traffic=None):
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kernel, "kernel_ms": float(kernel_ms),
            "bytes_per_launch": float(bytes_launch)}


def cpu_baseline(xy, iterations, threads=1):
    """The oracle's reference-semantics restatement (every city in NN order,
    sequential fp32 sums, farthest-unvisited fallback, aging, rank-based
    deposit) on the same instance: one colony per thread (independent seeds),
    `iterations` iterations each.  Returns (tours/s over all threads, wall
    seconds)."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    dist = orc.dist_from_int_coords(xy)
    cnn = orc.greedy_length(dist, 0)
    tau0 = float(np.float32(0.5 * 6 * 5 / (np.float64(np.float32(0.1)) * cnn)))

    def one(seed):
        return orc.run(dist, 1000, variant=orc.RBAS, mode=orc.MODE_REFSEQ, k=20, w=6, aging=True,
                       min_age=30, max_age=50, global_perf=10, min_pop=40,
                       iterations=iterations, alpha=1.0, beta=3.0, rho=0.1, tau0=tau0, seed=seed)

    t0 = time.perf_counter()
    if threads == 1:
        one(7)
    else:  # the C oracle runs without the GIL held (ctypes)
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, range(7, 7 + threads)))
    wall = time.perf_counter() - t0
    return 1000 * iterations * threads / wall, wall


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


# ---------------------------------------------------------------------------
def bench_c2(args, xy):
    n, m, k = 1000, 1000, 20
    p = c2_params(n, m)
    p.ants_per_block = args.ants_per_block
    log(f"bench: c2 colony n={n} m={m} k={k} RBAS w=6 aging")
    col = capi.Colony(p, xy=xy)
    col.initialize()
    cfg = col.config()
    settle = args.settle
    log(f"bench: c2 launch config {cfg}; settle {settle} + warmup {args.warmup} iterations")
    col.run(settle)
    col.run(args.warmup)
    log(f"bench: c2 timed {args.steps} iterations (hipGraph replay)")
    tours0 = col.counters()[2]
    t0 = time.perf_counter()
    ev_ms = col.run(args.steps)
    wall = time.perf_counter() - t0
    ib, gb, nv = col.stats()
    tours = col.counters()[2] - tours0  # retired ants build no tour
    tours_per_s = tours / wall
    log(f"bench: c2 {wall * 1e3 / args.steps:.4f} ms/iteration (events {ev_ms / args.steps:.4f}), "
        f"{tours / args.steps:.1f} tours/iteration, {tours_per_s:.4g} tours/s, best {gb}")

    # per-kernel durations: HIP events on the ctx stream
    fb0 = col.counters()
    kms = [float(x) for x in col.profile(args.profile_iters)]
    fb1 = col.counters()
    scanned = float(fb1[1] - fb0[1]) / args.profile_iters
    built = float(fb1[2] - fb0[2]) / args.profile_iters
    bytes_launch = construct_bytes(n, k, built, scanned)
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "construct_c2_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = hbm_roofline(KERNEL_NAMES[cfg["path"]], bytes_launch, kms[1], traffic)
    # latency floor: one wave repeating the step's dependent chain alone
    cyc, ns = col.probe_step(200000)
    actual_ns = kms[1] * 1e6 / (n - 1)
    roof["latency_floor"] = {"probe_ns_per_step": ns, "probe_cycles_per_step": cyc,
                             "kernel_ns_per_step": actual_ns, "frac": ns / actual_ns,
                             "what": "one wave alone repeating the staged step's dependent chain (LDS candidate "
                                     "read, ds_bpermute visited test, DPP scan, ballot, pick) vs the construction "
                                     "kernel's time per step (launch duration / (n - 1))"}
    log(f"bench: c2 kernel ms sweep {kms[0]:.4f} construct {kms[1]:.4f} stats {kms[2]:.4f} deposit {kms[3]:.4f}; "
        f"fallback steps/launch {(fb1[0] - fb0[0]) / args.profile_iters:.0f}, scanned/launch {scanned:.0f}; "
        f"step probe {ns:.1f} ns ({cyc:.0f} cycles) vs {actual_ns:.1f} ns per step")
    col.close()

    # the whole 2000-iteration run of the configuration on a fresh colony
    full = None
    if args.full_run_iters > 0:
        colr = capi.Colony(p, xy=xy)
        colr.initialize()
        colr.run(20)  # builds the iteration graphs
        colr.initialize()
        t0 = time.perf_counter()
        colr.run(args.full_run_iters)
        wall_r = time.perf_counter() - t0
        tr = colr.counters()[2]
        full = {"iterations": args.full_run_iters, "ms_per_iteration": wall_r * 1e3 / args.full_run_iters,
                "tours_per_s": tr / wall_r, "tours_per_iteration": tr / args.full_run_iters,
                "best_length": colr.stats()[1]}
        colr.close()
        log(f"bench: c2 whole run {full}")

    # the same colony with the reference's full-row candidate set on the GPU
    fullrow = None
    if args.fullrow_iters > 0:
        pf = c2_params(n, m)
        pf.mode = capi.MODE_FULLROW
        colf = capi.Colony(pf, xy=xy)
        colf.initialize()
        colf.run(2)
        tf0 = colf.counters()[2]
        t0 = time.perf_counter()
        colf.run(args.fullrow_iters)
        wall_f = time.perf_counter() - t0
        tours_f = colf.counters()[2] - tf0
        colf.close()
        fullrow = {"value": tours_f / wall_f, "unit": "tours/s", "iterations": args.fullrow_iters}
        log(f"bench: c2 full-row GPU {fullrow['value']:.4g} tours/s")
    return {
        "value": tours_per_s, "ms_per_step": wall * 1e3 / args.steps,
        "config": {"workload": "rank-based AS + aging, n=1000 clustered (10 Gaussian clusters, sigma 400), "
                               "m=1000, k=20, w=6, alpha 1, beta 3, rho 0.1",
                   "n": n, "m": m, "k": k, "iterations_timed": args.steps, "settle_iterations": settle,
                   "tours_per_iteration": tours / args.steps, "valid_ants_after": nv,
                   "construct_path": cfg["path"], "ants_per_block": cfg["ants_per_block"]},
        "roofline": roof,
        "kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2], "deposit": kms[3]},
        "fullrow_gpu": fullrow, "best_length": gb, "full_run": full,
    }


def bench_c3(args):
    """n = 20000 cities uniform in [0,100000]^2, m = 1024, k = 32, Ant System
    with fp32 atomic deposits (deposit_path 1), alpha 1, beta 2, rho 0.2,
    tau0 = m / C_nn, 300 iterations (the first 10 untimed)."""
    n, m, k, iters, warm = 20000, 1024, 32, args.c3_iters, 10
    xy = instances.uniform(n, 100000, seed=3)
    out = {}
    for path in (1, 2):
        p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.2, variant=capi.VARIANT_AS,
                                tau0_policy=capi.TAU0_AS, seed=5, deposit_path=path)
        col = capi.Colony(p, xy=xy)
        t0 = time.perf_counter()
        col.initialize()
        init_s = time.perf_counter() - t0
        if path == 2:  # secondary figure: the adjacency-gather deposit, a shorter run
            col.run(warm)
            kms = [float(x) for x in col.profile(3)]
            col.close()
            out["adjacency_deposit"] = {"kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2],
                                                      "deposit": kms[3]}}
            log(f"bench: c3 adjacency deposit kernels {kms}")
            break
        col.run(warm)
        t0 = time.perf_counter()
        col.run(iters - warm)
        wall = time.perf_counter() - t0
        fb0 = col.counters()
        kms = [float(x) for x in col.profile(3)]
        fb1 = col.counters()
        ib, gb, _ = col.stats()
        col.close()
        scanned = (fb1[1] - fb0[1]) / 3
        built = (fb1[2] - fb0[2]) / 3
        roof = hbm_roofline("k_construct_cand", construct_bytes(n, k, built, scanned), kms[1])
        pitch = (n + 63) // 64 * 64
        # HBM per element: tau read, choice and evaporated tau written (the
        # distances come from the coordinates, L2-resident: 8 B of L2 reads)
        sweep_bytes = 12.0 * n * pitch
        roof["sweep"] = {"bytes_per_launch": sweep_bytes, "kernel_ms": kms[0],
                         "achieved": sweep_bytes / (kms[0] * 1e-3) / 1e9, "frac": sweep_bytes / (kms[0] * 1e-3) / 1e9
                         / PEAK_HBM_GBS, "bytes_per_element": "4 read + 8 written (distances from coordinates)"}
        ms_it = wall * 1e3 / (iters - warm)
        log(f"bench: c3 {ms_it:.3f} ms/iteration, kernels {kms}, fallbacks/launch {(fb1[0] - fb0[0]) / 3:.0f} "
            f"({(fb1[3] - fb0[3]) / 3:.0f} from row top lists), init {init_s:.2f} s")
        out.update({
            "metric": "ms per iteration at n=20000", "value": ms_it, "unit": "ms", "higher_is_better": False,
            "ms_per_step": ms_it, "steps": iters - warm, "warmup": warm, "tours_per_s": m / (ms_it * 1e-3),
            "config": {"workload": "n=20000 uniform [0,100000]^2, m=1024, k=32 + row-scan fallback, Ant System "
                                   "with fp32 atomic deposits, alpha 1, beta 2, rho 0.2, 300 iterations",
                       "deposit_path": "fp32 atomics"},
            "roofline": roof,
            "kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2], "deposit": kms[3]},
            "init_seconds": init_s, "best_length": gb,
        })
    return out


def bench_c4(args):
    """4096 independent instances (n log-uniform in [50, 500], integer
    coordinates uniform in [0,1000]^2, m = n, k = 16, Ant System, alpha 1,
    beta 2, rho 0.5, tau0 = m / C_nn), 300 iterations each: one batch with the
    persistent construction launch vs. the same instances solved one after
    another with the single-instance solver (a sample of them, same seeds).
    Both timed regions replay already-built graphs from initialised state."""
    count, iters = args.batch_count, args.c4_iters
    sizes = instances.batched_sizes(count, 50, 500, seed=5)
    xys = [instances.uniform(int(n), 1000, seed=500000 + i) for i, n in enumerate(sizes)]
    p = capi.default_params(n=0, m=0, k=16, beta=2.0, rho=0.5, variant=capi.VARIANT_AS,
                            tau0_policy=capi.TAU0_AS, seed=99)
    log(f"bench: c4 batch of {count} instances, {int(sizes.sum())} cities, {iters} iterations")
    b = capi.Batch(p, sizes, xys)
    b.initialize()
    w = b.autotune(reps=2)
    b.initialize()
    b.run(4)  # builds the iteration graph
    b.initialize()
    cfg = b.config()
    t0 = time.perf_counter()
    b.run(iters)
    wall = time.perf_counter() - t0
    gb, _ = b.best()
    c0 = b.counters().sum(axis=0)
    kms = [float(x) for x in b.profile(3)]
    c1 = b.counters().sum(axis=0)
    b.close()
    batch_rate = count / wall
    ms_it = wall * 1e3 / iters
    roof = hbm_roofline("kb_construct_group" if cfg["ants_per_wave"] > 1 else "kb_construct_vreg",
                        construct_bytes(sizes, 16, sizes, float(c1[1] - c0[1]) / 3), kms[1])
    log(f"bench: c4 batched {wall:.3f} s ({ms_it:.3f} ms/iteration) -> {batch_rate:.1f} instances/s; kernels {kms}; "
        f"config {cfg}")
    rng = np.random.default_rng(1)
    sample = rng.choice(count, size=min(args.seq_sample, count), replace=False)
    seq_time = 0.0
    for i in sample:
        n = int(sizes[i])
        q = capi.default_params(n=n, m=n, k=16, beta=2.0, rho=0.5, variant=capi.VARIANT_AS,
                                tau0_policy=capi.TAU0_AS, seed=99, stream_id=int(i))
        c = capi.Colony(q, xy=xys[i])
        c.initialize()
        c.run(20)  # builds both iteration graphs
        c.initialize()
        t0 = time.perf_counter()
        c.run(iters)
        seq_time += time.perf_counter() - t0
        c.close()
    seq_rate = len(sample) / seq_time
    log(f"bench: c4 sequential {len(sample)} instances {seq_time:.3f} s -> {seq_rate:.1f} instances/s")
    tuned = None
    if args.tune_k:
        ks = [8, 12, 16, 20, 24]
        k, rep = capi.tune_k(p, sizes, xys, ks, trial_iterations=20, tolerance=0.005)
        tuned = {"chosen_k": k, "trial_iterations": 20, "tolerance": 0.005,
                 "per_k": {str(kk): {"ms_per_iteration": float(r[0]), "mean_best": float(r[1]),
                                     "ants_per_block": int(r[2])} for kk, r in zip(ks, rep)}}
        log(f"bench: c4 candidate-width tuner picks k={k}")
    return {
        "metric": "instances/sec, batched uneven instances", "value": batch_rate, "unit": "instances/s",
        "higher_is_better": True, "ms_per_step": ms_it, "steps": iters, "warmup": 0,
        "config": {"workload": f"{count} instances, n log-uniform [50,500], m=n, k=16, AS, alpha 1, beta 2, rho 0.5, "
                               f"{iters} iterations", "ants_per_block": w, "ants_per_wave": cfg["ants_per_wave"]},
        "roofline": roof,
        "kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2], "deposit": kms[3]},
        "sequential": {"value": seq_rate, "unit": "instances/s", "sample": int(len(sample)),
                       "kind": "single-instance solver, one instance after another, same seeds"},
        "speedup_vs_sequential": batch_rate / seq_rate,
        "mean_best_length": float(np.mean(gb)),
        "candidate_width_tuner": tuned,
    }


def time_to_gap(curve, ref, gap=0.02):
    """Per instance, the solver wall time at the first checkpoint whose
    best-so-far length is within `gap` of ref (inf if never)."""
    reach = np.full(len(ref), np.inf)
    for _, secs, gb in curve:
        hit = (gb <= (1 + gap) * ref) & np.isinf(reach)
        reach[hit] = secs
    return reach


def quantile_or_none(x, q):
    v = float(np.quantile(x, q)) if np.isfinite(x).mean() >= q else float("inf")
    return v if np.isfinite(v) else None


def bench_c5(args):
    """Learned edge prior: wall time to a 2% gap, prior vs plain Ant System.

    Training: 2000 instances n=200 (integer coordinates uniform in
    [0,1000]^2), batched Ant System m=n, k=20, alpha 1, beta 5, rho 0.5,
    tau0 = m/C_nn, 200 iterations; every candidate edge labelled by the
    best-so-far tours; 12 epochs of Adam (minibatch 65536) on the device.
    Evaluation: 256 instances n=1000 uniform in [0,1000]^2, m=1000, k=20,
    alpha 1, rho 0.5, 1000 iterations of (a) plain AS beta 2, (b) plain AS
    beta 5, (c) AS whose candidate-edge heuristic is the prior's exp(3 score)
    (fallback steps eta^5).  Best-known per instance = shortest tour of the
    three runs.  Time to 2% = solver wall time (for (c) including the prior's
    feature + forward pass) until best-so-far <= 1.02 best-known;
    best-so-far lengths are read after every iteration up to 50, then every
    10 iterations, outside the clock."""
    nt, ne, iters = 2000, args.prior_instances, args.c5_iters
    t0 = time.perf_counter()
    sizes = np.full(nt, 200, np.int32)
    xys = [instances.uniform(200, 1000, seed=700000 + i) for i in range(nt)]
    tp = capi.default_params(n=0, m=0, k=20, beta=5.0, rho=0.5, variant=capi.VARIANT_AS,
                             tau0_policy=capi.TAU0_AS, seed=3)
    tb = capi.Batch(tp, sizes, xys)
    tb.initialize()
    tb.run(200)
    gen_s = time.perf_counter() - t0
    pr = capi.Prior(capi.default_prior_config())
    t0 = time.perf_counter()
    pr.collect_batch(tb)
    tb.close()
    loss = pr.train(12)
    train_s = time.perf_counter() - t0
    fwd_ms, bwd_ms, edges = pr.profile(20)
    log(f"bench: c5 training data {gen_s:.2f} s, {pr.size()} edges, training {train_s:.2f} s, "
        f"loss {loss[0]:.4f} -> {loss[-1]:.4f}, forward {fwd_ms:.4f} ms / backward {bwd_ms:.4f} ms per {edges} edges")
    esz = np.full(ne, 1000, np.int32)
    exy = [instances.uniform(1000, 1000, seed=900000 + i) for i in range(ne)]
    solvers = [("plain AS beta 2", 2.0, False), ("plain AS beta 5", 5.0, False), ("AS + learned prior", 5.0, True)]
    curves, ms_it, kprof = {}, {}, {}
    for label, beta, use_prior in solvers:
        ep = capi.default_params(n=0, m=0, k=20, beta=beta, rho=0.5, variant=capi.VARIANT_AS,
                                 tau0_policy=capi.TAU0_AS, seed=11)
        e = capi.Batch(ep, esz, exy)
        e.initialize()
        e.run(2)  # builds the iteration graph
        e.initialize()
        el = 0.0
        if use_prior:
            t1 = time.perf_counter()
            pr.apply_batch(e)
            el += time.perf_counter() - t1
        cur, it = [], 0
        run_s = 0.0
        while it < iters:
            c = 1 if it < 50 else min(10, iters - it)
            t1 = time.perf_counter()
            e.run(c)
            dt = time.perf_counter() - t1
            el += dt
            run_s += dt
            it += c
            cur.append((it, el, e.best_lengths()))
            if it % 250 == 0:
                log(f"bench: c5 {label} iteration {it}, {el:.2f} s, mean best {cur[-1][2].mean():.1f}")
        if use_prior:
            c0 = e.counters().sum(axis=0)
            kprof["kernel_ms"] = [float(x) for x in e.profile(2)]
            c1 = e.counters().sum(axis=0)
            kprof["scanned"] = float(c1[1] - c0[1]) / 2
            kprof["apw"] = e.config()["ants_per_wave"]
        curves[label] = cur
        ms_it[label] = run_s * 1e3 / iters
        e.close()
    pr.close()
    ref = np.min(np.stack([c[-1][2] for c in curves.values()]), axis=0)
    table = {}
    for label, cur in curves.items():
        reach = time_to_gap(cur, ref)
        table[label] = {"t50_s": quantile_or_none(reach, 0.5), "t90_s": quantile_or_none(reach, 0.9),
                        "t100_s": quantile_or_none(reach, 1.0), "frac_within_2pct": float(np.isfinite(reach).mean()),
                        "ms_per_iteration": ms_it[label],
                        "final_mean_gap_pct": float((cur[-1][2] / ref - 1).mean() * 100)}
        log(f"bench: c5 {label}: {table[label]}")
    prior = table["AS + learned prior"]

    def speedup(key):
        plain = [table[k][key] for k in ("plain AS beta 2", "plain AS beta 5")]
        best = min((t for t in plain if t is not None), default=None)
        return best / prior[key] if (best and prior[key]) else None

    flop_fwd, flop_bwd = 2 * (8 * 64 + 64 * 64 + 64), 2 * (64 * 64 * 2 + 64 * 8 + 64)
    kms = kprof["kernel_ms"]
    roof = hbm_roofline("kb_construct_group" if kprof["apw"] > 1 else "kb_construct_vreg",
                        construct_bytes(esz, 20, esz, kprof["scanned"]), kms[1])
    bw_tflops = (flop_fwd + flop_bwd) * edges / (bwd_ms * 1e9)
    return {
        "metric": "wall-seconds to 2% gap, learned prior vs plain AS", "value": prior["t50_s"], "unit": "s",
        "higher_is_better": False, "steps": iters, "warmup": 0, "ms_per_step": prior["ms_per_iteration"],
        "dtype": "bf16 MLP (fp32 accumulate) / f32 ACO",
        "config": {"workload": f"{ne} instances n=1000 uniform [0,1000]^2, m=1000, k=20, AS alpha 1 rho 0.5, "
                               f"{iters} iterations; prior trained on 2000 n=200 AS runs (200 iterations)",
                   "statistic": "median over instances of the solver wall time until the instance is within 2% "
                                "of its best-known (shortest tour of all three runs)"},
        "solvers": table,
        "speedup_t50_vs_best_plain": speedup("t50_s"),
        "speedup_t90_vs_best_plain": speedup("t90_s"),
        "roofline": roof,
        "kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2], "deposit": kms[3]},
        "training": {"data_seconds": gen_s, "train_seconds": train_s, "edges": int(nt * 200 * 20),
                     "epochs": 12, "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
        "mlp_roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_BF16_TFLOPS, "kernel": "k_mlp_backward",
                         "achieved": bw_tflops, "frac": bw_tflops / PEAK_BF16_TFLOPS, "edges_per_launch": edges,
                         "backward_ms": bwd_ms, "forward_ms": fwd_ms,
                         "forward_tflops": flop_fwd * edges / (fwd_ms * 1e9)},
    }


def bench_ant_scaling(args):
    """Latency check of the construction kernel: the n=1000 rank-based
    workload (no aging, so every ant is an ordinary staged ant) with m = 64,
    256, 512 and 1000 ants; construction time per launch (HIP events, mean of
    `profile_iters` launches after 30 iterations)."""
    n = 1000
    xy = instances.clustered(n, clusters=10, side=10000, sigma=400.0, seed=1)
    rows = {}
    for m in (64, 256, 512, 1000):
        p = c2_params(n, m)
        p.aging = 0
        col = capi.Colony(p, xy=xy)
        col.initialize()
        col.run(30)
        k = [float(x) for x in col.profile(args.profile_iters)]
        cfg = col.config()
        col.close()
        rows[str(m)] = {"construct_ms": k[1], "us_per_step": k[1] * 1e3 / (n - 1),
                        "ants_per_block": cfg["ants_per_block"]}
        log(f"bench: m={m}: construct {k[1] * 1e3:.1f} us ({cfg})")
    base = rows["64"]["construct_ms"]
    return {
        "metric": "construction time vs ants, n=1000", "value": rows["1000"]["construct_ms"] / base, "unit": "x",
        "n_gpus": 1, "steps": args.profile_iters, "warmup": 30, "ms_per_step": rows["1000"]["construct_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "rank-based AS, n=1000 clustered, k=20, w=6, beta 3, rho 0.1, no aging; m varied",
                   "value_is": "construct time at m=1000 / at m=64"},
        "per_m": rows,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed iterations of the n=1000 line")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--settle", type=int, default=500,
                    help="n=1000 line: iterations run before the warm-up (aging retires ants from iteration 50 on)")
    ap.add_argument("--full-run-iters", type=int, default=2000,
                    help="n=1000 line: also time a whole run of this many iterations (0 = off)")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--fullrow-iters", type=int, default=10, help="0 skips the full-row GPU figure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ants-per-block", type=int, default=0)
    ap.add_argument("--only", choices=["c2", "c3", "c4", "c5", "ant-scaling"], default=None,
                    help="run one configuration and print its line alone (default: all, one line)")
    ap.add_argument("--skip", default="", help="comma-separated sub-objects to leave out of the default line")
    ap.add_argument("--c3-iters", type=int, default=300)
    ap.add_argument("--c4-iters", type=int, default=300)
    ap.add_argument("--c5-iters", type=int, default=1000)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--prior-instances", type=int, default=256)
    ap.add_argument("--tune-k", type=int, default=1, help="c4: also run the candidate-width tuner (0 = off)")
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances and ants shard without exchange; see the sharding notes)", file=sys.stderr)
        return 2

    common = {"n_gpus": 1, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    if args.only in ("c3", "c4", "c5", "ant-scaling"):
        fn = {"c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "ant-scaling": bench_ant_scaling}[args.only]
        line = {**common, **fn(args)}
        print(json.dumps(line), flush=True)
        return 0

    xy = instances.clustered(1000, clusters=10, side=10000, sigma=400.0, seed=1)
    if args.calibrate_copy:
        col = capi.Colony(c2_params(64, 8), xy=xy[:64])
        ms = col.calibrate_copy(args.calibrate_copy, reps=5)
        log(f"bench: copy of {args.calibrate_copy} bytes: {ms:.4f} ms, "
            f"{2 * args.calibrate_copy / (ms * 1e-3) / 1e9:.1f} GB/s (read + write)")
        col.close()
        return 0
    # CPU baselines first, before the GPU is touched
    cpu = cpu_multi = None
    model, ncpu = host_cpu()
    if not args.no_cpu_baseline:
        log(f"bench: host {model}, {ncpu} cores available")
        log(f"bench: CPU baseline, {args.cpu_iters} iterations x 1000 ants, reference semantics, 1 thread")
        rate, secs = cpu_baseline(xy, args.cpu_iters)
        cpu = {"value": rate, "unit": "tours/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iters} iterations x 1000 ants of the same n=1000 instance, reference "
                         f"full-row semantics (oracle refseq mode, {CPU_FLAGS}), {secs:.1f} s; host {model}, "
                         f"{ncpu} cores available"}
        log(f"bench: CPU {rate:.4g} tours/s ({secs:.1f} s)")
        threads = min(16, ncpu)
        if threads > 1:
            rate_n, secs_n = cpu_baseline(xy, 1, threads)
            cpu_multi = {"value": rate_n, "unit": "tours/s", "cores": threads,
                         "sample": f"{threads} independent colonies (one per thread), 1 iteration x 1000 ants each, "
                                   f"{secs_n:.1f} s ({CPU_FLAGS})"}
            log(f"bench: CPU {threads} threads {rate_n:.4g} tours/s ({secs_n:.1f} s)")

    c2 = bench_c2(args, xy)
    line = {
        "metric": "tours/sec at n=1000", "value": c2["value"], "unit": "tours/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": c2["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": c2["config"], "roofline": c2["roofline"], "kernel_ms": c2["kernel_ms"],
        "cpu_baseline": cpu, "cpu_multicore": cpu_multi, "fullrow_gpu": c2["fullrow_gpu"],
        "full_run": c2["full_run"], "best_length": c2["best_length"],
    }
    if args.only != "c2":
        skip = {s.strip() for s in args.skip.split(",") if s.strip()}
        for key, fn in (("c3", bench_c3), ("c4", bench_c4), ("c5", bench_c5)):
            if key in skip:
                continue
            t0 = time.perf_counter()
            line[key] = fn(args)
            line[key]["wall_seconds"] = time.perf_counter() - t0
            log(f"bench: {key} done in {line[key]['wall_seconds']:.1f} s")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
