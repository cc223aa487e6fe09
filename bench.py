#!/usr/bin/env python3
"""Benchmark: tours/s of the ACO hot path on one MI355X.

Workload (default): rank-based Ant System with ant aging, n = 1000 cities
from 10 Gaussian clusters (centres uniform in [0,10000]^2, sigma 400, clipped,
rounded), EUC_2D integer distances, m = 1000 ants, candidate lists k = 20,
w = 6 ranked ants, alpha 1, beta 3, rho 0.1, fp32 pheromone / choice_info.
One step = one ACO iteration (sweep, construction, statistics, deposit), all
on the device; the timed loop replays a hipGraph of the iteration with the
inputs resident in HBM.  tours/s counts the tours actually constructed: every
valid ant builds one per iteration, and aging retires ants (about 840 of the
1000 stay valid once the colony has converged).

Prints one JSON line.  `--gpus N` with N > 1 is refused: this build measures
one GPU.
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

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, vendor figure
KERNEL_NAMES = {0: "k_construct_cand", 1: "k_construct_full", 2: "k_construct_staged"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def c2_params(n=1000, m=1000):
    return capi.default_params(
        variant=capi.VARIANT_RBAS, mode=capi.MODE_CANDIDATE, n=n, m=m, k=20, w=6,
        alpha=1.0, beta=3.0, rho=0.1, tau0_policy=capi.TAU0_RBAS, aging=1,
        min_age=30, max_age=50, local_perf=10, global_perf=10, min_pop=40, max_pop=100,
        seed=20241203)


def construct_bytes_per_launch(n, k, tours, scanned):
    """Algorithmic bytes of one construction launch: per tour n steps of
    k candidate indices + k weights (8k B) and one tour write (4 B), one
    distance read per edge (4 B), plus 4 B per unvisited city scanned by
    fallback steps."""
    return tours * (n * (8 * k + 4) + 4 * n) + 4 * scanned


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


def bench_batched(args):
    """4096 independent instances (n log-uniform in [50, 500], integer
    coordinates uniform in [0,1000]^2, m = n, k = 16, Ant System, alpha 1,
    beta 2, rho 0.5, tau0 = m / C_nn), `steps` iterations each: one batch with
    the persistent construction launch vs. the same instances solved one after
    another with the single-instance solver (a sample of them, same seeds)."""
    count, iters = args.batch_count, args.steps
    sizes = instances.batched_sizes(count, 50, 500, seed=5)
    xys = [instances.uniform(int(n), 1000, seed=500000 + i) for i, n in enumerate(sizes)]
    p = capi.default_params(n=0, m=0, k=16, beta=2.0, rho=0.5, variant=capi.VARIANT_AS,
                            tau0_policy=capi.TAU0_AS, seed=99)
    log(f"bench: batch of {count} instances, {int(sizes.sum())} cities, {iters} iterations")
    b = capi.Batch(p, sizes, xys)
    b.initialize()
    w = b.autotune(reps=2)
    b.initialize()
    cfg = b.config()
    log(f"bench: autotuned ants per block {w}, config {cfg}")
    b.run(args.warmup)
    b.initialize()
    t0 = time.perf_counter()
    b.run(iters)
    wall = time.perf_counter() - t0
    gb, _ = b.best()
    b.close()
    batch_rate = count / wall
    log(f"bench: batched {wall:.3f} s -> {batch_rate:.1f} instances/s")
    rng = np.random.default_rng(1)
    sample = rng.choice(count, size=min(args.seq_sample, count), replace=False)
    seq_time = 0.0
    for i in sample:
        n = int(sizes[i])
        q = capi.default_params(n=n, m=n, k=16, beta=2.0, rho=0.5, variant=capi.VARIANT_AS,
                                tau0_policy=capi.TAU0_AS, seed=99, stream_id=int(i))
        c = capi.Colony(q, xy=xys[i])
        c.initialize()
        c.run(2)
        c.initialize()
        t0 = time.perf_counter()
        c.run(iters)
        seq_time += time.perf_counter() - t0
        c.close()
    seq_rate = len(sample) / seq_time
    log(f"bench: sequential {len(sample)} instances {seq_time:.3f} s -> {seq_rate:.1f} instances/s")
    tuned = None
    if args.tune_k:
        ks = [8, 12, 16, 20, 24]
        k, rep = capi.tune_k(p, sizes, xys, ks, trial_iterations=20, tolerance=0.005)
        tuned = {"chosen_k": k, "trial_iterations": 20, "tolerance": 0.005,
                 "per_k": {str(kk): {"ms_per_iteration": float(r[0]), "mean_best": float(r[1]),
                                     "ants_per_block": int(r[2])} for kk, r in zip(ks, rep)}}
        log(f"bench: candidate-width tuner picks k={k}: {tuned['per_k']}")
    return {
        "metric": "instances/sec, batched uneven instances", "value": batch_rate, "unit": "instances/s",
        "n_gpus": 1, "steps": iters, "warmup": args.warmup, "ms_per_step": wall * 1e3 / iters,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{count} instances, n log-uniform [50,500], m=n, k=16, AS, beta 2, rho 0.5",
                   "iterations": iters, "ants_per_block": w, "ants_per_wave": cfg["ants_per_wave"]},
        "sequential": {"value": seq_rate, "unit": "instances/s", "sample": int(len(sample)),
                       "kind": "single-instance solver, one instance after another"},
        "speedup_vs_sequential": batch_rate / seq_rate,
        "mean_best_length": float(np.mean(gb)),
        "candidate_width_tuner": tuned,
    }


def bench_large(args):
    """n = 20000 cities uniform in [0,100000]^2, m = 1024, k = 32, Ant System
    with fp32 atomic deposits, alpha 1, beta 2, rho 0.2, tau0 = m / C_nn."""
    n, m = 20000, 1024
    xy = instances.uniform(n, 100000, seed=3)
    p = capi.default_params(n=n, m=m, k=32, beta=2.0, rho=0.2, variant=capi.VARIANT_AS,
                            tau0_policy=capi.TAU0_AS, seed=5)
    col = capi.Colony(p, xy=xy)
    col.initialize()
    col.run(args.warmup)
    t0 = time.perf_counter()
    col.run(args.steps)
    wall = time.perf_counter() - t0
    fb0 = col.counters()
    kms = [float(x) for x in col.profile(2)]
    fb1 = col.counters()
    ib, gb, _ = col.stats()
    col.close()
    log(f"bench: n=20000 {wall * 1e3 / args.steps:.2f} ms/iteration, kernels {kms}, "
        f"fallback steps/launch {(fb1[0] - fb0[0]) / 2:.0f} ({(fb1[3] - fb0[3]) / 2:.0f} answered by row top lists)")
    return {
        "metric": "tours/sec at n=20000", "value": m * args.steps / wall, "unit": "tours/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "n=20000 uniform [0,100000]^2, m=1024, k=32, AS, beta 2, rho 0.2"},
        "kernel_ms": {"sweep": kms[0], "construct": kms[1], "stats": kms[2], "deposit": kms[3]},
        "best_length": gb,
    }


def bench_ant_scaling(args):
    """Latency check of the construction kernel: the n=1000 rank-based
    workload (no aging, so every ant is an ordinary staged ant) with m = 64,
    256, 512 and 1000 ants; construction time per launch (HIP events, mean of
    `profile_iters` launches after 30 iterations).  If the kernel were
    throughput-bound the time would grow with m; it stays near one ant's
    tour time (a chain of n dependent steps)."""
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


def bench_prior(args):
    """Learned edge prior: wall time to a 2% gap, prior vs plain Ant System.

    Training: 2000 instances n=200 (integer coordinates uniform in
    [0,1000]^2), batched Ant System m=n, k=20, alpha 1, beta 5, rho 0.5,
    tau0 = m/C_nn, 200 iterations; every candidate edge labelled by the
    best-so-far tours; 12 epochs of Adam (minibatch 65536) on the device.
    Evaluation: 256 instances n=1000 uniform in [0,1000]^2, m=1000, k=20,
    alpha 1, rho 0.5, `steps` iterations of (a) plain AS beta 2, (b) plain AS
    beta 5, (c) AS whose candidate-edge heuristic is the prior's exp(3 score)
    (fallback steps eta^5).  Best-known per instance = shortest tour of the
    three long runs.  Time to 2% = solver wall time (for (c) including the
    prior's feature + forward pass) until best-so-far <= 1.02 best-known;
    best-so-far lengths are read after every iteration up to 50, then every
    10 iterations, outside the clock."""
    nt, ne, iters = 2000, args.prior_instances, args.steps
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
    log(f"bench: training data {gen_s:.2f} s, {pr.size()} edges, training {train_s:.2f} s, "
        f"loss {loss[0]:.4f} -> {loss[-1]:.4f}, forward {fwd_ms:.3f} ms / backward {bwd_ms:.3f} ms per {edges} edges")
    esz = np.full(ne, 1000, np.int32)
    exy = [instances.uniform(1000, 1000, seed=900000 + i) for i in range(ne)]
    solvers = [("plain AS beta 2", 2.0, False), ("plain AS beta 5", 5.0, False), ("AS + learned prior", 5.0, True)]
    curves = {}
    ms_it = {}
    for label, beta, use_prior in solvers:
        ep = capi.default_params(n=0, m=0, k=20, beta=beta, rho=0.5, variant=capi.VARIANT_AS,
                                 tau0_policy=capi.TAU0_AS, seed=11)
        e = capi.Batch(ep, esz, exy)
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
            if it % 100 == 0:
                log(f"bench: {label} iteration {it}, {el:.2f} s, mean best {cur[-1][2].mean():.1f}")
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
        log(f"bench: {label}: {table[label]}")
    prior = table["AS + learned prior"]

    def speedup(key):
        plain = [table[k][key] for k in ("plain AS beta 2", "plain AS beta 5")]
        best = min((t for t in plain if t is not None), default=None)
        return best / prior[key] if (best and prior[key]) else None

    flop_fwd, flop_bwd = 2 * (8 * 64 + 64 * 64 + 64), 2 * (64 * 64 * 2 + 64 * 8 + 64)
    return {
        "metric": "wall-seconds to 2% gap, learned prior vs plain AS", "value": prior["t50_s"], "unit": "s",
        "n_gpus": 1, "steps": iters, "warmup": 0, "ms_per_step": prior["ms_per_iteration"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 MLP / f32 ACO",
        "data": "synthetic",
        "config": {"workload": f"{ne} instances n=1000 uniform [0,1000]^2, m=1000, k=20, AS alpha 1 rho 0.5, "
                               f"{iters} iterations; prior trained on 2000 n=200 AS runs (200 iterations)",
                   "statistic": "median over instances of the solver wall time until the instance is within 2% "
                                "of its best-known (shortest tour of all three long runs)"},
        "solvers": table,
        "speedup_t50_vs_best_plain": speedup("t50_s"),
        "speedup_t90_vs_best_plain": speedup("t90_s"),
            "training": {"data_seconds": gen_s, "train_seconds": train_s, "edges": int(nt * 200 * 20),
                     "epochs": 12, "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
        "mlp_kernels": {"edges_per_launch": edges, "forward_ms": fwd_ms, "backward_ms": bwd_ms,
                        "forward_tflops": flop_fwd * edges / (fwd_ms * 1e9),
                        "backward_tflops": (flop_fwd + flop_bwd) * edges / (bwd_ms * 1e9),
                        "peak_bf16_tflops": 2500.0},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--fullrow-iters", type=int, default=10, help="0 skips the full-row GPU figure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ants-per-block", type=int, default=0)
    ap.add_argument("--workload", choices=["c2", "batched", "large", "prior", "ant-scaling"], default="c2",
                    help="c2: the headline n=1000 tours/s line (default); batched: 4096 uneven instances; "
                         "large: n=20000; prior: learned edge prior time to 2% gap; ant-scaling: "
                         "construction time vs number of ants")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--prior-instances", type=int, default=256)
    ap.add_argument("--tune-k", type=int, default=1, help="batched: also run the candidate-width tuner (0 = off)")
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances and ants shard without exchange; see the sharding notes)", file=sys.stderr)
        return 2

    if args.workload != "c2":
        line = {"batched": bench_batched, "large": bench_large, "prior": bench_prior,
                "ant-scaling": bench_ant_scaling}[args.workload](args)
        print(json.dumps(line), flush=True)
        return 0

    n, m, k = 1000, 1000, 20
    xy = instances.clustered(n, clusters=10, side=10000, sigma=400.0, seed=1)
    if args.calibrate_copy:
        col = capi.Colony(c2_params(64, 8), xy=xy[:64])
        ms = col.calibrate_copy(args.calibrate_copy, reps=5)
        log(f"bench: copy of {args.calibrate_copy} bytes: {ms:.4f} ms, "
            f"{2 * args.calibrate_copy / (ms * 1e-3) / 1e9:.1f} GB/s (read + write)")
        col.close()
        return 0
    # CPU baselines first, before the GPU is touched
    cpu = None
    cpu_multi = None
    model, ncpu = host_cpu()
    if not args.no_cpu_baseline:
        log(f"bench: host {model}, {ncpu} cores available")
        log(f"bench: CPU baseline, {args.cpu_iters} iterations x {m} ants, reference semantics, 1 thread")
        rate, secs = cpu_baseline(xy, args.cpu_iters)
        cpu = {"value": rate, "unit": "tours/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iters} iterations x {m} ants of the same n=1000 instance, "
                         f"reference full-row semantics (oracle refseq mode), {secs:.1f} s; host {model}, "
                         f"{ncpu} cores available"}
        log(f"bench: CPU {rate:.4g} tours/s ({secs:.1f} s)")
        threads = min(16, ncpu)
        if threads > 1:
            rate_n, secs_n = cpu_baseline(xy, 1, threads)
            cpu_multi = {"value": rate_n, "unit": "tours/s", "cores": threads,
                         "sample": f"{threads} independent colonies (one per thread), 1 iteration x {m} ants each, "
                                   f"{secs_n:.1f} s"}
            log(f"bench: CPU {threads} threads {rate_n:.4g} tours/s ({secs_n:.1f} s)")

    p = c2_params(n, m)
    p.ants_per_block = args.ants_per_block
    log(f"bench: creating colony n={n} m={m} k={k} RBAS w=6 aging")
    col = capi.Colony(p, xy=xy)
    col.initialize()
    cfg = col.config()
    log(f"bench: launch config {cfg}")
    log(f"bench: warmup {args.warmup} iterations")
    col.run(args.warmup)
    log(f"bench: timed {args.steps} iterations (hipGraph replay)")
    tours0 = col.counters()[2]
    t0 = time.perf_counter()
    ev_ms = col.run(args.steps)
    wall = time.perf_counter() - t0
    ib, gb, nv = col.stats()
    # tours actually constructed: with aging, retired (invalid) ants build none
    tours = col.counters()[2] - tours0
    tours_per_s = tours / wall
    log(f"bench: {wall * 1e3 / args.steps:.3f} ms/iteration (events {ev_ms / args.steps:.3f}), "
        f"{tours_per_s:.4g} tours/s, best {gb}")

    # Per-kernel durations with HIP events on the ctx stream.
    fb0 = col.counters()
    kms = [float(x) for x in col.profile(args.profile_iters)]
    fb1 = col.counters()
    scanned_per_launch = float(fb1[1] - fb0[1]) / args.profile_iters
    fbsteps_per_launch = float(fb1[0] - fb0[0]) / args.profile_iters
    nvalid = col.stats()[2]
    bytes_launch = construct_bytes_per_launch(n, k, nvalid, scanned_per_launch)
    achieved = float(bytes_launch / (kms[1] * 1e-3) / 1e9)
    log(f"bench: kernel ms sweep {kms[0]:.4f} construct {kms[1]:.4f} stats {kms[2]:.4f} "
        f"deposit {kms[3]:.4f}; fallback steps/launch {fbsteps_per_launch:.0f}, "
        f"cities scanned/launch {scanned_per_launch:.0f}")

    traffic = None
    pmc = os.path.join(ROOT, "profiles", "construct_c2_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("bytes_per_launch")

    # The same colony with reference full-row semantics on the GPU (every
    # city in NN order each step: apples-to-apples work per tour).
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
        log(f"bench: full-row GPU {fullrow['value']:.4g} tours/s")

    line = {
        "metric": "tours/sec at n=1000",
        "value": tours_per_s,
        "unit": "tours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "rank-based AS + aging, n=1000 clustered (10 Gaussian clusters, "
                               "sigma 400), m=1000, k=20, w=6, alpha 1, beta 3, rho 0.1",
                   "n": n, "m": m, "k": k, "iterations_timed": args.steps,
                   "tours_per_iteration": tours / args.steps,
                   "construct_path": cfg["path"], "ants_per_block": cfg["ants_per_block"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAMES[cfg["path"]], "kernel_ms": float(kms[1]),
                     "bytes_per_launch": bytes_launch},
        "kernel_ms": {"sweep": float(kms[0]), "construct": float(kms[1]), "stats": float(kms[2]),
                      "deposit": float(kms[3])},
        "cpu_baseline": cpu,
        "cpu_multicore": cpu_multi,
        "fullrow_gpu": fullrow,
        "best_length": gb,
    }
    print(json.dumps(line), flush=True)
    col.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
