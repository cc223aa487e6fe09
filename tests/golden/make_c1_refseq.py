"""Writes c1_refseq.json: best tour lengths of the oracle's reference-semantics
run (every city in NN order, sequential fp32 sums, farthest-unvisited
fallback) over 30 Philox seeds on the n = 100 Ant System workload: integer
coordinates uniform in [0,1000]^2 (instance seed 1), m = 100, alpha 1, beta 2,
rho 0.5, tau0 = m / C_nn, 500 iterations.

Run from the repository root: python tests/golden/make_c1_refseq.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as orc  # noqa: E402
from acogpu_amd import instances  # noqa: E402

N, M, ITERS, SEEDS = 100, 100, 500, list(range(1, 31))


def c1_instance():
    xy = instances.uniform(N, 1000, seed=1)
    dist = orc.dist_from_int_coords(xy)
    tau0 = float(np.float32(M / orc.greedy_length(dist, 0)))
    return xy, dist, tau0


def refseq_best(dist, tau0, seed):
    _, gb, _, _ = orc.run(dist, M, variant=orc.AS, mode=orc.MODE_REFSEQ, iterations=ITERS, alpha=1.0,
                          beta=2.0, rho=0.5, tau0=tau0, seed=seed)
    return gb


def main():
    xy, dist, tau0 = c1_instance()
    best = [refseq_best(dist, tau0, s) for s in SEEDS]
    out = {"workload": "AS n=100 uniform [0,1000]^2 (instance seed 1), m=100, alpha 1, beta 2, rho 0.5, "
                       "tau0 = m/C_nn, 500 iterations, reference full-row semantics (oracle refseq mode)",
           "n": N, "m": M, "iterations": ITERS, "tau0": tau0, "seeds": SEEDS, "best": best}
    with open(os.path.join(HERE, "c1_refseq.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"mean {np.mean(best):.1f} min {min(best)} max {max(best)}")


if __name__ == "__main__":
    main()
