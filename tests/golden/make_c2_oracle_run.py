"""Writes c2_oracle_run.json: the oracle's whole fp32 run (reference CPU
operation order) of the n = 1000 rank-based Ant System workload with aging,
stopped after 20, 820 and 2000 iterations: best-so-far length, last
iteration's best length and best-so-far tour at each stop.

Workload: n = 1000 cities around 10 Gaussian clusters (centres uniform in
[0,10000]^2, sigma 400, instance seed 1), EUC_2D, m = 1000 ants, candidate
lists k = 20, w = 6 ranked ants, alpha 1, beta 3, rho 0.1, tau0 =
0.5 w (w - 1) / (rho C_nn), aging (min_age 30, max_age 50, global_perf 10,
min_pop 40), Philox seed 20241203.  Each stop is its own run from the start
(about 8 and 20 minutes on one core); the three run in parallel processes.

Run from the repository root: python tests/golden/make_c2_oracle_run.py
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as orc  # noqa: E402
from acogpu_amd import instances  # noqa: E402

N, M, K, W = 1000, 1000, 20, 6
BETA, RHO, SEED = 3.0, 0.1, 20241203
AGING = dict(min_age=30, max_age=50, global_perf=10, min_pop=40)
STOPS = [20, 820, 2000]


def c2_instance():
    xy = instances.clustered(N, clusters=10, side=10000, sigma=400.0, seed=1)
    dist = orc.dist_from_int_coords(xy)
    cnn = orc.greedy_length(dist, 0)
    tau0 = float(np.float32(0.5 * W * (W - 1) / (np.float64(np.float32(RHO)) * cnn)))
    return xy, dist, tau0


def oracle_stop(iterations):
    _, dist, tau0 = c2_instance()
    _, gb, ib, best = orc.run(dist, M, variant=orc.RBAS, mode=orc.MODE_CANDIDATE, k=K, w=W, aging=True,
                              iterations=iterations, alpha=1.0, beta=BETA, rho=RHO, tau0=tau0, seed=SEED,
                              **AGING)
    return {"iterations": iterations, "glob_best": gb, "iter_best": ib, "best_tour": best.tolist()}


def main(stops=STOPS):
    _, _, tau0 = c2_instance()
    with ProcessPoolExecutor(len(stops)) as ex:
        res = list(ex.map(oracle_stop, stops))
    out = {"workload": "RBAS + aging, n=1000 clustered (10 clusters, sigma 400, instance seed 1), m=1000, k=20, "
                       "w=6, alpha 1, beta 3, rho 0.1, tau0 = 0.5 w (w-1) / (rho C_nn), min_age 30, max_age 50, "
                       "global_perf 10, min_pop 40, seed 20241203; oracle fp32 run (reference CPU order)",
           "n": N, "m": M, "k": K, "w": W, "beta": BETA, "rho": RHO, "seed": SEED, "tau0": tau0, **AGING,
           "stops": res}
    with open(os.path.join(HERE, "c2_oracle_run.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    for r in res:
        print(r["iterations"], r["glob_best"], r["iter_best"])


if __name__ == "__main__":
    main()
