"""Oracle replay of one GPU iteration from the GPU's own state snapshot.

Test infrastructure: the oracle is the checker.  Given the state of a
Colony before an iteration (pheromone, aging arrays, best-so-far, iteration
counter), `snapshot()` captures it and `replay()` recomputes the iteration on
the CPU.  Tours, lengths, choice_info and row means must match bit for bit;
the pheromone after the iteration is compared with an FP64 evaporation +
deposit within a relative tolerance (fp32 atomics make its rounding order
non-deterministic).
"""
from __future__ import annotations

import numpy as np

import oracle as orc
from acogpu_amd import capi

MAXD = capi.MAX_DISTANCE


def snapshot(col: capi.Colony) -> dict:
    ag = col.aging()
    ib, gb, _ = col.stats()
    return {
        "tau": col.pheromone(),
        "age": ag["age"].copy(),
        "valid": ag["valid"].copy(),
        "perf": ag["perf"].copy(),
        "prev_len": ag["prev_len"].copy(),
        "num_valid": ag["num_valid"],
        "converged": ag["converged"],
        "gperf": ag["gperf"],
        "iter": ag["iter"],
        "glob_best": gb,
        "best_tour": col.best_tour(),
        "tours": col.tours(),
        "lengths": col.lengths(),
    }


def replay(p: capi.Params, dist: np.ndarray, snap: dict, heur_nn: np.ndarray | None = None) -> dict:
    """heur_nn: the candidate-edge heuristic installed with set_heuristic_nn
    (None = the default eta^beta)."""
    n, m = p.n, p.m
    aging = bool(p.aging)
    cand = p.mode == capi.MODE_CANDIDATE
    eta = orc.eta_table(p.beta, orc.max_offdiag(dist))
    nn = orc.nn_lists(dist, p.k if cand else n)
    age = snap["age"].copy()
    valid = snap["valid"].copy()
    if aging:
        orc.aging_start(age)
    choice, choice_nn, tau_evap, mean = orc.sweep(snap["tau"], dist, eta, nn if cand else None,
                                                   p.alpha, p.rho, want_mean=aging, heur_nn=heur_nn)
    tours = snap["tours"].copy()
    lengths = snap["lengths"].copy()
    tours, lengths, fb = orc.construct(
        orc.MODE_CANDIDATE if cand else orc.MODE_FULLROW, dist, nn, choice, choice_nn, snap["tau"], mean,
        eta, p.alpha, m, p.seed, snap["iter"], p.stream_id, age=age if aging else None,
        valid=valid if aging else None, converged=snap["converged"], min_age=p.min_age,
        tours=tours, lengths=lengths, heur_nn=heur_nn)
    perf = snap["perf"].copy()
    prev_len = snap["prev_len"].copy()
    st = np.array([MAXD - 1, snap["glob_best"], -1, snap["gperf"]], dtype=np.int64)
    best = snap["best_tour"].copy()
    w = min(p.w, n)
    dep_ant, dep_w = orc.stats(p.variant, tours, lengths, w, st, best,
                               valid=valid if aging else None, perf=perf if aging else None,
                               prev_len=prev_len if aging else None)
    tau64 = snap["tau"].astype(np.float64)
    orc.evaporate_f64(tau64, p.rho)
    orc.deposit_f64(tau64, tours, lengths, best, int(st[1]), dep_ant, dep_w)
    sc = np.array([snap["num_valid"], snap["converged"]], dtype=np.int32)
    if aging:
        orc.aging_control(age, valid, sc, int(st[3]), p.max_age, p.global_perf, p.min_pop)
    return {
        "choice": choice, "choice_nn": choice_nn, "mean": mean, "tau_evap": tau_evap,
        "tours": tours, "lengths": lengths, "iter_best": int(st[0]), "glob_best": int(st[1]),
        "gperf": int(st[3]), "best_tour": best, "tau_next": tau64, "age": age, "valid": valid,
        "perf": perf, "prev_len": prev_len, "num_valid": int(sc[0]), "converged": int(sc[1]),
        "fallback": fb, "nn": nn, "eta": eta,
    }


def active_ants(p: capi.Params, snap: dict) -> np.ndarray:
    """Ants that construct a tour in the replayed iteration."""
    if p.aging:
        return np.nonzero(snap["valid"])[0]
    return np.arange(p.m)


def check_iteration(col: capi.Colony, p: capi.Params, dist: np.ndarray, snap: dict, ref: dict,
                    tau_rtol: float = 1e-5, check_mean: bool = True) -> None:
    """Compare the GPU state after one iteration with the oracle replay."""
    act = active_ants(p, snap)
    np.testing.assert_array_equal(col.choice(), ref["choice"], err_msg="choice_info differs")
    if p.mode == capi.MODE_CANDIDATE:
        np.testing.assert_array_equal(col.choice_nn(), ref["choice_nn"], err_msg="choice_nn differs")
    if p.aging and check_mean:
        np.testing.assert_array_equal(col.mean(), ref["mean"], err_msg="row means differ")
    gt = col.tours()
    gl = col.lengths()
    bad = [a for a in act if not np.array_equal(gt[a], ref["tours"][a])]
    assert not bad, f"{len(bad)} of {len(act)} tours differ, first ant {bad[0]}: " \
                    f"gpu {gt[bad[0]][:12]} oracle {ref['tours'][bad[0]][:12]}"
    np.testing.assert_array_equal(gl[act], ref["lengths"][act], err_msg="tour lengths differ")
    for a in act[:8]:
        assert sorted(gt[a].tolist()) == list(range(p.n)), "tour is not a permutation"
    ib, gb, nv = col.stats()
    assert ib == ref["iter_best"]
    assert gb == ref["glob_best"]
    np.testing.assert_array_equal(col.best_tour(), ref["best_tour"])
    tau = col.pheromone().astype(np.float64)
    rel = np.abs(tau - ref["tau_next"]) / np.abs(ref["tau_next"])
    assert rel.max() <= tau_rtol, f"pheromone rel err {rel.max():.3e} > {tau_rtol:g}"
    if p.aging:
        ag = col.aging()
        np.testing.assert_array_equal(ag["age"], ref["age"])
        np.testing.assert_array_equal(ag["valid"], ref["valid"])
        np.testing.assert_array_equal(ag["perf"][act], ref["perf"][act])
        assert ag["num_valid"] == ref["num_valid"]
        assert ag["converged"] == ref["converged"]
        assert ag["gperf"] == ref["gperf"]
        assert nv == ref["num_valid"]
