"""Pins of the oracle's sweep, young-ant weights and full-row / reference-
sequential construction, as check functions that raise AssertionError.

tests/test_oracle.py runs every check against the oracle as built;
tests/test_oracle_mutants.py runs them against copies of the oracle with one
plausible mistake planted (on the CPU) and requires each copy to fail one of
them.  Expected values come from closed forms evaluated in numpy (fp32 where
the definition rounds to fp32, fp64 otherwise) and from brute force, never
from the oracle itself.

Reference lines pinned:
  choice = tau^alpha * eta          inc/Colony.cuh:91-99, src/Colony.cu:252-273
  evaporation max((1-rho) tau,1e-18) src/AntSystem.cu:9-11, 26-32
  young-ant weight                  inc/Colony.cuh:23-27
                                    ((tau - mean) * age/minAge + mean) * eta
  farthest-unvisited fallback       inc/Colony.cuh:33-43 (select_local_best)
  zero sum / overrun                src/Colony.cu:776-778, 832-839
"""
from __future__ import annotations

import numpy as np
from scipy import stats as sstats

import oracle as orc
from acogpu_amd import instances

F32 = np.float32


def eta_closed_form(d, beta):
    """1/d^beta for integer beta (d^beta exact in fp64, one fp64 division,
    rounded to fp32); 0 on the diagonal; d = 0 counts as d = 1/2."""
    d = np.asarray(d, dtype=np.int64)
    out = np.zeros(d.shape, dtype=F32)
    off = d != orc.MAX_DISTANCE
    dd = d[off].astype(np.float64)
    v = np.where(dd == 0, 2.0 ** beta, 1.0 / np.where(dd == 0, 1.0, dd) ** beta)
    out[off] = v.astype(F32)
    return out


def _sweep_instance(n=48, seed=21):
    xy = instances.uniform(n, 1000, seed=seed).reshape(-1, 2).copy()
    xy[7] = xy[3]  # duplicate city: d(3,7) = 0
    d = orc.dist_from_int_coords(xy)
    rng = np.random.default_rng(seed)
    tau = (10.0 ** rng.uniform(-6, -1, size=(n, n))).astype(F32)  # not symmetric on purpose
    # values at and around the evaporation floor, and zero
    tau[0, 1:8] = np.array([0.0, 1e-18, 1.5e-18, 2.0e-18, 2.1e-18, 1e-30, 3e-18], dtype=F32)
    return d, tau


def check_sweep_closed_forms():
    """choice, choice_nn and the evaporated tau of orc_sweep, elementwise."""
    beta, rho, k = 3.0, 0.5, 10
    d, tau = _sweep_instance()
    n = d.shape[0]
    eta_tab = orc.eta_table(beta, orc.max_offdiag(d))
    eta = eta_closed_form(d, beta)
    assert d[3, 7] == 0 and eta[3, 7] == F32(8.0) and eta[5, 5] == 0.0
    nn = orc.nn_lists(d, k)

    # alpha = 1: one fp32 product of tau (not the evaporated tau) and eta
    choice, cnn, tev, _ = orc.sweep(tau, d, eta_tab, nn, 1.0, rho)
    want = (tau * eta).astype(F32)
    np.testing.assert_array_equal(choice.view(np.uint32), want.view(np.uint32))
    assert np.all(np.diag(choice) == 0.0)
    np.testing.assert_array_equal(cnn, choice[np.arange(n)[:, None], nn])

    # evaporation: fp32 product with the fp32 factor 1 - rho, floored at 1e-18
    p = (F32(1.0 - F32(rho)) * tau).astype(F32)
    want_ev = np.where(p.astype(np.float64) > 1e-18, p, F32(1e-18)).astype(F32)
    np.testing.assert_array_equal(tev.view(np.uint32), want_ev.view(np.uint32))
    exact = np.maximum((1.0 - rho) * tau.astype(np.float64), 1e-18)
    assert np.all(np.abs(tev - exact) <= 0.5 * np.spacing(tev.astype(F32)).astype(np.float64) + 1e-45)
    assert tev[0, 1] == F32(1e-18) and tev[0, 2] == F32(1e-18) and tev[0, 6] == F32(1e-18)

    # alpha = 1.5: tau^1.5 within one fp32 rounding of the fp64 power, times eta
    choice15, cnn15, _, _ = orc.sweep(tau, d, eta_tab, nn, 1.5, rho)
    ref = tau.astype(np.float64) ** 1.5 * eta.astype(np.float64)
    nz = ref > 2.0 ** -100  # fp32 normal range (tau = 1e-30 underflows)
    rel = np.abs(choice15[nz] - ref[nz]) / ref[nz]
    assert rel.max() < 2.5 * 2.0 ** -24, rel.max()
    assert np.all(np.abs(choice15[~nz] - ref[~nz]) <= 2.0 ** -126)
    np.testing.assert_array_equal(cnn15, choice15[np.arange(n)[:, None], nn])

    # candidate-edge heuristic override: tau^alpha * h on the candidate edges
    h = np.random.default_rng(5).uniform(0.1, 2.0, size=(n, k)).astype(F32)
    _, cnnh, _, _ = orc.sweep(tau, d, eta_tab, nn, 1.0, rho, heur_nn=h)
    np.testing.assert_array_equal(cnnh, (tau[np.arange(n)[:, None], nn] * h).astype(F32))


# --------------------------------------------------------------------------
def _first_moves(tours):
    return tours[:, 0], tours[:, 1]


def _chi2_first_move(tours, weights, alpha_p=1e-4):
    """Pearson chi-square of the first move per start city against
    weights[c] (length n, zero where not selectable), summed over starts."""
    starts, firsts = _first_moves(tours)
    n = weights.shape[0]
    stat, dof = 0.0, 0
    for c in range(n):
        sel = firsts[starts == c]
        if len(sel) == 0:
            continue
        w = weights[c].astype(np.float64)
        assert np.all(w[sel] > 0), "a move of weight zero was taken"
        pos = np.flatnonzero(w > 0)
        exp = len(sel) * w[pos] / w[pos].sum()
        obs = np.bincount(sel, minlength=n)[pos]
        stat += float(((obs - exp) ** 2 / exp).sum())
        dof += len(pos) - 1
    assert dof > 0
    p = float(sstats.chi2.sf(stat, dof))
    assert p > alpha_p, f"chi2 {stat:.1f} on {dof} dof, p = {p:.2e}"
    return p


def _small_instance(n=12, seed=31, beta=2.0):
    xy = instances.uniform(n, 1000, seed=seed)
    d = orc.dist_from_int_coords(xy)
    eta_tab = orc.eta_table(beta, orc.max_offdiag(d))
    rng = np.random.default_rng(seed)
    tau = (10.0 ** rng.uniform(-3, -1, size=(n, n))).astype(F32)
    return d, eta_tab, eta_closed_form(d, beta), tau


def check_young_first_move(mode, alpha=1.0, with_heur=False):
    """Young ants (converged colony, age < min_age) choose their first move
    with probability proportional to ((tau - mean) * age/min_age + mean)^alpha
    times eta (or the candidate-edge heuristic h), over every unvisited
    candidate; ants of age >= min_age use tau^alpha eta."""
    n, m, age0, min_age = 12, 30000, 10, 30
    d, eta_tab, eta, tau = _small_instance(n=n)
    full = orc.nn_lists(d, n)
    nn = orc.nn_lists(d, n - 1)  # every other city is a candidate
    mean = tau.astype(np.float64).mean(axis=1).astype(F32)
    age = np.full(m, age0, dtype=np.int32)
    valid = np.ones(m, dtype=np.uint8)
    ratio = age0 / min_age
    pulled = (tau.astype(np.float64) - mean[:, None].astype(np.float64)) * ratio + mean[:, None]
    assert np.all(pulled > 0)
    hfull = None
    heur = None
    if with_heur:
        heur = np.random.default_rng(8).uniform(0.05, 3.0, size=(n, n - 1)).astype(F32)
        hfull = np.zeros((n, n))
        hfull[np.arange(n)[:, None], nn] = heur
        w = pulled ** alpha * hfull
    else:
        w = pulled ** alpha * eta.astype(np.float64)
    np.fill_diagonal(w, 0.0)
    choice, cnn, _, _ = orc.sweep(tau, d, eta_tab, nn, alpha, 0.5, heur_nn=heur)
    nnl = nn if mode == orc.MODE_CANDIDATE else full
    tours, _, _ = orc.construct(mode, d, nnl, choice, cnn, tau, mean, eta_tab, alpha, m, 17, 4,
                                age=age, valid=valid, converged=1, min_age=min_age, heur_nn=heur)
    _chi2_first_move(tours, w)
    # the same ants after min_age (age = min_age): plain weights
    age_old = np.full(m, min_age, dtype=np.int32)
    tours_o, _, _ = orc.construct(mode, d, nnl, choice, cnn, tau, mean, eta_tab, alpha, m, 17, 5,
                                  age=age_old, valid=valid, converged=1, min_age=min_age, heur_nn=heur)
    plain = tau.astype(np.float64) ** alpha * (hfull if with_heur else eta.astype(np.float64))
    np.fill_diagonal(plain, 0.0)
    _chi2_first_move(tours_o, plain)


def check_first_move_multichunk(mode):
    """Full-row and reference-sequential roulette over every unvisited city:
    n = 70 (two 64-lane chunks in the full-row scan), four cities per row
    with pheromone (ranks spread over both chunks), zero elsewhere; the
    first move follows their weights tau * eta."""
    n, m = 70, 30000
    xy = instances.uniform(n, 1000, seed=41)
    d = orc.dist_from_int_coords(xy)
    eta_tab = orc.eta_table(1.0, orc.max_offdiag(d))
    eta = eta_closed_form(d, 1.0)
    full = orc.nn_lists(d, n)
    rng = np.random.default_rng(41)
    tau = np.zeros((n, n), dtype=F32)
    for c in range(n):
        ranks = rng.choice(np.array([2, 30, 62, 63, 64, 65, 67, 68]), size=4, replace=False)
        tau[c, full[c, ranks]] = rng.uniform(0.2, 1.0, size=4).astype(F32) / eta[c, full[c, ranks]]
    choice, _, _, _ = orc.sweep(tau, d, eta_tab, None, 1.0, 0.5)
    tours, _, _ = orc.construct(mode, d, full, choice, None, tau, None, eta_tab, 1.0, m, 23, 1)
    w = tau.astype(np.float64) * eta.astype(np.float64)
    _chi2_first_move(tours, w)


def _farthest_walk(d, start, steps):
    """Brute force: from `start`, repeatedly the unvisited city last in the
    stable (distance, index) order of the current row."""
    n = d.shape[0]
    vis = {start}
    tour = [start]
    c = start
    for _ in range(steps):
        cand = [j for j in range(n) if j not in vis]
        nxt = max(cand, key=lambda j: (int(d[c, j]), j))
        tour.append(nxt)
        vis.add(nxt)
        c = nxt
    return tour


def check_zero_total_takes_farthest(mode):
    """A zero weight total (all pheromone 0) makes every step take the
    farthest unvisited city in the row's NN order."""
    n, m = 30, 12
    xy = instances.uniform(n, 60, seed=51)  # small grid: many equal distances
    d = orc.dist_from_int_coords(xy)
    eta_tab = orc.eta_table(2.0, orc.max_offdiag(d))
    full = orc.nn_lists(d, n)
    tau = np.zeros((n, n), dtype=F32)
    choice, _, _, _ = orc.sweep(tau, d, eta_tab, None, 1.0, 0.5)
    tours, lengths, fb = orc.construct(mode, d, full, choice, None, tau, None, eta_tab, 1.0, m, 3, 2)
    for a in range(m):
        assert tours[a].tolist() == _farthest_walk(d, int(tours[a, 0]), n - 1)
        assert lengths[a] == instances.tour_length(d, tours[a])
    assert fb[0] == m * (n - 1)


def check_overrun_takes_farthest(mode):
    """Weights whose sum overflows to +inf: the walk over the partial sums
    never passes u * sum, the overrun branch takes the farthest unvisited
    city (src/Colony.cu:832-839).  Checked on the steps with four or more
    unvisited cities, whose weights (1e38 to 3e38 each) always sum to +inf."""
    n, m = 20, 16
    rng = np.random.default_rng(61)
    d = rng.integers(1, 4, size=(n, n)).astype(np.int32)
    d = np.minimum(d, d.T)
    np.fill_diagonal(d, orc.MAX_DISTANCE)
    eta_tab = orc.eta_table(1.0, 3)
    full = orc.nn_lists(d, n)
    tau = np.full((n, n), 3.0e38, dtype=F32)
    choice, _, _, _ = orc.sweep(tau, d, eta_tab, None, 1.0, 0.5)
    tours, _, _ = orc.construct(mode, d, full, choice, None, tau, None, eta_tab, 1.0, m, 5, 3)
    for a in range(m):
        t = tours[a].tolist()
        assert sorted(t) == list(range(n))
        # every step with four or more unvisited candidates overflows
        want = _farthest_walk(d, t[0], n - 4)
        assert t[: n - 3] == want, (a, t, want)


# every check, by name (the mutation test requires each planted mistake to
# fail at least one of them)
CHECKS = {
    "sweep_closed_forms": check_sweep_closed_forms,
    "young_candidate": lambda: check_young_first_move(orc.MODE_CANDIDATE),
    "young_candidate_alpha": lambda: check_young_first_move(orc.MODE_CANDIDATE, alpha=1.5),
    "young_candidate_heur": lambda: check_young_first_move(orc.MODE_CANDIDATE, with_heur=True),
    "young_fullrow": lambda: check_young_first_move(orc.MODE_FULLROW),
    "young_refseq": lambda: check_young_first_move(orc.MODE_REFSEQ),
    "first_move_fullrow": lambda: check_first_move_multichunk(orc.MODE_FULLROW),
    "first_move_refseq": lambda: check_first_move_multichunk(orc.MODE_REFSEQ),
    "zero_total_fullrow": lambda: check_zero_total_takes_farthest(orc.MODE_FULLROW),
    "zero_total_refseq": lambda: check_zero_total_takes_farthest(orc.MODE_REFSEQ),
    "overrun_fullrow": lambda: check_overrun_takes_farthest(orc.MODE_FULLROW),
    "overrun_refseq": lambda: check_overrun_takes_farthest(orc.MODE_REFSEQ),
}
