"""CPU tests of the oracle: each function against a check that a plausible
mistake would fail (published known-answer vectors, closed forms, brute force,
statistics), not against a re-typing of its own formula."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc
import oracle_pins as pins
from acogpu_amd import instances

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123 published known-answer vectors for philox4x32-10."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(orc.philox(ctr, key)) == want


def test_rng_word_counter_layout():
    seed = (0x12345678 << 32) | 0x9ABCDEF0
    for it, ant, step, stream in [(0, 0, 0, 0), (5, 17, 6, 3), (1999, 999, 998, 0), (7, 3, 300, 1)]:
        blk = orc.philox([64 * (step // 256) + step % 64, ant, it, stream], [0x9ABCDEF0, 0x12345678])
        assert orc.rng_word(seed, it, ant, step, stream) == int(blk[(step // 64) % 4])
    # every (block, word) pair is used by exactly one step
    pairs = {(64 * (s // 256) + s % 64, (s // 64) % 4) for s in range(5000)}
    assert len(pairs) == 5000


def test_uniform_and_start_city_distribution():
    words = np.array([orc.rng_word(7, 3, a, s) for a in range(40) for s in range(250)], dtype=np.uint64)
    u = np.array([orc.uniform(int(w)) for w in words])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(u * 2 ** 24 == np.floor(u * 2 ** 24))  # 24-bit grid
    hist, _ = np.histogram(u, bins=20, range=(0, 1))
    exp = len(u) / 20
    assert ((hist - exp) ** 2 / exp).sum() < 45.3  # chi2(19), p = 0.001
    for n in (1, 2, 7, 1000, 20000):
        c = [orc.start_city(int(w), n) for w in words[:500]]
        assert min(c) >= 0 and max(c) < n
        assert c == [int((int(w) * n) >> 32) for w in words[:500]]


# ----------------------------------------------------------------- EUC_2D
def _nint_sqrt(D):
    """nearest integer of sqrt(D) by exact integer arithmetic."""
    r = math.isqrt(D)
    return r + 1 if D - r * r > r else r  # sqrt(D) >= r + 1/2  <=>  D >= r^2 + r + 1/4


def test_euc2d_exact_matches_reference_formula_and_integer_closed_form():
    rng = np.random.default_rng(0)
    d = rng.integers(-100000, 100001, size=(20000, 2))
    for dx, dy in d[:4000]:
        e = orc.euc2d_exact(int(dx), int(dy))
        assert e == _nint_sqrt(int(dx) ** 2 + int(dy) ** 2)
        assert e == orc.euc2d_ref(float(dx), 0.0, float(dy), 0.0)
    # the rounding boundary: D = r^2 + r rounds down to r, D = r^2 + r + 1 up to r + 1
    hits = 0
    for dx in range(0, 400):
        for dy in range(0, 400):
            D = dx * dx + dy * dy
            r = math.isqrt(D)
            if D == r * r + r:
                assert orc.euc2d_exact(dx, dy) == r
                hits += 1
            elif D == r * r + r + 1:
                assert orc.euc2d_exact(dx, dy) == r + 1
                hits += 1
    assert hits > 100
    assert orc.euc2d_exact(3, 4) == 5 and orc.euc2d_exact(1, 1) == 1 and orc.euc2d_exact(0, 0) == 0


def test_distance_matrix_from_coords():
    xy = instances.uniform(60, 1000, seed=2)
    d = orc.dist_from_int_coords(xy)
    assert (np.diag(d) == orc.MAX_DISTANCE).all()
    assert (d == d.T).all()
    brute = np.array([[_nint_sqrt(int((xy[i, 0] - xy[j, 0]) ** 2 + (xy[i, 1] - xy[j, 1]) ** 2))
                       for j in range(60)] for i in range(60)])
    np.fill_diagonal(brute, orc.MAX_DISTANCE)
    np.testing.assert_array_equal(d, brute)


def test_tsplib_reader_and_duplicate_city():
    name, x, y = instances.read_tsplib(os.path.join(GOLD, "tsplib", "a280.tsp"))
    assert name == "a280" and len(x) == 280
    d = orc.dist_from_float_coords(x, y)
    np.testing.assert_array_equal(d, instances.euc2d_matrix_from_float(x, y))
    # nodes 171 and 172 (1-based) share coordinates (release/input/a280.tsp:177-178)
    assert d[170, 171] == 0
    name, x, y = instances.read_tsplib(os.path.join(GOLD, "tsplib", "tsp225.tsp"))
    d = orc.dist_from_float_coords(x, y)
    np.testing.assert_array_equal(d, instances.euc2d_matrix_from_float(x, y))
    with pytest.raises(ValueError):
        bad = os.path.join(os.path.dirname(__file__), "golden", "_bad_type.tsp")
        with open(bad, "w") as f:
            f.write("NAME: x\nTYPE: ATSP\nDIMENSION: 2\n")
        try:
            instances.read_tsplib(bad)
        finally:
            os.remove(bad)


# ----------------------------------------------------------------- NN lists
@pytest.mark.parametrize("n,k,seed", [(50, 10, 1), (200, 20, 2), (300, 32, 3)])
def test_nn_lists_brute_force(n, k, seed):
    xy = instances.uniform(n, 60, seed=seed)  # small grid: many equal distances
    d = orc.dist_from_int_coords(xy)
    nn = orc.nn_lists(d, k)
    for i in range(n):
        cols = np.array([j for j in range(n) if j != i])
        order = cols[np.lexsort((cols, d[i, cols]))]
        np.testing.assert_array_equal(nn[i], order[:k])
    full = orc.nn_lists(d, n)
    assert (full[:, -1] == np.arange(n)).all()  # self last in the full row
    np.testing.assert_array_equal(full[:, :k], nn)


# ----------------------------------------------------------------- heuristic
def test_pow_alpha_against_fp64_power():
    """tau^alpha (alpha != 1) is defined by a fixed fp64 operation sequence
    shared with the kernels; it must agree with the fp64 power rounded to fp32
    (numpy) to one ulp everywhere and exactly almost everywhere, give exact
    results on powers of two, and x^0 = 1."""
    rng = np.random.default_rng(0)
    x = np.exp(rng.uniform(np.log(1e-30), np.log(1e3), 20000)).astype(np.float32)
    for a in (0.5, 1.5, 2.0, 3.7):
        ours = orc.pow_alpha(x, a)
        ref = (x.astype(np.float64) ** np.float64(np.float32(a))).astype(np.float32)
        ulps = np.abs(ours.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1, a
        assert (ulps == 0).mean() >= 0.9999, a
    np.testing.assert_array_equal(orc.pow_alpha(x[:500], 0.0), np.ones(500, np.float32))
    p2 = np.float32(2.0) ** np.arange(-40, 40).astype(np.float32)
    np.testing.assert_array_equal(orc.pow_alpha(p2, 3.0), (p2.astype(np.float64) ** 3).astype(np.float32))
    assert orc.pow_alpha(np.float32(0.0), 1.5)[0] == 0.0


def test_eta_table_closed_form():
    for beta in (2.0, 3.0, 5.0):
        t = orc.eta_table(beta, 5000)
        for d in (1, 2, 3, 7, 10, 999, 4096, 5000):
            exact = np.float32(float(Fraction(1, d ** int(beta))))
            assert abs(int(t[d].view(np.int32)) - int(exact.view(np.int32))) <= 1
        assert t[0] == np.float32(2.0 ** beta)  # duplicate cities: d taken as 1/2
        assert (np.diff(t[1:]) <= 0).all()


# ----------------------------------------------------------------- wave scan / roulette
def test_wave_scan_prefix_sums():
    rng = np.random.default_rng(4)
    v = rng.integers(0, 1000, size=64).astype(np.float32)  # exact in fp32
    np.testing.assert_array_equal(orc.wave_scan(v), np.cumsum(v.astype(np.int64)).astype(np.float32))
    for _ in range(20):
        v = rng.random(64).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 6, size=64)
        s = orc.wave_scan(v)
        ref = np.cumsum(v.astype(np.float64))
        assert np.all(np.abs(s - ref) <= 64 * 2 ** -23 * ref + 1e-40)


def test_select_candidate_frequencies():
    rng = np.random.default_rng(5)
    w = np.zeros(64, dtype=np.float32)
    w[:20] = rng.random(20).astype(np.float32) + np.float32(0.05)
    w[[3, 11, 17]] = 0.0  # visited
    counts = np.zeros(64)
    for i in range(40000):
        r, _ = orc.select_candidate(w, orc.uniform(orc.rng_word(1, 0, i, 1)))
        counts[r] += 1
    assert counts[[3, 11, 17]].sum() == 0 and counts[20:].sum() == 0
    nz = w > 0
    exp = 40000 * w[nz] / w[nz].sum()
    chi2 = ((counts[nz] - exp) ** 2 / exp).sum()
    assert chi2 < 40.8  # chi2(16), p = 0.0005


def test_select_candidate_boundaries():
    w = np.zeros(64, dtype=np.float32)
    assert orc.select_candidate(w, 0.5)[0] == -1  # all visited -> fallback
    w[5] = 2.0
    w[9] = 2.0
    assert orc.select_candidate(w, 0.0)[0] == 5
    assert orc.select_candidate(w, 0.49)[0] == 5
    assert orc.select_candidate(w, 0.5)[0] == 9  # prefix 2 > 2.0? no -> next lane
    assert orc.select_candidate(w, 1.0 - 2 ** -24)[0] == 9
    r, tot = orc.select_candidate(w, 0.75)
    assert r == 9 and tot == 4.0


# ----------------------------------------------------------------- construction
def _setup(n=120, k=12, seed=3, beta=2.0):
    xy = instances.uniform(n, 1000, seed=seed)
    d = orc.dist_from_int_coords(xy)
    eta = orc.eta_table(beta, orc.max_offdiag(d))
    nn = orc.nn_lists(d, k)
    tau = np.full((n, n), 0.01, dtype=np.float32)
    return xy, d, eta, nn, tau


@pytest.mark.parametrize("mode", [orc.MODE_CANDIDATE, orc.MODE_FULLROW, orc.MODE_REFSEQ])
def test_construct_tours_are_permutations_with_exact_lengths(mode):
    xy, d, eta, nn, tau = _setup()
    n = d.shape[0]
    nnl = nn if mode == orc.MODE_CANDIDATE else orc.nn_lists(d, n)
    choice, cnn, _, mean = orc.sweep(tau, d, eta, nn, 1.0, 0.5, want_mean=True)
    tours, lengths, fb = orc.construct(mode, d, nnl, choice, cnn, tau, mean, eta, 1.0, 16, 9, 0)
    for a in range(16):
        assert sorted(tours[a].tolist()) == list(range(n))
        assert lengths[a] == instances.tour_length(d, tours[a])


def test_construct_first_step_frequencies():
    """The first move from a fixed start city follows w/sum(w) (chi-square)."""
    xy, d, eta, nn, tau = _setup(n=60, k=8, seed=6)
    rng = np.random.default_rng(2)
    tau = rng.uniform(0.001, 0.01, size=(60, 60)).astype(np.float32)
    tau = ((tau + tau.T) / 2).astype(np.float32)
    choice, cnn, _, _ = orc.sweep(tau, d, eta, nn, 1.0, 0.5)
    m = 3000
    tours, _, _ = orc.construct(orc.MODE_CANDIDATE, d, nn, choice, cnn, tau, None, eta, 1.0, m, 11, 0)
    starts = tours[:, 0]
    c = int(np.bincount(starts).argmax())
    sel = tours[starts == c, 1]
    w = cnn[c]
    counts = np.array([(sel == j).sum() for j in nn[c]])
    exp = len(sel) * w / w.sum()
    assert counts.sum() == len(sel)
    assert ((counts - exp) ** 2 / exp).sum() < 26.1  # chi2(7), p = 0.0005


def test_fallback_is_argmax_of_unvisited():
    """With k = 1 every step whose single candidate is visited falls back to
    the unvisited city with the largest weight."""
    xy, d, eta, nn1, tau = _setup(n=40, k=1, seed=8)
    choice, cnn, _, _ = orc.sweep(tau, d, eta, nn1, 1.0, 0.5)
    tours, _, fb = orc.construct(orc.MODE_CANDIDATE, d, nn1, choice, cnn, tau, None, eta, 1.0, 4, 3, 0)
    assert fb[0] > 0
    for t in tours:
        vis = {int(t[0])}
        for s in range(1, 40):
            c, nxt = int(t[s - 1]), int(t[s])
            if int(nn1[c, 0]) not in vis:
                assert nxt == int(nn1[c, 0])
            else:
                cand = [j for j in range(40) if j not in vis]
                best = max(choice[c, j] for j in cand)
                assert nxt == min(j for j in cand if choice[c, j] == best)
            vis.add(nxt)


def test_aging_young_state_only_below_min_age_after_convergence():
    """Which ants are young: before convergence nobody, after it only ants
    below min_age (the weights young ants use are pinned by the first-move
    frequencies in test_oracle_pins_young_*)."""
    xy, d, eta, nn, tau = _setup(n=64, k=10, seed=9)
    rng = np.random.default_rng(1)
    tau = rng.uniform(0.001, 0.02, size=(64, 64)).astype(np.float32)
    choice, cnn, _, mean = orc.sweep(tau, d, eta, nn, 1.0, 0.5, want_mean=True)
    np.testing.assert_allclose(mean, tau.astype(np.float64).mean(axis=1), rtol=1e-5)
    age = np.array([1, 29, 30, 200], dtype=np.int32)
    valid = np.ones(4, dtype=np.uint8)
    t_conv, _, _ = orc.construct(orc.MODE_CANDIDATE, d, nn, choice, cnn, tau, mean, eta, 1.0, 4, 5, 2,
                                 age=age, valid=valid, converged=1, min_age=30)
    t_none, _, _ = orc.construct(orc.MODE_CANDIDATE, d, nn, choice, cnn, tau, mean, eta, 1.0, 4, 5, 2,
                                 age=age, valid=valid, converged=0, min_age=30)
    # ants at or above min_age see plain tau; before convergence nobody is young
    np.testing.assert_array_equal(t_conv[2:], t_none[2:])


# ----------------------------------------------------------------- bookkeeping
def test_stats_best_and_rbas_ranking():
    rng = np.random.default_rng(3)
    m, n, w = 40, 10, 6
    tours = np.array([rng.permutation(n) for _ in range(m)], dtype=np.int32)
    lengths = rng.integers(100, 120, size=m).astype(np.int64)
    valid = (rng.random(m) < 0.8).astype(np.uint8)
    st = np.array([orc.MAX_DISTANCE - 1, 10 ** 6, -1, 0], dtype=np.int64)
    best = np.zeros(n, dtype=np.int32)
    ants, ws = orc.stats(orc.RBAS, tours, lengths, w, st, best, valid=valid, perf=np.zeros(m, np.int32),
                         prev_len=np.zeros(m, np.int64))
    vidx = [a for a in range(m) if valid[a]]
    ranked = sorted(vidx, key=lambda a: (lengths[a], a))
    assert st[0] == lengths[ranked[0]] and st[2] == ranked[0] and st[1] == st[0]
    np.testing.assert_array_equal(best, tours[ranked[0]])
    assert list(ants[:-1]) == ranked[:w - 1] and list(ws[:-1]) == [5, 4, 3, 2, 1]
    assert ants[-1] == -1 and ws[-1] == w
    st2 = st.copy()
    ants_as, ws_as = orc.stats(orc.AS, tours, lengths, w, st2, best, valid=valid)
    assert list(ants_as) == vidx and set(ws_as) == {1}
    assert st2[3] == 1  # no improvement -> stagnation counter


def test_deposit_f64_matches_edge_accumulation():
    rng = np.random.default_rng(7)
    n, m = 30, 12
    tours = np.array([rng.permutation(n) for _ in range(m)], dtype=np.int32)
    lengths = rng.integers(500, 900, size=m).astype(np.int64)
    best = rng.permutation(n).astype(np.int32)
    tau0 = rng.uniform(1e-20, 1e-3, size=(n, n))
    tau = tau0.copy()
    orc.evaporate_f64(tau, 0.3)
    f = float(np.float32(1.0) - np.float32(0.3))
    np.testing.assert_array_equal(tau, np.maximum(f * tau0, 1e-18))
    dep_ant = np.array([0, 3, 5, -1], dtype=np.int32)
    dep_w = np.array([3, 2, 1, 4], dtype=np.int32)
    ref = tau.copy()
    orc.deposit_f64(tau, tours, lengths, best, 777, dep_ant, dep_w)
    for a, wt in zip(dep_ant, dep_w):
        t = tours[a] if a >= 0 else best
        L = lengths[a] if a >= 0 else 777
        i, j = t, np.roll(t, -1)
        np.add.at(ref, (i, j), wt / L)
        np.add.at(ref, (j, i), wt / L)
    np.testing.assert_allclose(tau, ref, rtol=1e-14)


def test_aging_retire_recruit_trace():
    """All ants start together, so with a large min_pop the lowest valid index
    among the first m/2 retires once per iteration from age max_age on, and is
    recruited back immediately (age reset): a closed-form cycle."""
    m, max_age = 8, 3
    age = np.zeros(m, dtype=np.int32)
    valid = np.ones(m, dtype=np.uint8)
    st = np.array([m, 0], dtype=np.int32)
    retired = []
    for it in range(12):
        orc.aging_start(age)
        before = valid.copy()
        orc.aging_control(age, valid, st, gperf=0, max_age=max_age, global_perf=100, min_pop=m)
        assert st[0] == m  # every retirement is recruited straight back
        r = [a for a in range(m) if age[a] == 0]
        retired.append(r[0] if r else None)
        assert (valid == before).all()
    # ages reach max_age at it = max_age - 1; a recruited ant is old again
    # max_age iterations later and, having the lowest index, retires first:
    # the retired ant at iteration t is (t - (max_age - 1)) mod max_age.
    assert retired[:2] == [None, None]
    assert retired[2:] == [(t - (max_age - 1)) % max_age for t in range(2, 12)]
    assert st[1] == 1
    # gperf threshold recruits even without retirement; min_pop triggers too
    age = np.full(m, 10, dtype=np.int32)
    valid = np.array([1, 1, 0, 1, 1, 1, 0, 1], dtype=np.uint8)
    st = np.array([6, 1], dtype=np.int32)
    orc.aging_control(age, valid, st, gperf=5, max_age=100, global_perf=5, min_pop=0)
    assert valid[2] == 1 and age[2] == 0 and st[0] == 7 and valid[6] == 0


def test_greedy_length_brute_force():
    xy = instances.uniform(30, 100, seed=4)
    d = orc.dist_from_int_coords(xy)
    vis = [0]
    while len(vis) < 30:
        c = vis[-1]
        vis.append(min((j for j in range(30) if j not in vis), key=lambda j: (d[c, j], j)))
    assert orc.greedy_length(d, 0) == instances.tour_length(d, np.array(vis))


# ----------------------------------------------------------------- whole runs
def test_refseq_run_respects_tsplib_optimum():
    opt = json.load(open(os.path.join(GOLD, "tsplib", "optima.json")))["optimum"]
    name, x, y = instances.read_tsplib(os.path.join(GOLD, "tsplib", "berlin52.tsp"))
    d = orc.dist_from_float_coords(x, y)
    secs, gb, ib, best = orc.run(d, 52, variant=orc.AS, mode=orc.MODE_REFSEQ, iterations=150, alpha=1.0,
                                 beta=5.0, rho=0.5, tau0=1e-6, seed=3)
    assert gb == instances.tour_length(d, best)
    assert opt["berlin52"] <= gb <= 1.10 * opt["berlin52"]


@pytest.mark.parametrize("inst", ["eil76", "kroA100"])
def test_candidate_rbas_run_respects_tsplib_optimum(inst):
    opt = json.load(open(os.path.join(GOLD, "tsplib", "optima.json")))["optimum"][inst]
    name, x, y = instances.read_tsplib(os.path.join(GOLD, "tsplib", f"{inst}.tsp"))
    d = orc.dist_from_float_coords(x, y)
    n = d.shape[0]
    cnn = orc.greedy_length(d, 0)
    tau0 = float(np.float32(0.5 * 6 * 5 / (np.float64(np.float32(0.1)) * cnn)))
    secs, gb, ib, best = orc.run(d, n, variant=orc.RBAS, mode=orc.MODE_CANDIDATE, k=20, w=6,
                                 iterations=300, beta=3.0, rho=0.1, tau0=tau0, seed=1)
    assert gb == instances.tour_length(d, best)
    assert opt <= gb <= 1.06 * opt


def test_c1_refseq_fixture_reproduces():
    """The committed 30-seed quality fixture is what the oracle computes
    (first seed re-run)."""
    import json
    import os
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_c1_refseq as mk

    ref = json.load(open(os.path.join(here, "golden", "c1_refseq.json")))
    _, dist, tau0 = mk.c1_instance()
    assert np.float32(tau0) == np.float32(ref["tau0"])
    assert mk.refseq_best(dist, tau0, ref["seeds"][0]) == ref["best"][0]


def test_c2_oracle_run_fixture_reproduces():
    """The committed whole-run fixture of the n = 1000 rank-based + aging
    workload is what the oracle computes (its 20-iteration stop re-run,
    about 15 s), and every stop's best tour is a permutation whose EUC_2D
    length is the recorded best."""
    import json
    import os
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_c2_oracle_run as mk

    ref = json.load(open(os.path.join(here, "golden", "c2_oracle_run.json")))
    _, dist, tau0 = mk.c2_instance()
    assert np.float32(tau0) == np.float32(ref["tau0"])
    for st in ref["stops"]:
        t = np.array(st["best_tour"], dtype=np.int32)
        assert sorted(t.tolist()) == list(range(mk.N))
        assert instances.tour_length(dist, t) == st["glob_best"] <= st["iter_best"]
    first = ref["stops"][0]
    assert mk.oracle_stop(first["iterations"]) == first


# ----------------------------------------------------------------- closed-form / brute-force pins
@pytest.mark.parametrize("name", sorted(pins.CHECKS))
def test_oracle_pins(name):
    """Sweep elementwise closed forms, young-ant weights (chi-square of the
    first move against numpy weights), full-row and reference-sequential
    roulette (chi-square over two scan chunks), and the farthest-unvisited
    fallback on zero and overflowing sums (brute force); see
    tests/oracle_pins.py."""
    pins.CHECKS[name]()
