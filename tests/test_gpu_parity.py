"""GPU parity: every kernel of the hot path against the CPU oracle.

Each test prints the case it runs (flushed) before touching the GPU.  All
calls go through the C ABI (include/acogpu.h) via ctypes.
"""
import numpy as np
import pytest

import oracle as orc
from acogpu_amd import capi, instances
from replay import MAXD, check_iteration, replay, snapshot

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make(params, xy=None, dist=None):
    col = capi.Colony(params, dist=dist, xy=xy)
    col.initialize()
    return col


def random_tau(rng, n, lo=1e-6, hi=1e-2):
    t = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, n))).astype(np.float32)
    t = np.triu(t, 1)
    return (t + t.T).astype(np.float32)


def test_inputs_dist_nn_eta():
    say("case: inputs n=300 uniform seed=3 k=20 beta=3")
    xy = instances.uniform(300, 1000, seed=3)
    p = capi.default_params(n=300, m=8, k=20, beta=3.0)
    with make(p, xy=xy) as col:
        ref = orc.dist_from_int_coords(xy)
        np.testing.assert_array_equal(col.dist(), ref)
        np.testing.assert_array_equal(col.nn(), orc.nn_lists(ref, 20))
        np.testing.assert_array_equal(col.eta(), orc.eta_table(3.0, orc.max_offdiag(ref)))


def test_inputs_host_distance_matrix():
    say("case: host dist n=130 duplicate city, diagonal forced")
    xy = instances.uniform(130, 500, seed=4)
    xy[7] = xy[3]  # duplicate point: d = 0
    d = orc.dist_from_int_coords(xy)
    d2 = d.copy()
    np.fill_diagonal(d2, 0)  # the diagonal is ignored by the ABI
    p = capi.default_params(n=130, m=4, k=10, beta=2.0)
    with make(p, dist=d2) as col:
        np.testing.assert_array_equal(col.dist(), d)
        np.testing.assert_array_equal(col.nn(), orc.nn_lists(d, 10))


PATHS = [1, 2]  # construction in global memory / staged in LDS


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n,m,k,alpha,beta,rho,seed", [
    (64, 32, 16, 1.0, 2.0, 0.5, 1),
    (200, 64, 20, 1.0, 3.0, 0.1, 2),
    (333, 100, 32, 1.0, 2.0, 0.2, 3),
    (150, 40, 64, 1.0, 5.0, 0.1, 4),
    (257, 70, 20, 1.5, 2.0, 0.3, 5),
])
def test_as_iteration_replay(n, m, k, alpha, beta, rho, seed, path):
    say(f"case: AS candidate n={n} m={m} k={k} alpha={alpha} beta={beta} rho={rho} seed={seed} path={path}")
    rng = np.random.default_rng(seed)
    xy = instances.uniform(n, 1000, seed=seed)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, alpha=alpha, beta=beta, rho=rho, seed=1000 + seed,
                            construct_path=path)
    with make(p, xy=xy) as col:
        assert col.config()["path"] == (0 if path == 1 else 2)
        col.set_pheromone(random_tau(rng, n))
        for it in range(3):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [127, 128, 129, 192])
def test_step_chunk_boundaries(n, path):
    """The construction runs its steps in chunks of 64 (a chunk's random words
    and tour entries one per lane): tours that end just before, at and just
    after a chunk boundary."""
    say(f"case: AS chunk boundary n={n} path={path}")
    rng = np.random.default_rng(n)
    xy = instances.uniform(n, 1000, seed=300 + n)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=40, k=12, beta=2.0, rho=0.3, seed=900 + n, construct_path=path)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        for it in range(2):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("path", PATHS)
def test_as_iteration_replay_fallback_heavy(path):
    say(f"case: AS candidate n=400 m=48 k=4 (fallback heavy) path={path}")
    xy = instances.uniform(400, 2000, seed=9)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=400, m=48, k=4, beta=2.0, rho=0.3, seed=77, construct_path=path)
    with make(p, xy=xy) as col:
        snap = snapshot(col)
        col.iterate()
        ref = replay(p, dist, snap)
        check_iteration(col, p, dist, snap, ref)
        steps, scanned, tours, listed = col.counters()
        assert listed == 0  # no top lists below 4096 cities
        assert tours == p.m
        assert steps == ref["fallback"][0] and scanned == ref["fallback"][1]
        assert steps > 0


@pytest.mark.parametrize("n,m,k,path,seed", [
    (5000, 24, 32, 1, 6),   # global-memory candidate rows, 32-lane scan, LDS bitmap
    (2500, 16, 40, 1, 8),   # global-memory candidate rows, 64-lane scan
    (3000, 16, 8, 2, 7),    # LDS-staged candidate rows with the LDS bitmap (n > 2048), fallback heavy
    (4500, 16, 4, 2, 9),    # LDS-staged rows, fallbacks answered by the row top lists
])
def test_as_iteration_replay_large_n(n, m, k, path, seed):
    say(f"case: AS large n={n} m={m} k={k} path={path} seed={seed}")
    rng = np.random.default_rng(seed)
    xy = instances.uniform(n, 100000, seed=seed)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.2, seed=500 + seed, construct_path=path)
    with make(p, xy=xy) as col:
        assert col.config()["path"] == (0 if path == 1 else 2)
        col.set_pheromone(random_tau(rng, n))
        for it in range(3):
            snap = snapshot(col)
            col.iterate()
            ref = replay(p, dist, snap)
            check_iteration(col, p, dist, snap, ref)
        assert ref["fallback"][0] > 0
        fb, _, _, listed = col.counters()
        if n >= 4096:  # row top lists (alpha = 1) answer most fallbacks, the rest scan rows
            assert 0 < listed < fb
        else:
            assert listed == 0


@pytest.mark.parametrize("beta", [1.0, 2.0, 3.0])
@pytest.mark.parametrize("span", [(1 << 20) - 1, 1 << 20])
def test_sweep_distances_from_extreme_coordinates(beta, span):
    """The sweep computes its rows' distances from the coordinates when every
    |coordinate| < 2^20 (fp32 square root, one exact fp64 correction) and
    reads the matrix otherwise; either way choice_info, tours and pheromone
    equal the oracle's, which uses the exact integer distances.  Coordinates
    reach +-span (corners included) with duplicate cities; n = 1100 covers a
    full 1024-column chunk and a tail; beta 1 and 2 use the sweep's own
    reciprocal, beta 3 (d^3 beyond 2^53) the table."""
    n, m, k = 1100, 8, 16
    say(f"case: sweep distances, coordinates up to {span}, beta {beta}")
    rng = np.random.default_rng(int(beta) * 7 + (span & 1))
    xy = rng.integers(-span, span + 1, size=(n, 2)).astype(np.int32)
    xy[:4] = [[-span, -span], [span, span], [-span, span], [span, -span]]
    xy[10:14] = xy[20]  # duplicates: distance 0
    xy[30] = xy[31] + np.array([1, 0], np.int32)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, beta=beta, rho=0.2, seed=77)
    with make(p, xy=xy) as col:
        np.testing.assert_array_equal(col.dist(), dist)
        col.set_pheromone(random_tau(rng, n))
        for _ in range(2):
            snap = snapshot(col)
            col.iterate()
            ref = replay(p, dist, snap)
            check_iteration(col, p, dist, snap, ref)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n,k,m", [(2, 1, 1), (3, 2, 3), (5, 4, 7), (33, 32, 5), (65, 64, 4)])
def test_tiny_instances_replay(n, k, m, path):
    """Smallest instances: one candidate, n below the wave width, k = n - 1."""
    say(f"case: tiny AS n={n} k={k} m={m} path={path}")
    rng = np.random.default_rng(n)
    xy = instances.uniform(n, 1000, seed=60 + n)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.5, seed=n, construct_path=path)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        for it in range(3):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("n,m", [(2, 1), (3, 4), (64, 3)])
def test_tiny_fullrow_replay(n, m):
    say(f"case: tiny full-row AS n={n} m={m}")
    rng = np.random.default_rng(n)
    xy = instances.uniform(n, 1000, seed=80 + n)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, mode=capi.MODE_FULLROW, beta=2.0, rho=0.5, seed=n)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        for it in range(2):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("path", PATHS)
def test_rbas_aging_replay_c2_shape(path):
    say(f"case: RBAS+aging clustered n=500 m=160 k=20 w=6 beta=3 rho=0.1 (C2 shape, small) path={path}")
    xy = instances.clustered(500, seed=11)
    dist = orc.dist_from_int_coords(xy)
    # max_age 6: retirements from the 6th iteration on; min_pop 159 makes every
    # retirement after the first trigger a recruitment, so young ants appear.
    p = capi.default_params(n=500, m=160, k=20, w=6, beta=3.0, rho=0.1, variant=capi.VARIANT_RBAS,
                            aging=1, min_age=4, max_age=6, global_perf=3, min_pop=159,
                            tau0_policy=capi.TAU0_RBAS, seed=5, construct_path=path)
    with make(p, xy=xy) as col:
        checked_young = False
        expect_tours = 0
        for it in range(14):
            snap = snapshot(col)
            expect_tours += snap["num_valid"]
            col.iterate()
            ref = replay(p, dist, snap)
            check_iteration(col, p, dist, snap, ref)
            if snap["converged"] and np.any((snap["age"] + 1 < p.min_age) & (snap["valid"] > 0)):
                checked_young = True
        assert checked_young, "no young ant after convergence was exercised"
        assert col.counters()[2] == expect_tours  # only valid ants construct


@pytest.mark.parametrize("path", PATHS)
def test_rbas_replay_forced_young_state(path):
    say(f"case: RBAS aging forced state n=256 m=64 k=24 path={path}")
    rng = np.random.default_rng(21)
    xy = instances.uniform(256, 5000, seed=21)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=256, m=64, k=24, w=4, beta=2.0, rho=0.25, variant=capi.VARIANT_RBAS,
                            aging=1, min_age=10, max_age=12, global_perf=5, min_pop=60, seed=99,
                            construct_path=path)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, 256))
        age = rng.integers(0, 14, size=64).astype(np.int32)
        valid = (rng.random(64) < 0.8).astype(np.uint8)
        col.set_aging(age=age, valid=valid, perf=np.zeros(64, np.int32), prev_len=np.zeros(64, np.int64),
                      scalars=np.array([int(valid.sum()), 1, 4, 17], np.int32))
        col.set_best(10 ** 9, np.arange(256, dtype=np.int32))
        snap = snapshot(col)
        col.iterate()
        check_iteration(col, p, dist, snap, replay(p, dist, snap))


def test_c2_full_size_settled_iterations_replay():
    """The n = 1000 rank-based + aging workload at its full size (10 Gaussian
    clusters, m = 1000, k = 20, w = 6, beta 3, rho 0.1, min_age 30, max_age
    50), settled past the retirement age like the benchmark's colony, then
    two iterations replayed through the oracle: young ants with their own
    staged blocks, fallbacks of both kinds, the rank-based deposit."""
    n, m = 1000, 1000
    say(f"case: C2 full size n={n} m={m} k=20 w=6 RBAS + aging, 505 iterations then 2 replayed")
    xy = instances.clustered(n, clusters=10, side=10000, sigma=400.0, seed=1)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(variant=capi.VARIANT_RBAS, mode=capi.MODE_CANDIDATE, n=n, m=m, k=20, w=6,
                            alpha=1.0, beta=3.0, rho=0.1, tau0_policy=capi.TAU0_RBAS, aging=1,
                            min_age=30, max_age=50, local_perf=10, global_perf=10, min_pop=40, max_pop=100,
                            seed=20241203)
    with make(p, xy=xy) as col:
        assert col.config()["path"] == 2  # LDS-staged construction
        col.run(505)
        young_seen = False
        for _ in range(2):
            snap = snapshot(col)
            young_seen = young_seen or bool(
                snap["converged"] and np.any((snap["age"] + 1 < p.min_age) & (snap["valid"] > 0)))
            fb0 = col.counters()[0]
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))
            assert col.counters()[0] > fb0
        assert young_seen, "no young ant in the settled colony"


# the register-bitmap fallback at each of its widths (8, 16 and 32 cities per
# lane for n <= 512, 1024, 2048), for ordinary and young ants, with few
# candidates so that most ants fall back many times
@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [509, 700, 1024, 1500, 2048])
def test_rbas_replay_forced_young_state_fallback_widths(n, path):
    m, k = 64, 6
    say(f"case: RBAS aging forced state n={n} m={m} k={k} (fallback heavy) path={path}")
    rng = np.random.default_rng(n)
    xy = instances.uniform(n, 20000, seed=n)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, w=4, beta=2.0, rho=0.25, variant=capi.VARIANT_RBAS,
                            aging=1, min_age=10, max_age=12, global_perf=5, min_pop=60, seed=n + 3,
                            construct_path=path)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        age = rng.integers(0, 14, size=m).astype(np.int32)
        valid = (rng.random(m) < 0.8).astype(np.uint8)
        col.set_aging(age=age, valid=valid, perf=np.zeros(m, np.int32), prev_len=np.zeros(m, np.int64),
                      scalars=np.array([int(valid.sum()), 1, 4, 17], np.int32))
        col.set_best(10 ** 12, np.arange(n, dtype=np.int32))
        snap = snapshot(col)
        col.iterate()
        check_iteration(col, p, dist, snap, replay(p, dist, snap))
        fb = col.counters()[0]
        assert fb > 0


@pytest.mark.parametrize("path", PATHS)
def test_heuristic_override_replay(path):
    """A candidate-edge heuristic replaces eta^beta on NN edges (old and young
    ants); the fallback keeps eta^beta; None restores the default."""
    say(f"case: heuristic override RBAS aging forced state n=300 m=64 k=16 path={path}")
    rng = np.random.default_rng(31)
    xy = instances.uniform(300, 3000, seed=31)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=300, m=64, k=16, w=4, beta=2.0, rho=0.25, variant=capi.VARIANT_RBAS,
                            aging=1, min_age=10, max_age=12, global_perf=5, min_pop=60, seed=123,
                            construct_path=path)
    with make(p, xy=xy) as col:
        default = col.heuristic_nn()  # eta^beta; pinned by the replay after the restore below
        h = np.exp(rng.uniform(np.log(1e-6), np.log(1.0), size=(300, 16))).astype(np.float32)
        h[:, 0] = 0.0  # zero weights: never chosen by roulette
        col.set_heuristic_nn(h)
        np.testing.assert_array_equal(col.heuristic_nn(), h)
        col.set_pheromone(random_tau(rng, 300))
        age = rng.integers(0, 14, size=64).astype(np.int32)
        valid = (rng.random(64) < 0.8).astype(np.uint8)
        col.set_aging(age=age, valid=valid, perf=np.zeros(64, np.int32), prev_len=np.zeros(64, np.int64),
                      scalars=np.array([int(valid.sum()), 1, 4, 17], np.int32))
        col.set_best(10 ** 9, np.arange(300, dtype=np.int32))
        for _ in range(2):
            snap = snapshot(col)
            col.iterate()
            ref = replay(p, dist, snap, heur_nn=h)
            check_iteration(col, p, dist, snap, ref)
        col.set_heuristic_nn(None)
        np.testing.assert_array_equal(col.heuristic_nn(), default)
        snap = snapshot(col)
        col.iterate()
        check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("m,w", [(5, 8), (64, 6), (64, 17), (64, 30), (4500, 6)])
def test_rbas_rank_selection_replay(m, w):
    """Rank-based deposit lists: two-round selection (w - 1 <= 16) and the
    general w - 1 passes, including fewer ants than w - 1."""
    say(f"case: RBAS ranking n=200 m={m} w={w}")
    rng = np.random.default_rng(50 + w)
    xy = instances.uniform(200, 2000, seed=50 + w)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=200, m=m, k=16, w=w, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS, seed=w)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, 200))
        for _ in range(2):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("aging", [0, 1])
def test_as_deposit_paths_agree_and_replay(aging):
    """Ant System deposit through fp32 atomics (path 1) and through the
    adjacency scatter + per-row LDS gather (path 2): same tours, pheromone
    within fp32 rounding of each other and within 1e-5 of the FP64 oracle;
    with aging, invalid ants deposit nothing."""
    say(f"case: AS deposit paths n=300 m=64 k=16 aging={aging}")
    rng = np.random.default_rng(41)
    xy = instances.uniform(300, 3000, seed=41)
    dist = orc.dist_from_int_coords(xy)
    tau = random_tau(rng, 300)
    age = rng.integers(0, 14, size=64).astype(np.int32)
    valid = (rng.random(64) < 0.7).astype(np.uint8)
    cols = []
    for path in (1, 2):
        p = capi.default_params(n=300, m=64, k=16, beta=2.0, rho=0.25, variant=capi.VARIANT_AS, aging=aging,
                                min_age=10, max_age=12, global_perf=5, min_pop=60, seed=8, deposit_path=path)
        col = make(p, xy=xy)
        col.set_pheromone(tau)
        if aging:
            col.set_aging(age=age, valid=valid, perf=np.zeros(64, np.int32), prev_len=np.zeros(64, np.int64),
                          scalars=np.array([int(valid.sum()), 1, 4, 17], np.int32))
        snap = snapshot(col)
        col.iterate()
        check_iteration(col, p, dist, snap, replay(p, dist, snap))
        cols.append(col)
    try:
        np.testing.assert_array_equal(cols[0].tours(), cols[1].tours())
        np.testing.assert_allclose(cols[0].pheromone(), cols[1].pheromone(), rtol=2e-6)
    finally:
        for c in cols:
            c.close()


@pytest.mark.parametrize("alpha,n", [(1.0, 4200), (1.5, 4200), (1.0, 8200)])
def test_large_n_rank_based_aging_young_ants(alpha, n):
    """n >= 4096 with aging: young ants (pulled-toward-the-mean weights) scan
    rows in their fallbacks while the others use the row top lists; with
    alpha != 1 there are no top lists at all.  At n = 8200 (n x pitch above
    2^26) the colony keeps no per-edge heuristic and young fallbacks gather
    the heuristic table."""
    m, k = 32, 6
    say(f"case: RBAS aging forced state n={n} m={m} k={k} alpha={alpha} (global-memory path)")
    rng = np.random.default_rng(23)
    xy = instances.uniform(n, 100000, seed=23)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, w=4, alpha=alpha, beta=2.0, rho=0.25, variant=capi.VARIANT_RBAS,
                            aging=1, min_age=10, max_age=12, global_perf=5, min_pop=30, seed=98,
                            construct_path=1)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        age = rng.integers(0, 14, size=m).astype(np.int32)
        valid = (rng.random(m) < 0.8).astype(np.uint8)
        col.set_aging(age=age, valid=valid, perf=np.zeros(m, np.int32), prev_len=np.zeros(m, np.int64),
                      scalars=np.array([int(valid.sum()), 1, 4, 17], np.int32))
        col.set_best(10 ** 12, np.arange(n, dtype=np.int32))
        for it in range(2):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))
        listed = col.counters()[3]
        assert (listed > 0) if alpha == 1.0 else (listed == 0)


def test_top_lists_rebuild_after_pheromone_set_from_outside():
    """A row top list keeps only weights above (1 - rho) times the previous
    sweep's 256th weight.  Pheromone scaled down from outside breaks that
    bound: every row must detect it (no list, row scans, same tours), and the
    next sweep selects from scratch."""
    n, m, k = 5000, 16, 6
    say(f"case: top lists after an outside pheromone change n={n} m={m} k={k}")
    rng = np.random.default_rng(51)
    xy = instances.uniform(n, 100000, seed=51)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.2, seed=77, construct_path=1)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        listed = []
        for it in range(3):
            if it == 1:
                col.set_pheromone((col.pheromone() * np.float32(1e-3)).astype(np.float32))
            snap = snapshot(col)
            before = col.counters()
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))
            after = col.counters()
            listed.append((after[0] - before[0], after[3] - before[3]))
        assert listed[0][1] > 0 and listed[2][1] > 0, listed
        assert listed[1][0] > 0 and listed[1][1] == 0, listed  # every row rejected its list


def test_as_deposit_paths_agree_large_n():
    """n > 10240: the adjacency scatter runs in row bands.  Both deposit
    paths from the same pheromone give the same tours and pheromone within
    fp32 rounding, and the banded one is within 1e-5 of an FP64 evaporation
    + deposit of those tours."""
    n, m = 10500, 48
    say(f"case: AS deposit paths n={n} m={m} k=8 (banded adjacency scatter)")
    rng = np.random.default_rng(43)
    xy = instances.uniform(n, 100000, seed=43)
    tau = random_tau(rng, n)
    cols = []
    try:
        for path in (1, 2):
            p = capi.default_params(n=n, m=m, k=8, beta=2.0, rho=0.2, variant=capi.VARIANT_AS, seed=12,
                                    deposit_path=path)
            col = make(p, xy=xy)
            cols.append(col)
            col.set_pheromone(tau)
            col.iterate()
        tours, lengths = cols[1].tours(), cols[1].lengths()
        np.testing.assert_array_equal(cols[0].tours(), tours)
        g1, g2 = cols[0].pheromone(), cols[1].pheromone()
        np.testing.assert_allclose(g1, g2, rtol=2e-6)
        del g1
        st = np.array([MAXD - 1, MAXD, -1, 0], dtype=np.int64)
        best = np.zeros(n, np.int32)
        dep_ant, dep_w = orc.stats(p.variant, tours, lengths, 1, st, best)
        tau64 = tau.astype(np.float64)
        orc.evaporate_f64(tau64, p.rho)
        orc.deposit_f64(tau64, tours, lengths, best, int(st[1]), dep_ant, dep_w)
        rel = np.abs(g2 - tau64) / np.abs(tau64)
        assert rel.max() <= 1e-5, f"pheromone rel err {rel.max():.3e}"
    finally:
        for c in cols:
            c.close()


@pytest.mark.parametrize("n,m,seed", [(70, 20, 1), (300, 24, 2)])
def test_fullrow_iteration_replay(n, m, seed):
    say(f"case: full-row AS n={n} m={m} seed={seed}")
    rng = np.random.default_rng(seed)
    xy = instances.uniform(n, 1000, seed=40 + seed)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, mode=capi.MODE_FULLROW, beta=2.0, rho=0.5, seed=seed)
    with make(p, xy=xy) as col:
        col.set_pheromone(random_tau(rng, n))
        for it in range(2):
            snap = snapshot(col)
            col.iterate()
            check_iteration(col, p, dist, snap, replay(p, dist, snap))


@pytest.mark.parametrize("mode,path", [(capi.MODE_CANDIDATE, 1), (capi.MODE_CANDIDATE, 2),
                                       (capi.MODE_FULLROW, 0)])
def test_deterministic_run_matches_oracle_run(mode, path):
    """RBAS with w=1 deposits one tour, so every pheromone cell receives at
    most one fp32 add per iteration: the whole multi-iteration run is
    deterministic and must equal the oracle's fp32 run bit for bit, through
    the hipGraph replay path."""
    say(f"case: w=1 RBAS graph run 20 iterations mode={mode} path={path}")
    xy = instances.uniform(120, 1000, seed=5)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=120, m=30, k=12, w=1, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS,
                            mode=mode, seed=3, tau0=0.01, construct_path=path)
    with make(p, xy=xy) as col:
        col.run(20)
        _, gb, _ = col.stats()
        secs, ogb, oib, obest = orc.run(dist, 30, variant=orc.RBAS,
                                        mode=orc.MODE_CANDIDATE if mode == capi.MODE_CANDIDATE else orc.MODE_FULLROW,
                                        k=12, w=1, iterations=20, alpha=1.0, beta=2.0, rho=0.2,
                                        tau0=0.01, seed=3)
        assert gb == ogb
        np.testing.assert_array_equal(col.best_tour(), obest)
        tau = col.pheromone()
        assert np.isfinite(tau).all()


# w = 12 and 17: lists longer than one deposit pass (8 entries)
@pytest.mark.parametrize("mode,path,aging,w", [(capi.MODE_CANDIDATE, 2, 0, 6), (capi.MODE_CANDIDATE, 2, 1, 6),
                                               (capi.MODE_CANDIDATE, 1, 1, 6), (capi.MODE_FULLROW, 0, 1, 6),
                                               (capi.MODE_CANDIDATE, 2, 0, 12), (capi.MODE_CANDIDATE, 2, 1, 17)])
def test_ranked_run_matches_oracle_run(mode, path, aging, w):
    """Rank-based AS: the deposit adds the w ranked tours to every pheromone
    cell in list order, as the oracle's CPU loop does, so the whole run
    (with aging: retirements, recruits and young ants) equals the oracle's
    fp32 run bit for bit, and two runs on the GPU are identical."""
    n, m, iters = 150, 40, 45
    say(f"case: w={w} RBAS graph run {iters} iterations mode={mode} path={path} aging={aging}")
    xy = instances.clustered(n, seed=21)
    dist = orc.dist_from_int_coords(xy)
    ag = dict(min_age=5, max_age=8, global_perf=3, min_pop=20)
    p = capi.default_params(n=n, m=m, k=12, w=w, beta=3.0, rho=0.1, variant=capi.VARIANT_RBAS, mode=mode,
                            aging=aging, seed=7, tau0=0.02, construct_path=path, **ag)
    res = []
    for _ in range(2):
        with make(p, xy=xy) as col:
            col.run(iters)
            ib, gb, nv = col.stats()
            res.append((ib, gb, col.best_tour(), col.pheromone()))
    assert res[0][:2] == res[1][:2]
    np.testing.assert_array_equal(res[0][3], res[1][3])
    secs, ogb, oib, obest = orc.run(dist, m, variant=orc.RBAS,
                                    mode=orc.MODE_CANDIDATE if mode == capi.MODE_CANDIDATE else orc.MODE_FULLROW,
                                    k=12, w=w, aging=bool(aging), iterations=iters, alpha=1.0, beta=3.0, rho=0.1,
                                    tau0=0.02, seed=7, **ag)
    ib, gb, best, _ = res[0]
    assert (gb, ib) == (ogb, oib)
    np.testing.assert_array_equal(best, obest)


@pytest.mark.parametrize("n,m,w,aging", [(2, 3, 2, 0), (3, 4, 3, 1), (5, 6, 5, 0), (17, 9, 8, 1), (40, 1, 4, 0),
                                         (64, 70, 8, 1)])
def test_ranked_run_small_matches_oracle_run(n, m, w, aging):
    """Small rank-based colonies through whole graph-replayed iterations:
    one deposit band (n < 16), lists of w = 8 (the most one fused statistics
    and deposit launch takes), a single ant, aging; equal to the oracle's
    fp32 run bit for bit."""
    iters = 30
    k = min(12, n - 1)
    say(f"case: RBAS n={n} m={m} w={w} aging={aging} k={k}, {iters} iterations vs oracle run")
    xy = instances.uniform(n, 1000, seed=60 + n)
    dist = orc.dist_from_int_coords(xy)
    ag = dict(min_age=3, max_age=6, global_perf=2, min_pop=max(1, m // 2))
    p = capi.default_params(n=n, m=m, k=k, w=w, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS, aging=aging,
                            seed=13, tau0=0.05, **ag)
    with make(p, xy=xy) as col:
        col.run(iters)
        ib, gb, _ = col.stats()
        best = col.best_tour()
    _, ogb, oib, obest = orc.run(dist, m, variant=orc.RBAS, mode=orc.MODE_CANDIDATE, k=k, w=w, aging=bool(aging),
                                 iterations=iters, alpha=1.0, beta=2.0, rho=0.2, tau0=0.05, seed=13, **ag)
    assert (gb, ib) == (ogb, oib)
    np.testing.assert_array_equal(best, obest)


def test_ranked_run_large_n_matches_oracle_run():
    """n > 4096 cities: the ordered deposit in 256-row bands; a short whole
    run equals the oracle's fp32 run, with one iteration also replayed
    against the FP64 pheromone."""
    n, m, iters = 5000, 16, 4
    say(f"case: w=6 RBAS n={n} m={m} {iters} iterations vs oracle run")
    xy = instances.uniform(n, 100000, seed=11)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=8, w=6, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS, seed=12,
                            tau0=0.001)
    with make(p, xy=xy) as col:
        col.run(iters - 1)
        snap = snapshot(col)
        col.iterate()
        check_iteration(col, p, dist, snap, replay(p, dist, snap))
        ib, gb, _ = col.stats()
        best = col.best_tour()
    _, ogb, oib, obest = orc.run(dist, m, variant=orc.RBAS, mode=orc.MODE_CANDIDATE, k=8, w=6, iterations=iters,
                                 alpha=1.0, beta=2.0, rho=0.2, tau0=0.001, seed=12)
    assert (gb, ib) == (ogb, oib)
    np.testing.assert_array_equal(best, obest)


def test_graph_and_eager_paths_agree():
    say("case: graph vs eager, AS m=1, 6 iterations")
    xy = instances.uniform(90, 1000, seed=8)
    p = capi.default_params(n=90, m=1, k=10, beta=2.0, rho=0.3, seed=4)
    with make(p, xy=xy) as a, make(p, xy=xy) as b:
        a.run(6)
        for _ in range(6):
            b.iterate()
        np.testing.assert_array_equal(a.tours(), b.tours())
        np.testing.assert_array_equal(a.pheromone(), b.pheromone())
        assert a.stats() == b.stats()


def test_tau0_policies():
    say("case: tau0 policies n=100")
    xy = instances.uniform(100, 1000, seed=12)
    dist = orc.dist_from_int_coords(xy)
    cnn = orc.greedy_length(dist, 0)
    p = capi.default_params(n=100, m=50, k=10, tau0_policy=capi.TAU0_AS)
    with make(p, xy=xy) as col:
        assert col.pheromone()[0, 1] == np.float32(50.0 / cnn)
    p = capi.default_params(n=100, m=50, k=10, w=6, rho=0.1, variant=capi.VARIANT_RBAS,
                            tau0_policy=capi.TAU0_RBAS)
    with make(p, xy=xy) as col:
        assert col.pheromone()[0, 1] == np.float32(0.5 * 6 * 5 / (np.float64(np.float32(0.1)) * cnn))


def test_c2_full_size_properties():
    say("case: C2 full size n=1000 m=1000 k=20 RBAS+aging, 80 iterations (past convergence)")
    xy = instances.clustered(1000, seed=2024)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=1000, m=1000, k=20, w=6, beta=3.0, rho=0.1, variant=capi.VARIANT_RBAS,
                            aging=1, tau0_policy=capi.TAU0_RBAS, seed=1)
    with make(p, xy=xy) as col:
        say(f"config {col.config()}")
        col.run(80)
        ib, gb, nv = col.stats()
        best = col.best_tour()
        assert sorted(best.tolist()) == list(range(1000))
        assert instances.tour_length(dist, best) == gb
        assert gb <= ib
        snap = snapshot(col)
        col.iterate()
        ref = replay(p, dist, snap)
        check_iteration(col, p, dist, snap, ref)


def test_c2_whole_run_matches_oracle_fixture():
    """The n = 1000 rank-based + aging workload at full size (m = 1000,
    k = 20, w = 6) through the hipGraph run path: stopped after 20, 820 and
    2000 iterations, the GPU's best-so-far length, last iteration's best and
    best-so-far tour equal the oracle's whole fp32 run
    (golden/c2_oracle_run.json, written by golden/make_c2_oracle_run.py)."""
    import json
    import os

    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_oracle_run.json")))
    say("case: C2 full size whole run vs oracle fixture, stops " + str([st["iterations"] for st in fx["stops"]]))
    xy = instances.clustered(1000, clusters=10, side=10000, sigma=400.0, seed=1)
    p = capi.default_params(n=1000, m=1000, k=20, w=6, alpha=1.0, beta=3.0, rho=0.1, variant=capi.VARIANT_RBAS,
                            mode=capi.MODE_CANDIDATE, tau0_policy=capi.TAU0_RBAS, aging=1, min_age=30, max_age=50,
                            local_perf=10, global_perf=10, min_pop=40, max_pop=100, seed=fx["seed"])
    with make(p, xy=xy) as col:
        assert col.pheromone()[0, 1] == np.float32(fx["tau0"])
        done = 0
        for st in fx["stops"]:
            col.run(st["iterations"] - done)
            done = st["iterations"]
            ib, gb, _ = col.stats()
            assert (gb, ib) == (st["glob_best"], st["iter_best"]), f"after {done} iterations"
            np.testing.assert_array_equal(col.best_tour(), np.array(st["best_tour"], dtype=np.int32))


def test_c3_full_size_properties():
    """The n = 20000 Ant System workload at full size (m = 1024, k = 32, row
    top lists on): every tour a permutation with its exact length, the best
    tour's length the best-so-far, and, after a few iterations, the next
    iteration's first 16 ants replayed bit for bit by the oracle from the
    GPU's own pheromone (ants are independent under the counter-based RNG).
    The candidate lists fed to the oracle are the GPU's, checked against a
    brute-force stable sort on sampled rows."""
    n, m, k = 20000, 1024, 32
    say(f"case: C3 full size n={n} m={m} k={k} AS, 3 iterations + replay of 16 ants")
    xy = instances.uniform(n, 100000, seed=3)
    dist = orc.dist_from_int_coords(xy)
    p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.2, variant=capi.VARIANT_AS,
                            tau0_policy=capi.TAU0_AS, seed=5)
    with make(p, xy=xy) as col:
        col.run(3)
        fb, _, tours_built, listed = col.counters()
        assert tours_built == 3 * m and fb > 0 and listed > 0
        tours = col.tours()
        np.testing.assert_array_equal(np.sort(tours, axis=1), np.broadcast_to(np.arange(n), (m, n)))
        lens = dist[tours, np.roll(tours, -1, axis=1)].astype(np.int64).sum(axis=1)
        np.testing.assert_array_equal(col.lengths(), lens)
        ib, gb, _ = col.stats()
        assert gb <= ib and instances.tour_length(dist, col.best_tour()) == gb
        nn = col.nn()
        rows = np.random.default_rng(7).choice(n, 40, replace=False)
        for i in rows:
            d = dist[i].astype(np.int64)
            d[i] = np.iinfo(np.int64).max
            np.testing.assert_array_equal(nn[i], np.lexsort((np.arange(n), d))[:k])
        tau = col.pheromone()
        it = col.aging()["iter"]
        col.iterate()
        eta = orc.eta_table(p.beta, orc.max_offdiag(dist))
        choice, choice_nn, _, _ = orc.sweep(tau, dist, eta, nn, p.alpha, p.rho)
        del tau
        np.testing.assert_array_equal(col.choice_nn(), choice_nn)
        q = 16
        ref_tours, ref_len, _ = orc.construct(orc.MODE_CANDIDATE, dist, nn, choice, choice_nn, None, None, eta,
                                              p.alpha, q, p.seed, it, p.stream_id)
        np.testing.assert_array_equal(col.tours()[:q], ref_tours)
        np.testing.assert_array_equal(col.lengths()[:q], ref_len)


def test_c3_deposit_paths_full_size():
    """The n = 20000 workload with the deposit the configuration names (fp32
    atomics, deposit_path 1) and with the adjacency gather: from the same
    pheromone both build the same tours, and after the deposit both pheromone
    matrices match an FP64 evaporation + deposit on sampled rows within 1e-5
    and each other within 2e-6."""
    n, m, k = 20000, 1024, 32
    say(f"case: C3 full size n={n} m={m} k={k}, fp32-atomic vs adjacency deposit, one iteration")
    xy = instances.uniform(n, 100000, seed=3)
    dist = orc.dist_from_int_coords(xy)
    rows = np.sort(np.random.default_rng(11).choice(n, 48, replace=False))
    taus, tours = {}, {}
    for path in (1, 2):
        p = capi.default_params(n=n, m=m, k=k, beta=2.0, rho=0.2, variant=capi.VARIANT_AS,
                                tau0_policy=capi.TAU0_AS, seed=5, deposit_path=path)
        with make(p, xy=xy) as col:
            tau0 = float(col.pheromone()[0, 1])
            col.iterate()
            tours[path] = col.tours()
            lens = col.lengths()
            taus[path] = col.pheromone()[rows].astype(np.float64)
    np.testing.assert_array_equal(tours[1], tours[2])
    lens_ref = dist[tours[1], np.roll(tours[1], -1, axis=1)].astype(np.int64).sum(axis=1)
    np.testing.assert_array_equal(lens, lens_ref)
    # FP64: evaporation of the uniform tau0, then 1/L on both directions of every tour edge
    pos = {int(r): q for q, r in enumerate(rows)}
    ref = np.full((len(rows), n), float(np.float32(np.float32(0.8) * np.float32(tau0))), dtype=np.float64)
    a, b = tours[1], np.roll(tours[1], -1, axis=1)
    w = np.broadcast_to((1.0 / lens_ref.astype(np.float64))[:, None], a.shape)
    for x, y in ((a, b), (b, a)):
        sel = np.isin(x, rows)
        np.add.at(ref, (np.array([pos[int(v)] for v in x[sel]]), y[sel]), w[sel])
    for path in (1, 2):
        rel = np.abs(taus[path] - ref) / ref
        say(f"deposit_path {path}: max rel err vs FP64 {rel.max():.2e}")
        assert rel.max() <= 1e-5
    mutual = (np.abs(taus[1] - taus[2]) / ref).max()
    say(f"atomic vs adjacency: max rel difference {mutual:.2e}")
    assert mutual <= 2e-6


@pytest.mark.parametrize("mode", [capi.MODE_CANDIDATE, capi.MODE_FULLROW])
def test_run_until_halts_on_device_like_the_reference_loop(mode):
    """aco_run_until: the optimum test on the iteration best (src/TSP.cpp:
    168-171) stops the run after the first iteration that reaches it; the
    state equals a colony run for exactly that many iterations, and the colony
    continues normally afterwards.  Rank-based w = 1 makes runs deterministic."""
    say(f"case: run_until RBAS w=1 n=120 m=30 mode={mode}")
    xy = instances.uniform(120, 1000, seed=15)
    p = capi.default_params(n=120, m=30, k=12, w=1, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS, mode=mode,
                            seed=6, tau0=0.01)
    with make(p, xy=xy) as a:
        ib = []
        for _ in range(30):
            a.iterate()
            ib.append(a.stats()[0])
    target = ib[11]
    first = next(t for t, v in enumerate(ib) if v <= target)
    with make(p, xy=xy) as b, make(p, xy=xy) as c:
        ran, halted = b.run_until(30, target, chunk=8)
        assert (ran, halted) == (first + 1, True)
        c.run(first + 1)
        assert b.stats() == c.stats()
        np.testing.assert_array_equal(b.pheromone(), c.pheromone())
        np.testing.assert_array_equal(b.best_tour(), c.best_tour())
        b.run(3)
        c.run(3)
        assert b.stats() == c.stats()
        np.testing.assert_array_equal(b.pheromone(), c.pheromone())
    with make(p, xy=xy) as d:
        assert d.run_until(10, 0, chunk=4) == (10, False)
        assert d.stats()[0] == ib[9]
