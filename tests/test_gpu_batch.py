"""GPU parity of the batched solver: every instance of a batch must behave as
a single colony with stream_id = instance index (itself pinned against the
oracle), through the persistent construction launch and the graph replay."""
import numpy as np
import pytest

import oracle as orc
from acogpu_amd import capi, instances

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_batch_inputs(count, lo, hi, seed, side=1000):
    sizes = instances.batched_sizes(count, lo, hi, seed=seed)
    xys = [instances.uniform(int(n), side, seed=seed * 100000 + i) for i, n in enumerate(sizes)]
    return sizes, xys


def batch_params(**kw):
    base = dict(n=0, m=0, k=16, beta=2.0, rho=0.5, variant=capi.VARIANT_AS, tau0_policy=capi.TAU0_AS, seed=77)
    base.update(kw)
    return capi.default_params(**base)


# (ants per block, ants per wave, k): one ant per wave, and 2 or 4 ants
# sharing a wave (k <= 32 or 16), where k below the group width leaves lanes
# idle
@pytest.mark.parametrize("apb,apw,k", [(1, 1, 16), (4, 1, 16), (8, 1, 16), (4, 2, 16), (8, 4, 16), (16, 4, 16),
                                       (2, 2, 19), (8, 4, 9)])
def test_batch_first_iterations_equal_single_colonies(apb, apw, k):
    say(f"case: batch of 10 vs single colonies, ants_per_block={apb} ants_per_wave={apw} k={k}")
    sizes, xys = make_batch_inputs(10, 20, 300, seed=1)
    p = batch_params(ants_per_block=apb, ants_per_wave=apw, k=k)
    with capi.Batch(p, sizes, xys) as b:
        b.initialize()
        assert b.config()["ants_per_wave"] == apw
        b.run(1)
        gb, best = b.best()
        for i in (0, 3, 7, 9):
            n = int(sizes[i])
            q = batch_params(n=n, m=n, stream_id=i, k=k)
            with capi.Colony(q, xy=xys[i]) as c:
                c.initialize()
                c.iterate()
                np.testing.assert_array_equal(b.tours(i), c.tours())
                np.testing.assert_array_equal(b.lengths(i), c.lengths())
                assert gb[i] == c.stats()[1]
                np.testing.assert_array_equal(best[i], c.best_tour())
                np.testing.assert_allclose(b.pheromone(i), c.pheromone(), rtol=1e-6)


# two ants per wave over its whole range: n in (512, 1024] (bitmap words 16-31
# of a group, visited lookups for cities >= 512, the 32-word fallback), k 20
# and 32, a fallback-heavy k = 3, sizes not multiples of 32
@pytest.mark.parametrize("k,sizes", [(20, [530, 611, 777, 803, 999, 1000]), (32, [517, 577, 630]),
                                     (3, [700, 901, 1000])])
def test_two_ants_per_wave_large_instances_equal_one_ant_per_wave(k, sizes):
    say(f"case: ants_per_wave 2 vs 1, k={k}, n={sizes}, 3 iterations")
    sizes = np.array(sizes, dtype=np.int32)
    xys = [instances.uniform(int(n), 1000, seed=900 + 7 * i + k) for i, n in enumerate(sizes)]
    got = {}
    for apw in (1, 2):
        with capi.Batch(batch_params(k=k, ants_per_wave=apw, variant=capi.VARIANT_RBAS, w=1,
                                     tau0_policy=capi.TAU0_CONST, tau0=0.001), sizes, xys) as b:
            b.initialize()
            assert b.config()["ants_per_wave"] == apw
            b.run(3)
            got[apw] = ([b.tours(i) for i in range(len(sizes))], [b.lengths(i) for i in range(len(sizes))],
                        b.counters(), b.best()[0])
    for i in range(len(sizes)):
        np.testing.assert_array_equal(got[1][0][i], got[2][0][i], err_msg=f"instance {i} tours")
        np.testing.assert_array_equal(got[1][1][i], got[2][1][i])
    np.testing.assert_array_equal(got[1][2], got[2][2])
    np.testing.assert_array_equal(got[1][3], got[2][3])
    if k == 3:
        assert (got[2][2][:, 0] > 0).all()  # fallbacks on every instance
    # the largest instance as a single colony (stream = instance index)
    i = len(sizes) - 1
    n = int(sizes[i])
    q = batch_params(n=n, m=n, k=k, stream_id=i, variant=capi.VARIANT_RBAS, w=1, tau0_policy=capi.TAU0_CONST,
                     tau0=0.001)
    with capi.Colony(q, xy=xys[i]) as c:
        c.initialize()
        c.run(3)
        np.testing.assert_array_equal(c.tours(), got[2][0][i])


def test_batch_above_2048_cities_equals_single_colonies():
    """n > 2048: the visited set no longer fits one register per lane and
    the batch takes its LDS-bitmap construction kernel."""
    say("case: batch of 2, n in {2100, 2300}, m = 8, k = 8 vs single colonies")
    sizes = np.array([2100, 2300], dtype=np.int32)
    xys = [instances.uniform(int(n), 10000, seed=700 + i) for i, n in enumerate(sizes)]
    p = batch_params(k=8)
    with capi.Batch(p, sizes, xys, [8, 8]) as b:
        b.initialize()
        assert b.config()["ants_per_wave"] == 1 and b.config()["ants_per_block"] <= 8
        b.run(1)
        for i in range(2):
            n = int(sizes[i])
            with capi.Colony(batch_params(n=n, m=8, k=8, stream_id=i), xy=xys[i]) as c:
                c.initialize()
                c.iterate()
                np.testing.assert_array_equal(b.tours(i), c.tours())
                np.testing.assert_array_equal(b.lengths(i), c.lengths())
        with pytest.raises(capi.AcoError):
            capi.Batch(batch_params(k=8, ants_per_block=16), sizes, xys, [8, 8])


@pytest.mark.parametrize("deposit_path", [1, 2])
def test_batch_iteration_replay_through_oracle(deposit_path):
    say(f"case: batch of 16, iteration 4 replayed through the oracle, deposit_path={deposit_path}")
    sizes, xys = make_batch_inputs(16, 30, 250, seed=2)
    p = batch_params(deposit_path=deposit_path)
    with capi.Batch(p, sizes, xys) as b:
        b.initialize()
        b.run(3)
        gb0, best0 = b.best()
        taus = {i: b.pheromone(i) for i in (1, 5, 11, 15)}
        b.run(1)
        gb1, best1 = b.best()
        for i, tau in taus.items():
            n = int(sizes[i])
            dist = orc.dist_from_int_coords(xys[i])
            eta = orc.eta_table(2.0, orc.max_offdiag(dist))
            nn = orc.nn_lists(dist, 16)
            choice, cnn, _, _ = orc.sweep(tau, dist, eta, nn, 1.0, 0.5)
            tours, lengths, _ = orc.construct(orc.MODE_CANDIDATE, dist, nn, choice, cnn, tau, None, eta, 1.0, n,
                                              77, 3, stream=i)
            np.testing.assert_array_equal(b.tours(i), tours)
            np.testing.assert_array_equal(b.lengths(i), lengths)
            st = np.array([orc.MAX_DISTANCE - 1, gb0[i], -1, 0], dtype=np.int64)
            bt = best0[i].copy()
            dep_ant, dep_w = orc.stats(orc.AS, tours, lengths, 1, st, bt)
            assert gb1[i] == st[1]
            np.testing.assert_array_equal(best1[i], bt)
            tau64 = tau.astype(np.float64)
            orc.evaporate_f64(tau64, 0.5)
            orc.deposit_f64(tau64, tours, lengths, bt, int(st[1]), dep_ant, dep_w)
            got = b.pheromone(i).astype(np.float64)
            assert (np.abs(got - tau64) / tau64).max() <= 1e-5


def test_batch_deterministic_rank_based_run_matches_oracle():
    """Rank-based with w = 1 deposits one tour per instance: the 12-iteration
    run is deterministic and equals the oracle's fp32 run per instance."""
    say("case: batch of 8, RBAS w=1, 12 iterations vs oracle runs")
    sizes, xys = make_batch_inputs(8, 40, 200, seed=3)
    p = batch_params(variant=capi.VARIANT_RBAS, w=1, tau0_policy=capi.TAU0_CONST, tau0=0.01, rho=0.2)
    with capi.Batch(p, sizes, xys) as b:
        b.initialize()
        b.run(12)
        gb, best = b.best()
        for i in range(8):
            dist = orc.dist_from_int_coords(xys[i])
            n = int(sizes[i])
            _, ogb, _, obest = orc.run(dist, n, variant=orc.RBAS, mode=orc.MODE_CANDIDATE, k=16, w=1,
                                       iterations=12, alpha=1.0, beta=2.0, rho=0.2, tau0=0.01, seed=77, stream=i)
            assert gb[i] == ogb
            np.testing.assert_array_equal(best[i], obest)


def test_batch_autotune_changes_only_the_launch():
    say("case: autotune on a batch of 24 vs one ant per wave")
    sizes, xys = make_batch_inputs(24, 50, 500, seed=4)
    with capi.Batch(batch_params(ants_per_block=1, ants_per_wave=1), sizes, xys) as a, \
            capi.Batch(batch_params(ants_per_block=1), sizes, xys) as b:
        a.initialize()
        b.initialize()
        w = b.autotune(reps=2)
        assert w in (1, 2, 4, 8, 16) and b.config()["ants_per_block"] == w
        assert b.config()["ants_per_wave"] in (1, 4)
        a.run(1)
        b.run(1)
        for i in range(24):
            np.testing.assert_array_equal(a.tours(i), b.tours(i))


@pytest.mark.parametrize("apw", [2, 4])
def test_batch_ants_sharing_a_wave_through_fallbacks(apw):
    """k = 3 leaves most steps with every candidate visited: the argmax
    fallback and the rounding path run for some groups of a wave while the
    others take the roulette.  Tours, lengths and the fallback counters equal
    the one-ant-per-wave launch's; n from 5 (fewer steps than a group has
    lanes) up, m not a multiple of the ants per item."""
    say(f"case: batch of 40, k=3, n in [5, 200], ants_per_wave={apw} vs 1, 3 iterations")
    sizes, xys = make_batch_inputs(40, 5, 200, seed=11)
    ms = [max(1, int(n) // 3 + 1) for n in sizes]
    with capi.Batch(batch_params(k=3, ants_per_wave=1, ants_per_block=4), sizes, xys, ms) as a, \
            capi.Batch(batch_params(k=3, ants_per_wave=apw, ants_per_block=4), sizes, xys, ms) as b:
        assert b.config()["ants_per_wave"] == apw
        a.initialize()
        b.initialize()
        for it in range(3):
            a.run(1)
            b.run(1)
            for i in range(40):
                np.testing.assert_array_equal(a.tours(i), b.tours(i))
                np.testing.assert_array_equal(a.lengths(i), b.lengths(i))
            np.testing.assert_array_equal(a.counters(), b.counters())
            # the next iteration starts from the same pheromone in both
            for i in range(40):
                b.set_pheromone(i, a.pheromone(i))
        i = int(np.argmax(sizes))
        dist = orc.dist_from_int_coords(xys[i])
        for t, l in zip(b.tours(i), b.lengths(i)):
            assert sorted(t.tolist()) == list(range(int(sizes[i])))
            assert instances.tour_length(dist, t) == l


def test_batch_4096_uneven_instances_properties():
    say("case: 4096 instances n log-uniform [50, 500], m = n, k = 16, 4 iterations")
    sizes, xys = make_batch_inputs(4096, 50, 500, seed=5)
    p = batch_params()
    with capi.Batch(p, sizes, xys) as b:
        b.initialize()
        say(f"config {b.config()}")
        b.run(4)
        gb, best = b.best()
        rng = np.random.default_rng(0)
        for i in rng.choice(4096, size=24, replace=False):
            n = int(sizes[i])
            assert sorted(best[i].tolist()) == list(range(n))
            dist = orc.dist_from_int_coords(xys[i])
            assert instances.tour_length(dist, best[i]) == gb[i]


def test_tune_k_picks_fastest_k_within_tolerance():
    say("case: candidate-width tuner over k in {6, 10, 16, 24} on 48 instances, 8 trial iterations")
    sizes, xys = make_batch_inputs(48, 40, 300, seed=6)
    p = batch_params()
    ks = [6, 10, 16, 24]
    k, rep = capi.tune_k(p, sizes, xys, ks, trial_iterations=8, tolerance=0.01)
    say(f"chosen k={k}; report {rep.tolist()}")
    assert k in ks
    assert (rep[:, 0] > 0).all() and (rep[:, 1] > 0).all() and set(rep[:, 2].tolist()) <= {1.0, 2.0, 4.0, 8.0, 16.0}
    ok = rep[:, 1] <= rep[:, 1].min() * 1.01 * (1 + 1e-6)
    assert ok[ks.index(k)]
    assert rep[ks.index(k), 0] <= rep[ok, 0].min() * (1 + 1e-6)
    # the trial means are those of ordinary batches with that k
    q = batch_params(k=k)
    with capi.Batch(q, sizes, xys) as b:
        b.initialize()
        b.run(8)
        gb, _ = b.best()
    assert abs(gb.mean() - rep[ks.index(k), 1]) <= 0.02 * gb.mean()


def test_islands_exchange_equals_single_colonies_given_the_shared_best():
    """Island model: 6 copies of one instance (island i = stream i).  After an
    exchange every island holds the shortest best-so-far tour; the following
    iterations equal single colonies whose best-so-far was set to it
    (rank-based w = 1: deterministic)."""
    say("case: 6 islands of n=150, RBAS w=1, exchange after 5 iterations, 4 more")
    xy = instances.uniform(150, 1000, seed=77)
    p = batch_params(variant=capi.VARIANT_RBAS, w=1, k=12, tau0_policy=capi.TAU0_CONST, tau0=0.01, rho=0.2)
    with capi.Batch(p, [150] * 6, [xy] * 6, m_list=[40] * 6) as b:
        b.initialize()
        b.run(5)
        gb0, best0 = b.best()
        win = int(np.argmin(gb0))
        b.exchange_best()
        gb1, best1 = b.best()
        assert (gb1 == gb0[win]).all()
        for t in best1:
            np.testing.assert_array_equal(t, best0[win])
        b.run(4)
        gb2, best2 = b.best()
        for i in (0, 2, 5):
            q = capi.default_params(n=150, m=40, k=12, variant=capi.VARIANT_RBAS, w=1, beta=2.0, rho=0.2,
                                    tau0_policy=capi.TAU0_CONST, tau0=0.01, seed=77, stream_id=i)
            with capi.Colony(q, xy=xy) as c:
                c.initialize()
                c.run(5)
                assert c.stats()[1] == gb0[i]
                c.set_best(int(gb0[win]), best0[win])
                c.run(4)
                np.testing.assert_array_equal(b.tours(i), c.tours())
                assert c.stats()[1] == gb2[i]
                np.testing.assert_array_equal(c.best_tour(), best2[i])


@pytest.mark.parametrize("w", [6, 11])
def test_ranked_batch_equals_single_colonies(w):
    """Rank-based AS: the batch deposit adds each instance's ranked tours in
    list order, as the single colony does, so whole runs of uneven instances
    (several sharing one 64-row deposit band, and w = 11 taking two deposit
    passes) equal single colonies bit for bit."""
    sizes = np.array([50, 57, 61, 130, 333], dtype=np.int32)
    say(f"case: RBAS w={w} batch n={sizes.tolist()}, 15 iterations vs single colonies")
    xys = [instances.uniform(int(n), 1000, seed=300 + i) for i, n in enumerate(sizes)]
    kw = dict(variant=capi.VARIANT_RBAS, w=w, k=12, tau0_policy=capi.TAU0_RBAS, rho=0.1)
    with capi.Batch(batch_params(**kw), sizes, xys) as b:
        b.initialize()
        b.run(15)
        gb, best = b.best()
        for i, n in enumerate(sizes):
            with capi.Colony(batch_params(n=int(n), m=int(n), stream_id=i, **kw), xy=xys[i]) as c:
                c.initialize()
                c.run(15)
                np.testing.assert_array_equal(b.tours(i), c.tours(), err_msg=f"instance {i}")
                assert c.stats()[1] == gb[i]
                np.testing.assert_array_equal(c.best_tour(), best[i])


def test_run_islands_ends_with_a_shared_best():
    say("case: run_islands 8 islands n=120, 12 iterations, exchange every 4")
    xy = instances.uniform(120, 1000, seed=78)
    p = batch_params(variant=capi.VARIANT_RBAS, w=6, k=12, tau0_policy=capi.TAU0_RBAS, rho=0.1)
    with capi.Batch(p, [120] * 8, [xy] * 8) as b:
        b.initialize()
        b.run_islands(12, 4)
        gb, best = b.best()
        assert (gb == gb.min()).all()
        dist = orc.dist_from_int_coords(xy)
        for t in best:
            assert sorted(t.tolist()) == list(range(120))
            assert instances.tour_length(dist, t) == gb[0]
    sizes, xys = make_batch_inputs(3, 50, 80, seed=9)
    with capi.Batch(batch_params(), sizes, xys) as b:
        b.initialize()
        if len(set(sizes.tolist())) > 1:
            with pytest.raises(capi.AcoError):
                b.exchange_best()
