"""CPU pins of the learned-edge-prior oracle (oracle/prior_oracle.c).

The prior has no counterpart in the reference, so the oracle is pinned by
brute force, library routines (scipy's normal CDF), closed-form special
cases of the network and finite differences of its own loss."""
import numpy as np
import pytest
from scipy.special import ndtr

import oracle as orc
from acogpu_amd import instances

W1, B1, W2, B2, W3, B3 = 0, 512, 576, 4672, 4736, 4800


def random_params(rng, scale=0.4):
    p = (rng.standard_normal(orc.PRIOR_PARAMS) * scale).astype(np.float32)
    return p


def test_features_brute_force():
    xy = instances.uniform(60, 500, seed=7)
    dist = orc.dist_from_int_coords(xy)
    k = 8
    nn = orc.nn_lists(dist, k)
    got = orc.prior_features(dist, nn)
    n = 60
    rows = {i: list(nn[i]) for i in range(n)}
    nearest = {i: int(dist[i, rows[i][0]]) for i in range(n)}
    kth = {i: int(dist[i, rows[i][-1]]) for i in range(n)}
    total = sum(nearest.values())
    f32 = np.float32
    for i in range(n):
        for r, j in enumerate(rows[i]):
            d = int(dist[i, j])
            rank = rows[j].index(i) if i in rows[j] else k
            want = [f32(d) / f32(kth[i] + 1), f32(d) / f32(kth[j] + 1), f32(r) / f32(k), f32(rank) / f32(k),
                    f32(d * n / (total + n)), f32(nearest[i]) / f32(kth[i] + 1), f32(nearest[j]) / f32(kth[j] + 1),
                    f32(d - nearest[i]) / f32(kth[i] - nearest[i] + 1)]
            np.testing.assert_array_equal(got[i, r], np.array(want, dtype=np.float32), err_msg=f"edge {i},{r}")


def test_labels_full_lists_mark_every_tour_edge():
    n = 40
    xy = instances.uniform(n, 300, seed=3)
    dist = orc.dist_from_int_coords(xy)
    nn = orc.nn_lists(dist, n - 1)
    rng = np.random.default_rng(0)
    tour = rng.permutation(n).astype(np.int32)
    y = orc.prior_labels(nn, tour)
    assert y.sum() == 2 * n
    assert (y.sum(axis=1) == 2).all()
    on = {(int(tour[t]), int(tour[(t + 1) % n])) for t in range(n)}
    on |= {(b, a) for a, b in on}
    for i in range(n):
        for r in range(n - 1):
            assert (y[i, r] == 1.0) == ((i, int(nn[i, r])) in on)


def test_labels_short_lists_subset():
    n = 50
    dist = orc.dist_from_int_coords(instances.uniform(n, 300, seed=4))
    nn_full = orc.nn_lists(dist, n - 1)
    tour = np.random.default_rng(1).permutation(n).astype(np.int32)
    y_full = orc.prior_labels(nn_full, tour)
    y5 = orc.prior_labels(nn_full[:, :5].copy(), tour)
    np.testing.assert_array_equal(y5, y_full[:, :5])


def test_bf16_rounding():
    assert orc.bf16(1.0) == 1.0
    assert orc.bf16(1.0 + 2 ** -8) == 1.0               # tie -> even
    assert orc.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # tie -> even (up)
    assert orc.bf16(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    assert orc.bf16(-3.0e-3) == -orc.bf16(3.0e-3)
    for v in np.random.default_rng(2).standard_normal(200):
        r = orc.bf16(v)
        assert abs(r - v) <= abs(v) * 2 ** -8
        # representable: 8 significant bits
        m, e = np.frexp(r)
        assert float(m * 256) == round(m * 256)


def gelu_as(v):
    """x Phi(x) with Phi from Abramowitz & Stegun 7.1.26, evaluated in numpy
    (the GELU definition; within 7.5e-8 |x| of x ndtr(x))."""
    v = np.asarray(v, dtype=np.float64)
    a1, a2, a3, a4, a5, pp = 0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.3275911
    t = 1.0 / (1.0 + pp * np.abs(v) / np.sqrt(2.0))
    q = 0.5 * (a1 * t + a2 * t ** 2 + a3 * t ** 3 + a4 * t ** 4 + a5 * t ** 5) * np.exp(-0.5 * v * v)
    return v * np.where(v >= 0, 1.0 - q, q)


def test_gelu_against_normal_cdf():
    """GELU with Phi from the erf approximation 7.1.26 of Abramowitz & Stegun
    (|error of erf| <= 1.5e-7): within 7.5e-8 |x| of x ndtr(x) everywhere;
    its gradient Phi(x) + x phi(x) (exact density) is within 7.5e-8 of the
    exact GELU derivative ndtr(x) + x phi(x), and within 6e-7 of central
    finite differences of the approximate GELU (whose own derivative is off
    from the exact one by at most 5.4e-7).  A wrong coefficient, a dropped
    term or a missing 1/sqrt(2 pi) breaks a bound."""
    xs = np.linspace(-14, 14, 2801)
    g = np.array([orc.gelu(x) for x in xs])
    assert np.all(np.abs(g - xs * ndtr(xs)) <= 7.5e-8 * np.abs(xs) + 1e-300)
    assert orc.gelu(0.0) == 0.0 and orc.gelu_grad(0.0) == pytest.approx(0.5, abs=1e-7)
    dg = np.array([orc.gelu_grad(x) for x in xs])
    exact = ndtr(xs) + xs * np.exp(-0.5 * xs * xs) / np.sqrt(2 * np.pi)
    assert np.all(np.abs(dg - exact) <= 7.5e-8)
    for x in np.linspace(-8, 8, 97):
        h = 1e-5
        fd = (orc.gelu(x + h) - orc.gelu(x - h)) / (2 * h)
        assert abs(orc.gelu_grad(x) - fd) <= 6e-7


def test_forward_zero_network_is_b3():
    p = np.zeros(orc.PRIOR_PARAMS, np.float32)
    p[B3] = 0.375
    x = np.random.default_rng(0).random((9, 8)).astype(np.float32)
    for bf in (False, True):
        np.testing.assert_array_equal(orc.mlp_forward(p, x, bf), np.full(9, 0.375))


def test_forward_one_path_closed_form():
    """Only W1[3][5], b1[3], W2[7][3], b2[7], w3[7], b3 nonzero: every other
    hidden unit is gelu(0) = 0, so the output is a scalar chain.  A
    transposed W1 or W2 operand would read input 3 / unit 5 instead."""
    a, b, c, d0, w, e = 1.25, -0.5, 0.75, 0.125, -1.5, 0.25
    p = np.zeros(orc.PRIOR_PARAMS, np.float32)
    p[W1 + 3 * 8 + 5] = a
    p[B1 + 3] = b
    p[W2 + 7 * 64 + 3] = c
    p[B2 + 7] = d0
    p[W3 + 7] = w
    p[B3] = e
    x = np.random.default_rng(5).random((16, 8)).astype(np.float32) * 2
    g = gelu_as
    x64 = x.astype(np.float64)
    assert np.all(np.abs(g(x64) - x64 * ndtr(x64)) <= 7.5e-8 * np.abs(x64))
    want = w * g(c * g(a * x[:, 5].astype(np.float64) + b) + d0) + e
    np.testing.assert_allclose(orc.mlp_forward(p, x, False), want, rtol=1e-12, atol=1e-14)
    # bf16 emulation: x and the hidden activations rounded
    xb = np.array([orc.bf16(v) for v in x[:, 5]])
    h1 = np.array([orc.bf16(g(a * v + b)) for v in xb])
    h2 = np.array([orc.bf16(g(c * v + d0)) for v in h1])
    np.testing.assert_allclose(orc.mlp_forward(p, x, True), w * h2 + e, rtol=1e-12, atol=1e-14)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    p = random_params(rng, 0.3)
    x = (rng.random((23, 8)) * 2).astype(np.float32)
    y = (rng.random(23) < 0.3).astype(np.float32)
    pw = 2.5
    loss, g = orc.mlp_grad(p, x, y, pw, bf16_emulation=False)
    assert np.isfinite(loss) and loss > 0
    idx = np.concatenate([rng.choice(orc.PRIOR_PARAMS, 360, replace=False), [B3, W3, B2, B1, W1]])

    def loss_at(q):
        return orc.mlp_grad(q, x, y, pw, bf16_emulation=False)[0]

    scale = np.abs(g).max()
    for i in idx:
        q1 = p.copy()
        q2 = p.copy()
        q1[i] = np.float32(p[i] + 1e-3)
        q2[i] = np.float32(p[i] - 1e-3)
        delta = float(q1[i]) - float(q2[i])
        fd = (loss_at(q1) - loss_at(q2)) / delta
        assert abs(fd - g[i]) <= 1e-6 * scale + 1e-5 * abs(g[i]), f"param {i}: fd {fd} analytic {g[i]}"


def test_loss_matches_forward_and_bf16_gradient_is_close():
    rng = np.random.default_rng(12)
    p = random_params(rng, 0.3)
    x = (rng.random((40, 8)) * 2).astype(np.float32)
    y = (rng.random(40) < 0.5).astype(np.float32)
    out = orc.mlp_forward(p, x, False)
    s = 1 / (1 + np.exp(-out))
    want = np.mean(-3.0 * y * np.log(s) - (1 - y) * np.log(1 - s))
    loss, g = orc.mlp_grad(p, x, y, 3.0, bf16_emulation=False)
    assert loss == pytest.approx(want, rel=1e-12)
    _, gb = orc.mlp_grad(p, x, y, 3.0, bf16_emulation=True)
    assert np.linalg.norm(gb - g) <= 0.05 * np.linalg.norm(g)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(13)
    p0 = random_params(rng)
    g = (rng.standard_normal(orc.PRIOR_PARAMS) * 0.1).astype(np.float32)
    g[::7] = 0.0
    p = p0.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    orc.adam_step(p, g, m, v, 1, 1e-3)
    want = p0.astype(np.float64) - 1e-3 * g / (np.abs(g) + 1e-8)
    # one fp32 rounding of the parameter plus relative 1e-6 of the step
    assert (np.abs(p - want) <= np.spacing(np.abs(p0)) + 1e-9).all()
    np.testing.assert_array_equal(p[::7], p0[::7])
