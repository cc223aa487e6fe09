"""ctypes binding of the C oracle (oracle/aco_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing
measured as the product.  The C source cites the reference file:line of every
function it restates.
"""
from __future__ import annotations

import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

MAX_DISTANCE = 0x7FFFFFFF
MODE_CANDIDATE, MODE_FULLROW, MODE_REFSEQ = 0, 1, 2
AS, RBAS = 0, 1

_lib = None


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def load(path):
    """A separately built copy of the oracle library (declared, not installed)."""
    L = ctypes.CDLL(path)
    _declare(L)
    return L


@contextlib.contextmanager
def using(path):
    """Routes every call of this module to the library at `path` for the
    duration of the block (the tests' mutation checks build deliberately
    broken copies of the oracle and load them this way)."""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = saved


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, i32, u32, u64, f32, f64 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_float, ctypes.c_double)
    L.orc_philox4x32_10.argtypes = [vp, vp, vp]
    L.orc_rng_word.argtypes = [u64, u32, u32, u32, u32]
    L.orc_rng_word.restype = u32
    L.orc_uniform.argtypes = [u32]
    L.orc_uniform.restype = f32
    L.orc_start_city.argtypes = [u32, i32]
    L.orc_start_city.restype = i32
    L.orc_euc2d_exact.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.orc_euc2d_exact.restype = i32
    L.orc_euc2d_ref.argtypes = [f32, f32, f32, f32]
    L.orc_euc2d_ref.restype = i32
    L.orc_dist_from_int_coords.argtypes = [vp, i32, vp]
    L.orc_dist_from_float_coords.argtypes = [vp, vp, i32, vp]
    L.orc_nn_lists.argtypes = [vp, i32, i32, vp]
    L.orc_pow_alpha.argtypes = [ctypes.c_float, ctypes.c_float]
    L.orc_pow_alpha.restype = ctypes.c_float
    L.orc_eta_table.argtypes = [f64, i32, vp]
    L.orc_eta_table.restype = i32
    L.orc_wave_scan.argtypes = [vp]
    L.orc_select_candidate.argtypes = [vp, f32, vp]
    L.orc_select_candidate.restype = i32
    L.orc_sweep.argtypes = [i32, i32, vp, vp, vp, i32, vp, vp, f32, f32, i32, vp, vp, vp, vp]
    L.orc_construct.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, f32,
                                vp, vp, i32, i32, u64, u32, u32, vp, vp, vp]
    L.orc_stats.argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_stats.restype = i32
    L.orc_deposit_f64.argtypes = [i32, vp, vp, vp, vp, ctypes.c_int64, vp, vp, i32]
    L.orc_evaporate_f64.argtypes = [i32, vp, f32]
    L.orc_aging_start.argtypes = [i32, vp]
    L.orc_aging_control.argtypes = [i32, vp, vp, vp, vp, i32, i32, i32, i32]
    L.orc_greedy_length.argtypes = [vp, i32, i32]
    L.orc_greedy_length.restype = i32
    L.orc_run.argtypes = [i32, i32, vp, vp, vp, u64, vp, vp]
    L.orc_prior_features.argtypes = [i32, i32, vp, vp, vp]
    L.orc_prior_labels.argtypes = [i32, i32, vp, vp, vp]
    for name in ("orc_bf16", "orc_gelu", "orc_gelu_grad"):
        getattr(L, name).argtypes = [f64]
        getattr(L, name).restype = f64
    L.orc_mlp_forward.argtypes = [vp, vp, ctypes.c_int64, i32, vp]
    L.orc_mlp_grad.argtypes = [vp, vp, vp, ctypes.c_int64, f32, i32, vp]
    L.orc_mlp_grad.restype = f64
    L.orc_adam_step.argtypes = [i32, vp, vp, vp, vp, i32, f32, f32, f32, f32]
    L.orc_run.restype = f64


# --------------------------------------------------------------------------
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def rng_word(seed, it, ant, step, stream=0):
    return int(lib().orc_rng_word(seed, it, ant, step, stream))


def uniform(word):
    return float(lib().orc_uniform(word))


def start_city(word, n):
    return int(lib().orc_start_city(word, n))


def euc2d_exact(dx, dy):
    return int(lib().orc_euc2d_exact(int(dx), int(dy)))


def euc2d_ref(xi, xj, yi, yj):
    return int(lib().orc_euc2d_ref(xi, xj, yi, yj))


def dist_from_int_coords(xy):
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1, 2)
    n = xy.shape[0]
    d = np.empty((n, n), dtype=np.int32)
    lib().orc_dist_from_int_coords(_p(xy), n, _p(d))
    return d


def dist_from_float_coords(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n = x.shape[0]
    d = np.empty((n, n), dtype=np.int32)
    lib().orc_dist_from_float_coords(_p(x), _p(y), n, _p(d))
    return d


def nn_lists(dist, k):
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    n = dist.shape[0]
    nn = np.empty((n, k), dtype=np.int32)
    lib().orc_nn_lists(_p(dist), n, k, _p(nn))
    return nn


def max_offdiag(dist):
    d = dist[dist != MAX_DISTANCE]
    return int(d.max()) if d.size else 0


def eta_table(beta, dmax):
    t = np.empty(dmax + 1, dtype=np.float32)
    lib().orc_eta_table(float(beta), int(dmax), _p(t))
    return t


def pow_alpha(x, alpha):
    """tau^alpha as the kernels define it (one call per element)."""
    return np.array([lib().orc_pow_alpha(float(v), float(alpha)) for v in np.ravel(x)], dtype=np.float32)


def wave_scan(v):
    a = np.zeros(64, dtype=np.float32)
    a[: len(v)] = np.asarray(v, dtype=np.float32)
    lib().orc_wave_scan(_p(a))
    return a


def select_candidate(w, u):
    a = np.zeros(64, dtype=np.float32)
    a[: len(w)] = np.asarray(w, dtype=np.float32)
    tot = np.zeros(1, dtype=np.float32)
    r = lib().orc_select_candidate(_p(a), float(u), _p(tot))
    return int(r), float(tot[0])


def sweep(tau, dist, eta, nn, alpha, rho, want_mean=False, heur_nn=None):
    tau = np.ascontiguousarray(tau, dtype=np.float32)
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    n = tau.shape[0]
    k = 0 if nn is None else nn.shape[1]
    choice = np.empty((n, n), dtype=np.float32)
    choice_nn = np.empty((n, max(k, 1)), dtype=np.float32)
    tau_evap = np.empty((n, n), dtype=np.float32)
    mean = np.zeros(n, dtype=np.float32)
    nn_c = None if nn is None else np.ascontiguousarray(nn, dtype=np.int32)
    h = None if heur_nn is None else np.ascontiguousarray(heur_nn, dtype=np.float32)
    lib().orc_sweep(n, k, _p(tau), _p(dist), _p(eta), len(eta), _p(nn_c), _p(h), float(alpha), float(rho),
                    int(want_mean), _p(choice), _p(choice_nn) if k else None, _p(tau_evap), _p(mean))
    return choice, (choice_nn[:, :k] if k else None), tau_evap, mean


def construct(mode, dist, nn, choice, choice_nn, tau, mean, eta, alpha, m, seed, it,
              stream=0, age=None, valid=None, converged=0, min_age=1, tours=None, lengths=None, heur_nn=None):
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    n = dist.shape[0]
    k = nn.shape[1] if nn is not None else 0
    if tours is None:
        tours = np.zeros((m, n), dtype=np.int32)
    if lengths is None:
        lengths = np.zeros(m, dtype=np.int64)
    fb = np.zeros(2, dtype=np.int64)
    nn_c = np.ascontiguousarray(nn, dtype=np.int32) if nn is not None else None
    cand = nn_c if mode == MODE_CANDIDATE else None
    full = nn_c if mode != MODE_CANDIDATE else None
    cnn = np.ascontiguousarray(choice_nn, dtype=np.float32) if choice_nn is not None else None
    age_c = np.ascontiguousarray(age, dtype=np.int32) if age is not None else None
    val_c = np.ascontiguousarray(valid, dtype=np.uint8) if valid is not None else None
    tau_c = np.ascontiguousarray(tau, dtype=np.float32) if tau is not None else None
    mean_c = np.ascontiguousarray(mean, dtype=np.float32) if mean is not None else None
    ch = np.ascontiguousarray(choice, dtype=np.float32) if choice is not None else None
    h = None if heur_nn is None else np.ascontiguousarray(heur_nn, dtype=np.float32)
    lib().orc_construct(mode, n, m, k, _p(dist), _p(cand), _p(full), _p(ch), _p(cnn), _p(h), _p(tau_c),
                        _p(mean_c), _p(eta), len(eta), float(alpha), _p(age_c), _p(val_c),
                        int(converged), int(min_age), int(seed), int(it), int(stream), _p(tours),
                        _p(lengths), _p(fb))
    return tours, lengths, fb


def stats(variant, tours, lengths, w, state64, best_tour, valid=None, perf=None, prev_len=None):
    m, n = tours.shape
    dep_ant = np.zeros(m + w + 1, dtype=np.int32)
    dep_w = np.zeros(m + w + 1, dtype=np.int32)
    val_c = np.ascontiguousarray(valid, dtype=np.uint8) if valid is not None else None
    nd = lib().orc_stats(variant, n, m, w, _p(tours), _p(lengths), _p(val_c), _p(perf),
                         _p(prev_len), _p(state64), _p(best_tour), _p(dep_ant), _p(dep_w))
    return dep_ant[:nd].copy(), dep_w[:nd].copy()


def evaporate_f64(tau64, rho):
    lib().orc_evaporate_f64(tau64.shape[0], _p(tau64), float(rho))


def deposit_f64(tau64, tours, lengths, best_tour, best_len, dep_ant, dep_w):
    dep_ant = np.ascontiguousarray(dep_ant, dtype=np.int32)
    dep_w = np.ascontiguousarray(dep_w, dtype=np.int32)
    lib().orc_deposit_f64(tau64.shape[0], _p(tau64), _p(tours), _p(lengths), _p(best_tour),
                          int(best_len), _p(dep_ant), _p(dep_w), len(dep_ant))


def aging_start(age):
    lib().orc_aging_start(len(age), _p(age))


def aging_control(age, valid, state, gperf, max_age, global_perf, min_pop):
    """state: int32 array [num_valid, converged] updated in place."""
    nv = np.array([state[0]], dtype=np.int32)
    cv = np.array([state[1]], dtype=np.int32)
    lib().orc_aging_control(len(age), _p(age), _p(valid), _p(nv), _p(cv), int(gperf),
                            int(max_age), int(global_perf), int(min_pop))
    state[0], state[1] = int(nv[0]), int(cv[0])


def greedy_length(dist, start=0):
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    return int(lib().orc_greedy_length(_p(dist), dist.shape[0], start))


def run(dist, m, *, variant=AS, mode=MODE_CANDIDATE, k=20, w=6, aging=False, min_age=30,
        max_age=50, global_perf=10, min_pop=40, iterations=10, alpha=1.0, beta=2.0, rho=0.5,
        tau0=1e-6, seed=1, stream=0):
    """Whole run with the reference CPU operation order; returns
    (seconds, glob_best, last_iter_best, best_tour)."""
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    n = dist.shape[0]
    params = np.array([variant, mode, k, w, int(aging), min_age, max_age, global_perf, min_pop,
                       iterations, stream], dtype=np.int32)
    fparams = np.array([alpha, beta, rho, tau0], dtype=np.float32)
    out = np.zeros(2, dtype=np.int64)
    best = np.zeros(n, dtype=np.int32)
    secs = lib().orc_run(n, m, _p(dist), _p(params), _p(fparams), int(seed), _p(out), _p(best))
    if secs < 0:
        raise MemoryError("oracle run: allocation failed")
    return secs, int(out[0]), int(out[1]), best


# ---------------------------------------------------------------------------
# Learned edge prior (prior_oracle.c)

PRIOR_PARAMS = 4801


def prior_features(dist, nn):
    n, k = nn.shape
    d = np.ascontiguousarray(dist, dtype=np.int32)
    q = np.ascontiguousarray(nn, dtype=np.int32)
    out = np.zeros((n, k, 8), dtype=np.float32)
    lib().orc_prior_features(n, k, _p(d), _p(q), _p(out))
    return out


def prior_labels(nn, tour):
    n, k = nn.shape
    q = np.ascontiguousarray(nn, dtype=np.int32)
    t = np.ascontiguousarray(tour, dtype=np.int32)
    out = np.zeros((n, k), dtype=np.float32)
    lib().orc_prior_labels(n, k, _p(q), _p(t), _p(out))
    return out


def bf16(x):
    return lib().orc_bf16(float(x))


def gelu(x):
    return lib().orc_gelu(float(x))


def gelu_grad(x):
    return lib().orc_gelu_grad(float(x))


def mlp_forward(params, x, bf16_emulation=True):
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 8)
    out = np.zeros(len(xx), dtype=np.float64)
    lib().orc_mlp_forward(_p(p), _p(xx), len(xx), int(bf16_emulation), _p(out))
    return out


def mlp_grad(params, x, y, pos_weight=1.0, bf16_emulation=True):
    """Returns (loss, gradient) of the mean weighted BCE with logits."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 8)
    yy = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
    g = np.zeros(PRIOR_PARAMS, dtype=np.float64)
    loss = lib().orc_mlp_grad(_p(p), _p(xx), _p(yy), len(xx), float(pos_weight), int(bf16_emulation), _p(g))
    return loss, g


def adam_step(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """In place on float32 params, m, v."""
    g = np.ascontiguousarray(grad, dtype=np.float32)
    lib().orc_adam_step(len(params), _p(params), _p(g), _p(m), _p(v), int(step), float(lr), float(beta1),
                        float(beta2), float(eps))
