"""ctypes binding of the C ABI (include/acogpu.h) of libacogpu_amd.so.

This is the Python host side used by the tests, smoke() and bench.py.  It
loads the in-tree library built by acogpu_amd/build.py and fails loudly when
it is missing: there is no CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libacogpu_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "acogpu.h")

VARIANT_AS, VARIANT_RBAS = 0, 1
MODE_CANDIDATE, MODE_FULLROW = 0, 1
TAU0_CONST, TAU0_AS, TAU0_RBAS = 0, 1, 2
MAX_DISTANCE = 0x7FFFFFFF


class AcoError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32), ("mode", ctypes.c_int32), ("n", ctypes.c_int32),
        ("m", ctypes.c_int32), ("k", ctypes.c_int32), ("w", ctypes.c_int32),
        ("alpha", ctypes.c_float), ("beta", ctypes.c_float), ("rho", ctypes.c_float),
        ("tau0_policy", ctypes.c_int32), ("tau0", ctypes.c_float), ("aging", ctypes.c_int32),
        ("min_age", ctypes.c_int32), ("max_age", ctypes.c_int32), ("local_perf", ctypes.c_int32),
        ("global_perf", ctypes.c_int32), ("min_pop", ctypes.c_int32), ("max_pop", ctypes.c_int32),
        ("broker", ctypes.c_int32), ("ants_per_block", ctypes.c_int32), ("seed", ctypes.c_uint64),
        ("stream_id", ctypes.c_uint32), ("construct_path", ctypes.c_int32),
        ("deposit_path", ctypes.c_int32), ("ants_per_wave", ctypes.c_int32), ("reserved", ctypes.c_int32 * 4),
    ]


_lib = None

# every entry point declared in include/acogpu.h: name -> (restype, argtypes)
_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
SIGNATURES = {
    "aco_params_init": (ctypes.c_int, [_vp]),
    "aco_create": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "aco_initialize": (ctypes.c_int, [_vp]),
    "aco_iterate": (ctypes.c_int, [_vp, _vp]),
    "aco_run": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_run_until": (ctypes.c_int, [_vp, _i32, _i64, _i32, _vp, _vp]),
    "aco_phase": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_phase_end": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "aco_iterate_stats": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "aco_probe_step": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_profile": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_run_until_time": (ctypes.c_int, [_vp, _i32, _i64, ctypes.c_float, _i32, _vp, _vp, _vp]),
    "aco_get_stats": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "aco_get_best_tour": (ctypes.c_int, [_vp, _vp]),
    "aco_get_tours": (ctypes.c_int, [_vp, _vp]),
    "aco_get_lengths": (ctypes.c_int, [_vp, _vp]),
    "aco_get_pheromone": (ctypes.c_int, [_vp, _vp]),
    "aco_set_pheromone": (ctypes.c_int, [_vp, _vp]),
    "aco_get_choice": (ctypes.c_int, [_vp, _vp]),
    "aco_get_choice_nn": (ctypes.c_int, [_vp, _vp]),
    "aco_get_nn": (ctypes.c_int, [_vp, _vp]),
    "aco_get_mean": (ctypes.c_int, [_vp, _vp]),
    "aco_get_dist": (ctypes.c_int, [_vp, _vp]),
    "aco_get_eta": (ctypes.c_int, [_vp, _vp, _vp]),
    "aco_get_aging": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "aco_set_aging": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "aco_set_best": (ctypes.c_int, [_vp, _i64, _vp]),
    "aco_get_counters": (ctypes.c_int, [_vp, _vp]),
    "aco_set_heuristic_nn": (ctypes.c_int, [_vp, _vp]),
    "aco_get_heuristic_nn": (ctypes.c_int, [_vp, _vp]),
    "aco_batch_set_heuristic_nn": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_get_config": (ctypes.c_int, [_vp, _vp]),
    "aco_profile": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_calibrate_copy": (ctypes.c_int, [_vp, _i64, _i32, _vp]),
    "aco_destroy": (ctypes.c_int, [_vp]),
    "aco_batch_create": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "aco_batch_initialize": (ctypes.c_int, [_vp]),
    "aco_batch_run": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_get_best": (ctypes.c_int, [_vp, _vp, _vp]),
    "aco_batch_get_tours": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_get_lengths": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_get_pheromone": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_set_pheromone": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_get_config": (ctypes.c_int, [_vp, _vp]),
    "aco_batch_get_counters": (ctypes.c_int, [_vp, _vp]),
    "aco_batch_autotune": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_batch_tune_k": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _i32, _i32, ctypes.c_float, _vp, _vp]),
    "aco_batch_exchange_best": (ctypes.c_int, [_vp]),
    "aco_batch_run_islands": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "aco_batch_destroy": (ctypes.c_int, [_vp]),
    "aco_batch_last_error": (ctypes.c_char_p, [_vp]),
    "aco_last_error": (ctypes.c_char_p, [_vp]),
    "aco_version": (ctypes.c_char_p, []),
    "aco_prior_config_init": (None, [_vp]),
    "aco_prior_create": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_get_params": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_set_params": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_collect": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_collect_batch": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_clear": (ctypes.c_int, [_vp]),
    "aco_prior_size": (ctypes.c_int64, [_vp]),
    "aco_prior_train": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_prior_apply": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_apply_batch": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_apply_noalloc": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_features": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "aco_prior_forward": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "aco_prior_gradient": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "aco_prior_adam_step": (ctypes.c_int, [_vp, _vp]),
    "aco_prior_profile": (ctypes.c_int, [_vp, _i32, _vp]),
    "aco_prior_destroy": (ctypes.c_int, [_vp]),
    "aco_prior_last_error": (ctypes.c_char_p, [_vp]),
}

PRIOR_PARAMS = 4801


class PriorConfig(ctypes.Structure):
    """aco_prior_config (include/acogpu.h)."""
    _fields_ = [("seed", ctypes.c_uint64), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("pos_weight", ctypes.c_float),
                ("gamma", ctypes.c_float), ("minibatch", ctypes.c_int32), ("transform", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 7)]


def default_prior_config(**kw) -> PriorConfig:
    c = PriorConfig()
    lib().aco_prior_config_init(ctypes.byref(c))
    for key, val in kw.items():
        if not hasattr(c, key):
            raise AttributeError(key)
        setattr(c, key, val)
    return c


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AcoError(f"{LIB_PATH} is missing: run `python -m acogpu_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_params(**kw) -> Params:
    p = Params()
    lib().aco_params_init(ctypes.byref(p))
    for key, val in kw.items():
        if not hasattr(p, key):
            raise AttributeError(key)
        setattr(p, key, val)
    return p


class Colony:
    """One aco_ctx: the Colony/AntSystem/RankBasedAntSystem state on the GPU."""

    def __init__(self, params: Params, dist=None, xy=None, dist_device_ptr=None):
        """dist: host n x n int32 matrix; xy: host coordinates; dist_device_ptr:
        address of an n x n int32 matrix in device memory (the reference's
        GPU convention)."""
        self._L = lib()
        self.params = params
        self.n, self.m = params.n, params.m
        self._ctx = ctypes.c_void_p()
        d = None if dist is None else np.ascontiguousarray(dist, dtype=np.int32)
        c = None if xy is None else np.ascontiguousarray(xy, dtype=np.int32)
        dp = ctypes.c_void_p(int(dist_device_ptr)) if dist_device_ptr is not None else _p(d)
        rc = self._L.aco_create(ctypes.byref(params), dp, _p(c), ctypes.byref(self._ctx))
        if rc != 0:
            raise AcoError(f"aco_create failed ({rc}): {self._L.aco_last_error(None).decode()}")

    # -- lifecycle ---------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            raise AcoError(f"({rc}) {self._L.aco_last_error(self._ctx).decode()}")

    def close(self):
        if self._ctx:
            self._L.aco_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def initialize(self):
        self._chk(self._L.aco_initialize(self._ctx))

    def iterate(self):
        t = np.zeros(6, dtype=np.float32)
        self._chk(self._L.aco_iterate(self._ctx, _p(t)))
        return t

    def iterate_stats(self):
        """(timers[6], iter_best, glob_best, valid ants) of aco_iterate_stats."""
        t = np.zeros(6, dtype=np.float32)
        ib, gb, nv = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
        self._chk(self._L.aco_iterate_stats(self._ctx, _p(t), ctypes.byref(ib), ctypes.byref(gb), ctypes.byref(nv)))
        return t, ib.value, gb.value, nv.value

    def phase(self, phase, wait=True):
        """One phase; wait=False enqueues it without waiting (complete the
        iteration with phase_end)."""
        if not wait:
            self._chk(self._L.aco_phase(self._ctx, int(phase), None))
            return None
        ms = ctypes.c_float()
        self._chk(self._L.aco_phase(self._ctx, int(phase), ctypes.byref(ms)))
        return ms.value

    def phase_end(self):
        """(timers[6], iter_best, glob_best, valid ants) of aco_phase_end."""
        t = np.zeros(6, dtype=np.float32)
        ib, gb, nv = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
        self._chk(self._L.aco_phase_end(self._ctx, _p(t), ctypes.byref(ib), ctypes.byref(gb), ctypes.byref(nv)))
        return t, ib.value, gb.value, nv.value

    def run(self, iterations):
        ms = ctypes.c_float()
        self._chk(self._L.aco_run(self._ctx, int(iterations), ctypes.byref(ms)))
        return ms.value

    def run_until(self, max_iterations, stop_at, chunk=32):
        """(iterations run, halted) of aco_run_until."""
        ran, halted = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self._L.aco_run_until(self._ctx, int(max_iterations), int(stop_at), int(chunk), ctypes.byref(ran),
                                        ctypes.byref(halted)))
        return ran.value, bool(halted.value)

    def run_until_time(self, max_iterations, stop_at, time_limit_ms, chunk=32):
        """(iterations run, halted code 0/1 optimum/2 time, elapsed ms) of aco_run_until_time."""
        ran, halted, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        self._chk(self._L.aco_run_until_time(self._ctx, int(max_iterations), int(stop_at), float(time_limit_ms),
                                             int(chunk), ctypes.byref(ran), ctypes.byref(halted), ctypes.byref(ms)))
        return ran.value, halted.value, ms.value

    def profile(self, iterations):
        out = np.zeros(4, dtype=np.float32)
        self._chk(self._L.aco_profile(self._ctx, int(iterations), _p(out)))
        return out

    def calibrate_copy(self, nbytes, reps=20):
        ms = ctypes.c_float()
        self._chk(self._L.aco_calibrate_copy(self._ctx, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def probe_step(self, iters=200000):
        """(core clock cycles per step, ns per step) of the one-wave step-chain probe."""
        out = np.zeros(2, dtype=np.float32)
        self._chk(self._L.aco_probe_step(self._ctx, int(iters), _p(out)))
        return float(out[0]), float(out[1])

    # -- state -------------------------------------------------------------
    def stats(self):
        ib, gb, nv = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint32()
        self._chk(self._L.aco_get_stats(self._ctx, ctypes.byref(ib), ctypes.byref(gb), ctypes.byref(nv)))
        return ib.value, gb.value, nv.value

    def best_tour(self):
        t = np.zeros(self.n, dtype=np.int32)
        self._chk(self._L.aco_get_best_tour(self._ctx, _p(t)))
        return t

    def set_best(self, glob_best, tour=None):
        t = None if tour is None else np.ascontiguousarray(tour, dtype=np.int32)
        self._chk(self._L.aco_set_best(self._ctx, int(glob_best), _p(t)))

    def tours(self):
        t = np.zeros((self.m, self.n), dtype=np.int32)
        self._chk(self._L.aco_get_tours(self._ctx, _p(t)))
        return t

    def lengths(self):
        t = np.zeros(self.m, dtype=np.int64)
        self._chk(self._L.aco_get_lengths(self._ctx, _p(t)))
        return t

    def pheromone(self):
        t = np.zeros((self.n, self.n), dtype=np.float32)
        self._chk(self._L.aco_get_pheromone(self._ctx, _p(t)))
        return t

    def set_pheromone(self, tau):
        t = np.ascontiguousarray(tau, dtype=np.float32)
        self._chk(self._L.aco_set_pheromone(self._ctx, _p(t)))

    def choice(self):
        t = np.zeros((self.n, self.n), dtype=np.float32)
        self._chk(self._L.aco_get_choice(self._ctx, _p(t)))
        return t

    def choice_nn(self):
        t = np.zeros((self.n, self.params.k), dtype=np.float32)
        self._chk(self._L.aco_get_choice_nn(self._ctx, _p(t)))
        return t

    def nn(self):
        cols = self.params.k if self.params.mode == MODE_CANDIDATE else self.n
        t = np.zeros((self.n, cols), dtype=np.int32)
        self._chk(self._L.aco_get_nn(self._ctx, _p(t)))
        return t

    def mean(self):
        t = np.zeros(self.n, dtype=np.float32)
        self._chk(self._L.aco_get_mean(self._ctx, _p(t)))
        return t

    def dist(self):
        t = np.zeros((self.n, self.n), dtype=np.int32)
        self._chk(self._L.aco_get_dist(self._ctx, _p(t)))
        return t

    def eta(self):
        ln = ctypes.c_int32()
        self._chk(self._L.aco_get_eta(self._ctx, None, ctypes.byref(ln)))
        t = np.zeros(ln.value, dtype=np.float32)
        self._chk(self._L.aco_get_eta(self._ctx, _p(t), ctypes.byref(ln)))
        return t

    def aging(self):
        age = np.zeros(self.m, dtype=np.int32)
        valid = np.zeros(self.m, dtype=np.uint8)
        perf = np.zeros(self.m, dtype=np.int32)
        prev = np.zeros(self.m, dtype=np.int64)
        sc = np.zeros(4, dtype=np.int32)
        self._chk(self._L.aco_get_aging(self._ctx, _p(age), _p(valid), _p(perf), _p(prev), _p(sc)))
        return {"age": age, "valid": valid, "perf": perf, "prev_len": prev, "num_valid": int(sc[0]),
                "converged": int(sc[1]), "gperf": int(sc[2]), "iter": int(sc[3])}

    def set_aging(self, age=None, valid=None, perf=None, prev_len=None, scalars=None):
        conv = lambda a, dt: None if a is None else np.ascontiguousarray(a, dtype=dt)
        self._chk(self._L.aco_set_aging(self._ctx, _p(conv(age, np.int32)), _p(conv(valid, np.uint8)),
                                        _p(conv(perf, np.int32)), _p(conv(prev_len, np.int64)),
                                        _p(conv(scalars, np.int32))))

    def set_heuristic_nn(self, h=None):
        t = None if h is None else np.ascontiguousarray(h, dtype=np.float32)
        self._chk(self._L.aco_set_heuristic_nn(self._ctx, _p(t)))

    def heuristic_nn(self):
        t = np.zeros((self.n, self.params.k), dtype=np.float32)
        self._chk(self._L.aco_get_heuristic_nn(self._ctx, _p(t)))
        return t

    def config(self):
        t = np.zeros(4, dtype=np.int32)
        self._chk(self._L.aco_get_config(self._ctx, _p(t)))
        return {"path": int(t[0]), "ants_per_block": int(t[1]), "young_blocks": int(t[2]),
                "max_lds": int(t[3])}

    def counters(self):
        """(fallback steps, cities compared by fallback steps, tours constructed,
        fallback steps answered by a row top list)"""
        t = np.zeros(4, dtype=np.uint64)
        self._chk(self._L.aco_get_counters(self._ctx, _p(t)))
        return int(t[0]), int(t[1]), int(t[2]), int(t[3])


class Batch:
    """One aco_batch: many independent instances solved together on the GPU."""

    def __init__(self, params: Params, n_list, xy_list, m_list=None):
        self._L = lib()
        self.params = params
        self.n = np.ascontiguousarray(n_list, dtype=np.int32)
        self.count = len(self.n)
        self.m = self.n.copy() if m_list is None else np.ascontiguousarray(m_list, dtype=np.int32)
        xy = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32).reshape(-1) for c in xy_list]))
        self._ctx = ctypes.c_void_p()
        mp = None if m_list is None else self.m
        rc = self._L.aco_batch_create(ctypes.byref(params), self.count, _p(self.n), _p(mp), _p(xy),
                                      ctypes.byref(self._ctx))
        if rc != 0:
            raise AcoError(f"aco_batch_create failed ({rc}): {self._L.aco_batch_last_error(None).decode()}")
        self.city_off = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)

    def _chk(self, rc):
        if rc != 0:
            raise AcoError(f"({rc}) {self._L.aco_batch_last_error(self._ctx).decode()}")

    def close(self):
        if self._ctx:
            self._L.aco_batch_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def initialize(self):
        self._chk(self._L.aco_batch_initialize(self._ctx))

    def run(self, iterations):
        ms = ctypes.c_float()
        self._chk(self._L.aco_batch_run(self._ctx, int(iterations), ctypes.byref(ms)))
        return ms.value

    def profile(self, iterations):
        """Average HIP-event ms of {sweep, construction, statistics, deposit}."""
        out = np.zeros(4, dtype=np.float32)
        self._chk(self._L.aco_batch_profile(self._ctx, int(iterations), _p(out)))
        return out

    def autotune(self, reps=3):
        w = ctypes.c_int32()
        self._chk(self._L.aco_batch_autotune(self._ctx, int(reps), ctypes.byref(w)))
        return w.value

    def exchange_best(self):
        self._chk(self._L.aco_batch_exchange_best(self._ctx))

    def run_islands(self, iterations, exchange_every):
        ms = ctypes.c_float()
        self._chk(self._L.aco_batch_run_islands(self._ctx, int(iterations), int(exchange_every), ctypes.byref(ms)))
        return ms.value

    def best_lengths(self):
        gb = np.zeros(self.count, dtype=np.int64)
        self._chk(self._L.aco_batch_get_best(self._ctx, _p(gb), None))
        return gb

    def best(self):
        gb = np.zeros(self.count, dtype=np.int64)
        tours = np.zeros(int(self.n.sum()), dtype=np.int32)
        self._chk(self._L.aco_batch_get_best(self._ctx, _p(gb), _p(tours)))
        return gb, [tours[o:o + n] for o, n in zip(self.city_off, self.n)]

    def tours(self, i):
        t = np.zeros((int(self.m[i]), int(self.n[i])), dtype=np.int32)
        self._chk(self._L.aco_batch_get_tours(self._ctx, int(i), _p(t)))
        return t

    def lengths(self, i):
        t = np.zeros(int(self.m[i]), dtype=np.int64)
        self._chk(self._L.aco_batch_get_lengths(self._ctx, int(i), _p(t)))
        return t

    def pheromone(self, i):
        t = np.zeros((int(self.n[i]), int(self.n[i])), dtype=np.float32)
        self._chk(self._L.aco_batch_get_pheromone(self._ctx, int(i), _p(t)))
        return t

    def set_pheromone(self, i, tau):
        t = np.ascontiguousarray(tau, dtype=np.float32)
        self._chk(self._L.aco_batch_set_pheromone(self._ctx, int(i), _p(t)))

    def set_heuristic_nn(self, i, h):
        t = np.ascontiguousarray(h, dtype=np.float32)
        self._chk(self._L.aco_batch_set_heuristic_nn(self._ctx, int(i), _p(t)))

    def counters(self):
        """(count, 2) uint64: fallback steps and cities scanned by them, per instance."""
        t = np.zeros((self.count, 2), dtype=np.uint64)
        self._chk(self._L.aco_batch_get_counters(self._ctx, _p(t)))
        return t

    def config(self):
        t = np.zeros(5, dtype=np.int32)
        self._chk(self._L.aco_batch_get_config(self._ctx, _p(t)))
        return {"ants_per_block": int(t[0]), "blocks": int(t[1]), "items": int(t[2]), "lds_bytes": int(t[3]),
                "ants_per_wave": int(t[4])}


def tune_k(params: Params, n_list, xy_list, ks, trial_iterations=20, tolerance=0.005, m_list=None):
    """aco_batch_tune_k: (chosen k, report rows {ms/iteration, mean best, ants per block})."""
    L = lib()
    n = np.ascontiguousarray(n_list, dtype=np.int32)
    m = None if m_list is None else np.ascontiguousarray(m_list, dtype=np.int32)
    xy = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32).reshape(-1) for c in xy_list]))
    k = np.ascontiguousarray(ks, dtype=np.int32)
    rep = np.zeros((len(k), 3), dtype=np.float32)
    chosen = ctypes.c_int32()
    rc = L.aco_batch_tune_k(ctypes.byref(params), len(n), _p(n), _p(m), _p(xy), _p(k), len(k), int(trial_iterations),
                            float(tolerance), ctypes.byref(chosen), _p(rep))
    if rc != 0:
        raise AcoError(f"aco_batch_tune_k failed ({rc}): {L.aco_batch_last_error(None).decode()}")
    return int(chosen.value), rep


class Prior:
    """One aco_prior: the learned edge prior (MLP on candidate-edge features)."""

    def __init__(self, config: PriorConfig | None = None):
        self._L = lib()
        self.config = config if config is not None else default_prior_config()
        self._ctx = ctypes.c_void_p()
        rc = self._L.aco_prior_create(ctypes.byref(self.config), ctypes.byref(self._ctx))
        if rc != 0:
            raise AcoError(f"aco_prior_create failed ({rc}): {self._L.aco_prior_last_error(None).decode()}")

    def _chk(self, rc):
        if rc != 0:
            raise AcoError(f"({rc}) {self._L.aco_prior_last_error(self._ctx).decode()}")

    def close(self):
        if self._ctx:
            self._L.aco_prior_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self):
        t = np.zeros(PRIOR_PARAMS, dtype=np.float32)
        self._chk(self._L.aco_prior_get_params(self._ctx, _p(t)))
        return t

    def set_params(self, params):
        t = np.ascontiguousarray(params, dtype=np.float32)
        assert t.shape == (PRIOR_PARAMS,)
        self._chk(self._L.aco_prior_set_params(self._ctx, _p(t)))

    def collect(self, colony: "Colony"):
        self._chk(self._L.aco_prior_collect(self._ctx, colony._ctx))

    def collect_batch(self, batch: "Batch"):
        self._chk(self._L.aco_prior_collect_batch(self._ctx, batch._ctx))

    def clear(self):
        self._chk(self._L.aco_prior_clear(self._ctx))

    def size(self) -> int:
        return int(self._L.aco_prior_size(self._ctx))

    def train(self, epochs: int):
        loss = np.zeros(max(epochs, 1), dtype=np.float32)
        self._chk(self._L.aco_prior_train(self._ctx, int(epochs), _p(loss)))
        return loss[:epochs]

    def apply(self, colony: "Colony"):
        self._chk(self._L.aco_prior_apply(self._ctx, colony._ctx))

    def apply_batch(self, batch: "Batch"):
        self._chk(self._L.aco_prior_apply_batch(self._ctx, batch._ctx))

    def apply_noalloc(self, colony: "Colony"):
        """aco_prior_apply with no device allocation (library scratch array)."""
        self._chk(self._L.aco_prior_apply_noalloc(self._ctx, colony._ctx))

    def features(self, colony: "Colony"):
        n, k = colony.n, colony.params.k
        f = np.zeros((n, k, 8), dtype=np.float32)
        y = np.zeros((n, k), dtype=np.float32)
        self._chk(self._L.aco_prior_features(self._ctx, colony._ctx, _p(f), _p(y)))
        return f, y

    def forward(self, x):
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(len(xx), dtype=np.float32)
        self._chk(self._L.aco_prior_forward(self._ctx, _p(xx), len(xx), _p(out)))
        return out

    def gradient(self, x, y):
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 8)
        yy = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        g = np.zeros(PRIOR_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        self._chk(self._L.aco_prior_gradient(self._ctx, _p(xx), _p(yy), len(xx), _p(g), ctypes.byref(loss)))
        return float(loss.value), g

    def profile(self, reps=20):
        """(forward ms, backward ms, edges) per launch over one minibatch."""
        t = np.zeros(3, dtype=np.float32)
        self._chk(self._L.aco_prior_profile(self._ctx, int(reps), _p(t)))
        return float(t[0]), float(t[1]), int(t[2])

    def adam_step(self, grad):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        self._chk(self._L.aco_prior_adam_step(self._ctx, _p(g)))
