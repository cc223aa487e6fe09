"""GPU tests of the drop-in boundary: the reference's device-pointer
distance matrix (src/Setup.cpp:60-66), the one-wait phased iteration behind
Colony::constructSolution, the time limit of TSP::terminationCondition
(src/TSP.cpp:168-171) on the device-resident run loop, the full-row size
limit of the C++ Colony, and create-time limits of the batched solver."""
import ctypes
import re
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from acogpu_amd import capi, instances

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "acogpu_amd", "_build")
TSPLIB = os.path.join(ROOT, "tests", "golden", "tsplib")


def say(msg):
    print(msg, flush=True)


class DeviceBuffer:
    """hipMalloc'd copy of a host array (plain HIP runtime calls)."""

    def __init__(self, host):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        host = np.ascontiguousarray(host)
        assert self.L.hipMalloc(ctypes.byref(self.ptr), host.nbytes) == 0
        assert self.L.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), host.nbytes, 1) == 0  # H2D

    def close(self):
        if self.ptr:
            self.L.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


@pytest.mark.parametrize("mode", [capi.MODE_CANDIDATE, capi.MODE_FULLROW])
def test_device_pointer_distances_equal_host_pointer(mode):
    """aco_create with the matrix in device memory (the reference's GPU
    build) gives the same colony as the host pointer, bit for bit."""
    n = 150
    say(f"case: device-pointer distance matrix n={n} mode={mode}")
    xy = instances.uniform(n, 1000, seed=81)
    dist = orc.dist_from_int_coords(xy)
    dist[np.arange(n), np.arange(n)] = 12345  # the diagonal is ignored either way
    p = capi.default_params(n=n, m=40, k=12, beta=2.0, rho=0.3, mode=mode, variant=capi.VARIANT_AS, seed=4)
    buf = DeviceBuffer(dist)
    try:
        with capi.Colony(p, dist=dist) as host, capi.Colony(p, dist_device_ptr=buf.ptr.value) as dev:
            np.testing.assert_array_equal(host.dist(), dev.dist())
            host.initialize()
            dev.initialize()
            host.run(3)
            dev.run(3)
            np.testing.assert_array_equal(host.tours(), dev.tours())
            np.testing.assert_array_equal(host.lengths(), dev.lengths())
            np.testing.assert_array_equal(host.pheromone(), dev.pheromone())  # identical tours: same sums
            assert host.stats() == dev.stats()
    finally:
        buf.close()


def _iteration_rows(path):
    rows = []
    for line in open(path).read().splitlines():
        parts = line.split()
        if len(parts) == 12 and parts[0].isdigit():
            rows.append(parts[:4])  # iteration, ants, iteration best, best so far (no times)
    return rows


def test_acouav_device_and_host_distances_agree(tmp_path):
    """acouav hands the C++ AntSystem a device copy of the matrix by default,
    as the reference's GPU build; --host-distances passes the host matrix.
    Both runs print the same iterations."""
    rows = {}
    for flag in ([], ["--host-distances"]):
        out = tmp_path / f"run{len(flag)}.txt"
        args = [os.path.join(BUILD, "acouav_gpu_aging_nobroker"), "-m", "100", "-s", "25", "-i",
                os.path.join(TSPLIB, "kroA100.tsp"), "-p", str(out), "-b", "3", "-r", "0.1", "-j", "3", "-k", "5",
                "--candidates", "20", "--rank-based", "--w", "1", *flag]  # one deposit: deterministic
        say("case: " + " ".join(args[1:]))
        r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        rows[len(flag)] = _iteration_rows(out)
    assert len(rows[0]) == 25 and rows[0] == rows[1]


def test_colony_full_row_limit_warns_and_runs_on_candidate_lists(tmp_path):
    """The C++ Colony's default candidate set (every city in NN order) has a
    size limit on the device; beyond it the colony says so on stderr and
    runs on candidate lists."""
    n = 8200
    xy = instances.uniform(n, 100000, seed=5).reshape(-1, 2)
    tsp = tmp_path / "u8200.tsp"
    with open(tsp, "w") as f:
        f.write(f"NAME: u8200\nTYPE: TSP\nDIMENSION: {n}\nEDGE_WEIGHT_TYPE: EUC_2D\nNODE_COORD_SECTION\n")
        for i, (x, y) in enumerate(xy):
            f.write(f"{i + 1} {x} {y}\n")
        f.write("EOF\n")
    out = tmp_path / "u8200.txt"
    say(f"case: acouav full-row default at n={n}, 2 iterations")
    r = subprocess.run([os.path.join(BUILD, "acouav_gpu_noaging_nobroker"), "-m", "64", "-s", "2", "-i", str(tsp),
                        "-p", str(out), "-b", "2", "-r", "0.2"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"running with candidate lists of the \d+ nearest cities", r.stderr), r.stderr
    assert len(_iteration_rows(out)) == 2


def test_async_phases_equal_iterate_with_one_wait():
    """aco_phase without a wait and aco_phase_end give the iteration of
    aco_iterate: same tours, pheromone and statistics; the timers are the
    phases' event times."""
    say("case: asynchronous phases + aco_phase_end vs aco_iterate, RBAS w=1 aging n=90")
    xy = instances.uniform(90, 1000, seed=33)
    p = capi.default_params(n=90, m=24, k=10, w=1, beta=2.0, rho=0.3, variant=capi.VARIANT_RBAS, aging=1,
                            min_age=2, max_age=3, seed=8)
    with capi.Colony(p, xy=xy) as a, capi.Colony(p, xy=xy) as b:
        a.initialize()
        b.initialize()
        for it in range(8):
            a.iterate()
            for ph in (0, 1, 2, 3, 4):
                b.phase(ph, wait=False)
            t, ib, gb, nv = b.phase_end()
            assert (ib, gb, nv) == a.stats() == b.stats()
            # StatsTimer slots: placement (inside the tour kernel) reads 0
            assert t[0] == 0 and t[1] > 0 and t[2] > 0 and t[3] >= 0 and t[4] > 0 and t[5] > 0
            np.testing.assert_array_equal(a.tours(), b.tours())
            np.testing.assert_array_equal(a.pheromone(), b.pheromone())
        with pytest.raises(capi.AcoError, match="aco_phase_end"):
            b.phase_end()


@pytest.mark.parametrize("w", [1, 6])
def test_iterate_stats_equals_phases(w):
    """aco_iterate_stats (the iteration's kernels launched back to back with
    their own dispatch timestamps, statistics written to host memory by the
    statistics kernel, one wait; rank-based: statistics and deposit as one
    launch) gives the iteration of the phase-by-phase path (two launches),
    with its statistics, over both pheromone buffers; the timers of the
    phases that run kernels are positive."""
    say(f"case: aco_iterate_stats vs waiting phases, RBAS w={w} aging n=200")
    xy = instances.uniform(200, 1000, seed=37)
    p = capi.default_params(n=200, m=64, k=16, w=w, beta=2.0, rho=0.2, variant=capi.VARIANT_RBAS, aging=1,
                            min_age=2, max_age=4, global_perf=2, seed=9)
    with capi.Colony(p, xy=xy) as a, capi.Colony(p, xy=xy) as b:
        a.initialize()
        b.initialize()
        for it in range(9):
            t, ib, gb, nv = a.iterate_stats()
            for ph in (0, 1, 2, 3, 4):
                b.phase(ph)
            assert (ib, gb, nv) == a.stats() == b.stats(), it
            assert t[0] == 0 and t[3] == 0 and t[1] > 0 and t[2] > 0 and t[4] > 0 and t[5] > 0
            np.testing.assert_array_equal(a.tours(), b.tours())
            np.testing.assert_array_equal(a.pheromone(), b.pheromone())
        assert a.aging()["num_valid"] < 64  # retirements happened and were reported


def test_run_until_time_limit():
    """The time test of TSP::terminationCondition on the device-resident loop:
    a zero limit runs nothing, a positive limit stops at the first chunk
    boundary at or past it, and the colony equals one run for exactly the
    iterations reported (rank-based w = 1: deterministic)."""
    n = 1000
    say(f"case: run_until_time n={n} m=256 RBAS w=1")
    xy = instances.clustered(n, clusters=10, side=10000, sigma=400.0, seed=1)
    p = capi.default_params(n=n, m=256, k=20, w=1, beta=3.0, rho=0.1, variant=capi.VARIANT_RBAS, seed=3,
                            tau0=0.001)
    with capi.Colony(p, xy=xy) as a, capi.Colony(p, xy=xy) as b:
        a.initialize()
        b.initialize()
        assert a.run_until_time(100, -1, 0.0, chunk=8)[:2] == (0, 2)
        ran, halted, ms = a.run_until_time(10000, -1, 3.0, chunk=8)
        say(f"ran {ran} iterations, halted {halted}, {ms:.3f} ms")
        assert halted == 2 and ran % 8 == 0 and ran >= 8 and ms >= 3.0
        ran2, halted2, _ = a.run_until_time(16, -1, 1e6, chunk=8)
        assert (ran2, halted2) == (16, 0)
        b.run(ran + ran2)
        np.testing.assert_array_equal(a.tours(), b.tours())
        np.testing.assert_array_equal(a.pheromone(), b.pheromone())
        assert a.stats() == b.stats()


def test_batch_create_refuses_instances_beyond_lds():
    """A batch whose largest instance's candidate table cannot be staged in
    LDS is refused at create with the limit, not at the first run."""
    say("case: batch create n=3000 k=16 (table beyond LDS)")
    sizes = np.array([100, 3000], dtype=np.int32)
    xys = [instances.uniform(int(n), 1000, seed=i) for i, n in enumerate(sizes)]
    p = capi.default_params(n=0, m=0, k=16, beta=2.0, rho=0.5)
    with pytest.raises(capi.AcoError, match="bytes of LDS"):
        capi.Batch(p, sizes, xys)
