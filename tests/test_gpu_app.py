"""GPU tests of the drop-in surfaces: the acouav command line (C++ Colony /
AntSystem / RankBasedAntSystem over the C ABI) and the phase-by-phase ABI."""
import json
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
OPT = json.load(open(os.path.join(TSPLIB, "optima.json")))["optimum"]


def run_acouav(prog, args, cwd):
    print(f"case: {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BUILD, prog), *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def parse_iterations(path):
    lines = open(path).read().splitlines()
    hdr = [i for i, l in enumerate(lines) if l.startswith("Iter#")]
    assert hdr, "missing column header"
    rows = []
    for l in lines[hdr[-1] + 1:]:
        parts = l.split()
        if len(parts) == 12 and parts[0].isdigit():
            rows.append(parts)
    return lines, rows


def test_acouav_reference_flags_berlin52(tmp_path):
    os.makedirs(tmp_path / "aco" / "release" / "output")
    out = tmp_path / "berlin52.txt"
    r = run_acouav("acouav_gpu_noaging_nobroker",
                   ["-m", "52", "-s", "60", "-o", "7542", "-i", os.path.join(TSPLIB, "berlin52.tsp"), "-p", str(out),
                    "-a", "1.0", "-b", "2.0", "-r", "0.5"], tmp_path)
    lines, rows = parse_iterations(out)
    assert "PLATFORM: GPU" in lines and "ACO: Ant System" in lines and "Ants Aging: FALSE" in lines
    assert 1 <= len(rows) <= 60
    gbest = [int(r[3]) for r in rows]
    assert all(b >= OPT["berlin52"] for b in gbest)
    assert all(x >= y for x, y in zip(gbest, gbest[1:]))  # best-so-far never increases
    assert [int(r[0]) for r in rows] == list(range(len(rows)))
    # summaries in aco/release/output relative to the working directory
    log = (tmp_path / "aco" / "release" / "output" / "summary.log").read_text().splitlines()
    csv = (tmp_path / "aco" / "release" / "output" / "summary.csv").read_text().splitlines()
    assert log[0].split()[0] == "platform" and log[1].split()[0] == "GPU"
    assert csv[0].startswith("platform,cities,ants") and csv[1].startswith("GPU,52,52,")
    assert r.stdout.split()[0] == "GPU"


def test_acouav_tour_line_when_optimum_reached(tmp_path):
    out = tmp_path / "eil76.txt"
    # optimum set loose so the run stops early and prints the tour
    run_acouav("acouav_gpu_noaging_nobroker",
               ["-m", "76", "-s", "200", "-o", "600", "-i", os.path.join(TSPLIB, "eil76.tsp"), "-p", str(out),
                "-b", "3", "-r", "0.2", "--candidates", "20"], tmp_path)
    lines, rows = parse_iterations(out)
    tour_line = lines[-1]
    assert tour_line.endswith(",")
    tour = [int(v) - 1 for v in tour_line.rstrip(",").split(",")]
    assert sorted(tour) == list(range(76))
    name, x, y = instances.read_tsplib(os.path.join(TSPLIB, "eil76.tsp"))
    d = orc.dist_from_float_coords(x, y)
    assert instances.tour_length(d, tour) == int(rows[-1][3]) <= 600


def test_acouav_aging_variant_and_rank_based(tmp_path):
    out = tmp_path / "kroA100.txt"
    run_acouav("acouav_gpu_aging_nobroker",
               ["-m", "100", "-s", "40", "-i", os.path.join(TSPLIB, "kroA100.tsp"), "-p", str(out), "-b", "3",
                "-r", "0.1", "-j", "3", "-k", "5", "-x", "98", "--rank-based", "--w", "6", "--dorigo-tau0",
                "--candidates", "20"], tmp_path)
    lines, rows = parse_iterations(out)
    assert "Ants Aging: TRUE" in lines and "ACO: Rank-based Ant System" in lines
    assert len(rows) == 40
    ants = [int(r[1]) for r in rows]
    assert min(ants) < 100  # retirements happened
    assert int(rows[-1][3]) >= OPT["kroA100"]


def test_acouav_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.tsp"
    bad.write_text("NAME: bad\nTYPE: TSP\nDIMENSION: 3\nEDGE_WEIGHT_TYPE: GEO\n")
    r = subprocess.run([os.path.join(BUILD, "acouav"), "-i", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Invalid edge weight type" in r.stdout


def test_phase_api_matches_iterate():
    print("case: phase API vs aco_iterate, w=1 RBAS n=80", flush=True)
    xy = instances.uniform(80, 1000, seed=31)
    p = capi.default_params(n=80, m=16, k=10, w=1, beta=2.0, rho=0.3, variant=capi.VARIANT_RBAS, seed=8)
    with capi.Colony(p, xy=xy) as a, capi.Colony(p, xy=xy) as b:
        a.initialize()
        b.initialize()
        for _ in range(4):
            a.iterate()
            for ph in (0, 1, 2, 3, 4):
                b.phase(ph)
        np.testing.assert_array_equal(a.tours(), b.tours())
        np.testing.assert_array_equal(a.pheromone(), b.pheromone())
        assert a.stats() == b.stats()
        b.phase(0)
        with pytest.raises(capi.AcoError, match="out of order"):
            b.phase(0)
        with pytest.raises(capi.AcoError, match="phased iteration"):
            b.iterate()
