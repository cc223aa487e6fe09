"""The oracle pins catch plausible mistakes: each case below builds a copy of
the oracle (on the CPU, into a temporary directory) with one line changed and
requires the named checks of tests/oracle_pins.py to fail on it, while the
oracle as built passes them (tests/test_oracle.py)."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle as orc
import oracle_pins as pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
CFLAGS = ["-O3", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-w"]

# (name, original text in aco_oracle.c, planted mistake, checks that must fail)
MUTANTS = [
    ("choice from the evaporated pheromone",
     "if (choice) choice[(int64_t)i * n + j] = apply_alpha(t, alpha) * e;",
     "if (choice) choice[(int64_t)i * n + j] = apply_alpha(one_minus_rho * t, alpha) * e;",
     ["sweep_closed_forms"]),
    ("evaporation without the 1e-18 floor",
     "tau_evap[(int64_t)i * n + j] = ((double)p > 1e-18) ? p : (float)1e-18;",
     "tau_evap[(int64_t)i * n + j] = p;",
     ["sweep_closed_forms"]),
    ("candidate weights read from the transposed choice matrix",
     ": choice[(int64_t)i * n + nn[e]];",
     ": choice[(int64_t)nn[e] * n + i];",
     ["sweep_closed_forms"]),
    ("young factor 1 instead of age / min_age",
     "wc.ratio = wc.young ? (float)age[a] / (float)min_age : 0.0f;",
     "wc.ratio = wc.young ? 1.0f : 0.0f;",
     ["young_candidate", "young_fullrow", "young_refseq"]),
    ("young weight without the pull toward the row mean",
     "float p = ((w->tau[xy] - m) * w->ratio) + m;",
     "float p = (w->tau[xy] * w->ratio) + m;",
     ["young_candidate", "young_fullrow", "young_refseq"]),
    ("young weight without alpha",
     "return apply_alpha(p, w->alpha) * orc_eta(w->eta, w->eta_len, w->dist[xy]);",
     "return p * orc_eta(w->eta, w->eta_len, w->dist[xy]);",
     ["young_candidate_alpha"]),
    ("young weight with the candidate-edge heuristic ignores age",
     "float p = ((tau[xy] - mc) * wc.ratio) + mc;",
     "float p = tau[xy];",
     ["young_candidate_heur"]),
    ("nearest instead of farthest unvisited city",
     "for (int32_t r = n - 1; r >= 0; r--)",
     "for (int32_t r = 0; r < n; r++)",
     ["zero_total_fullrow", "zero_total_refseq", "overrun_fullrow", "overrun_refseq"]),
    ("overrun takes the row's farthest city, visited or not",
     "next = (i < 0) ? farthest_unvisited(row, vis, n) : row[i];",
     "next = (i < 0) ? row[n - 2] : row[i];",
     ["overrun_refseq"]),
    ("full-row prefix without the earlier chunks' total",
     "float P = tprev[q] + prow[r];",
     "float P = prow[r];",
     ["first_move_fullrow"]),
]


def _build(tmp, idx, old, new):
    src = open(os.path.join(ORACLE_DIR, "aco_oracle.c")).read()
    assert src.count(old) == 1, f"mutation anchor not unique or missing: {old!r}"
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d)
    with open(os.path.join(d, "aco_oracle.c"), "w") as f:
        f.write(src.replace(old, new))
    for name in ("prior_oracle.c", "aco_oracle.h"):
        shutil.copy(os.path.join(ORACLE_DIR, name), d)
    out = os.path.join(d, "liboracle_mutant.so")
    subprocess.run(["gcc", *CFLAGS, "-shared", "-o", out, os.path.join(d, "aco_oracle.c"),
                    os.path.join(d, "prior_oracle.c"), "-lm"], check=True)
    return out


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(8) as ex:
        futs = [ex.submit(_build, tmp, i, old, new) for i, (_, old, new, _) in enumerate(MUTANTS)]
        return [f.result() for f in futs]


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pins_fail_on_planted_mistake(mutant_libs, idx):
    name, _, _, must_fail = MUTANTS[idx]
    with orc.using(mutant_libs[idx]):
        for check in must_fail:
            with pytest.raises(AssertionError):
                pins.CHECKS[check]()
