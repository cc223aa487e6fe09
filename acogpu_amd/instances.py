"""Seeded synthetic TSP instances for the benchmark configurations.

Integer coordinates, interleaved (x0, y0, x1, y1, ...), generated on the host
with numpy's PCG64 generator.  The same arrays feed the GPU library and the
CPU oracle, so both sides see identical inputs.
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, side: int, seed: int) -> np.ndarray:
    """n cities, integer coordinates uniform in [0, side]^2."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, side + 1, size=(n, 2), dtype=np.int64).astype(np.int32)


def clustered(n: int, clusters: int = 10, side: int = 10000, sigma: float = 400.0,
              seed: int = 0) -> np.ndarray:
    """n cities around `clusters` Gaussian centres uniform in [0, side]^2,
    per-axis N(0, sigma^2), clipped to [0, side] and rounded to integers."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, side, size=(clusters, 2))
    which = rng.integers(0, clusters, size=n)
    pts = centres[which] + rng.normal(0.0, sigma, size=(n, 2))
    return np.rint(np.clip(pts, 0.0, side)).astype(np.int32)


def batched_sizes(count: int, lo: int = 50, hi: int = 500, seed: int = 0) -> np.ndarray:
    """Instance sizes log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int32)


def tour_length(dist: np.ndarray, tour: np.ndarray) -> int:
    t = np.asarray(tour, dtype=np.int64)
    return int(dist[t, np.roll(t, -1)].astype(np.int64).sum())


def read_tsplib(path: str):
    """TSPLIB EUC_2D reader (TSPReader::read, src/TSPReader.cu:25-100):
    returns (name, x float32[n], y float32[n]).  Raises ValueError on a
    non-TSP type or a non-EUC_2D edge weight type, like the reference."""
    name, n, xs, ys = "", 0, [], []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].rstrip()
        i += 1
        if len(line) <= 1:
            continue
        if ":" in line:
            tag, value = line.split(":", 1)
        else:
            tag, value = line, ""
        tag, value = tag.replace(" ", ""), value.replace(" ", "")
        if tag == "NAME":
            name = value
        elif tag == "TYPE" and value not in ("TSP", "STSP"):
            raise ValueError(f"Invalid problem type: {value}")
        elif tag == "DIMENSION":
            n = int(value)
        elif tag == "EDGE_WEIGHT_TYPE" and value != "EUC_2D":
            raise ValueError(f"Invalid edge weight type: {value}")
        elif tag == "NODE_COORD_SECTION":
            while len(xs) < n and i < len(lines):
                parts = lines[i].split()
                i += 1
                if not parts or parts[0] == "EOF":
                    break
                xs.append(float(parts[1]))
                ys.append(float(parts[2]))
        elif tag == "EOF":
            break
    if len(xs) != n:
        raise ValueError(f"expected {n} coordinates, read {len(xs)}")
    return name, np.asarray(xs, dtype=np.float32), np.asarray(ys, dtype=np.float32)


def euc2d_matrix_from_float(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Distance matrix with the reference's EUC_2D arithmetic
    (Utils::round_distance src/Utils.cu:212-218: fp32 coordinate difference,
    fp64 sqrt, +0.5, truncation), MAX_DISTANCE on the diagonal."""
    x = np.asarray(x, dtype=np.float32)
    y = np.asarray(y, dtype=np.float32)
    dx = (x[:, None] - x[None, :]).astype(np.float64)
    dy = (y[:, None] - y[None, :]).astype(np.float64)
    d = (np.sqrt(dx * dx + dy * dy) + 0.5).astype(np.int64).astype(np.int32)
    np.fill_diagonal(d, 0x7FFFFFFF)
    return d
