"""Pins the oracle's neighbor search (NeighborhoodSearchSparse, SPH/main.py:764-853):
hand-derived known answers, lattice counts, brute force with the reference's FP32
test and serial list order, symmetry, and the caps."""
import json
import os

import numpy as np
import pytest

from nonnewton_amd.params import KIND_FLUID, KIND_STATIC, SceneParams, params_from_config
from oracle import oracle as O

from conftest import GOLDEN

# test_neighbor.ply: 3 x 3 points at 0.03 spacing, h = 0.04; axis pairs (0.03) are
# neighbors, diagonals (0.0424) are not -- derived by hand.
KAT = {0: [1, 3], 1: [0, 2, 4], 2: [1, 5], 3: [0, 4, 6], 4: [1, 3, 5, 7], 5: [2, 4, 8], 6: [3, 7], 7: [4, 6, 8], 8: [5, 7]}


def search(points, params, kind=None):
    pts = np.asarray(points, dtype=np.float32)
    kind = np.zeros(len(pts), np.int32) if kind is None else kind
    s = O.OracleSim(params, pts, kind)
    st = s.search()
    nbr, cnt = s.neighbors()
    return nbr, cnt, st


def brute_force_serial(pts, h, gn):
    """Every pair with the FP32 norm test, ordered as the reference's serial run
    lists them: 27 cell offsets in ti.ndrange order, then particle index."""
    pts = np.asarray(pts, dtype=np.float32)
    n = len(pts)
    hf = np.float32(h)
    cell = (pts / hf).astype(np.int32)
    out = []
    for i in range(n):
        d = pts[i] - pts
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        js = [j for j in np.nonzero(dist < hf)[0] if j != i]
        off = cell[js] - cell[i]
        order = sorted(range(len(js)), key=lambda k: ((off[k][0] + 1) * 9 + (off[k][1] + 1) * 3 + (off[k][2] + 1), js[k]))
        out.append([int(js[k]) for k in order])
    return out


def test_neighbor_ply_known_answer():
    pts = np.load(os.path.join(GOLDEN, "test_neighbor_points.npy"))
    with open(os.path.join(GOLDEN, "scene_test_neighbor.json")) as f:
        p, _ = params_from_config(json.load(f))
    nbr, cnt, st = search(pts, p)
    assert st == 0
    assert int(cnt.sum()) == 24
    for i, exp in KAT.items():
        assert sorted(nbr[i, : cnt[i]].tolist()) == exp
        assert np.all(nbr[i, cnt[i]:] == -1)


def test_lattice_interior_has_26():
    p = SceneParams(domain_end=(0.4, 0.4, 0.4))
    g = np.arange(7) * 0.02 + 0.1
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    nbr, cnt, st = search(pts, p)
    center = 3 * 49 + 3 * 7 + 3
    assert cnt[center] == 26  # the 6 lattice points at exactly h are excluded (strict <)


def test_bruteforce_sets_and_serial_order():
    rng = np.random.default_rng(7)
    p = SceneParams(domain_end=(0.5, 0.5, 0.5))
    g = np.arange(10) * 0.02 + 0.05
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts + rng.uniform(-0.002, 0.002, pts.shape)
    pts = rng.permutation(pts).astype(np.float32)
    nbr, cnt, st = search(pts, p)
    ref = brute_force_serial(pts, p.support_radius, p.grid_num)
    for i in range(len(pts)):
        assert nbr[i, : cnt[i]].tolist() == ref[i]


def test_symmetric():
    rng = np.random.default_rng(3)
    p = SceneParams(domain_end=(0.5, 0.5, 0.5))
    pts = rng.uniform(0.05, 0.3, (800, 3)).astype(np.float32)
    nbr, cnt, _ = search(pts, p)
    pairs = {(i, int(j)) for i in range(len(pts)) for j in nbr[i, : cnt[i]]}
    assert all((j, i) in pairs for i, j in pairs)


def test_cell_overflow_reported():
    p = SceneParams(domain_end=(0.2, 0.2, 0.2))
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.041, 0.079, (51, 3))  # 51 particles in cell (1, 1, 1)
    _, _, st = search(pts, p)
    assert st & O.ST_CELL_OVERFLOW


def test_neighbor_overflow_reported():
    p = SceneParams(domain_end=(0.4, 0.4, 0.4))
    g = np.arange(5) * 0.01 + 0.1  # 125 points at spacing r: ~100 neighbors each
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    _, cnt, st = search(pts, p)
    assert st & O.ST_NEIGHBOR_OVERFLOW
    assert cnt.max() == 60


def test_out_of_grid_reported_and_truncation_kept():
    p = SceneParams(domain_end=(0.2, 0.2, 0.2))
    _, _, st = search([[-0.05, 0.1, 0.1], [0.1, 0.1, 0.1]], p)
    assert st & O.ST_OUT_OF_GRID
    # x in (-h, 0) truncates to cell 0 (pos_to_index casts toward zero, :807)
    _, _, st = search([[-0.01, 0.1, 0.1], [0.1, 0.1, 0.1]], p)
    assert st == 0


def test_squared_threshold_equivalent_to_sqrt_test():
    """sqrt_rn(s) < h  <=>  s < T, T the least float with sqrt_rn(T) >= h (the GPU's
    test), checked exhaustively over 2^21 floats around h^2 for h = 0.04."""
    h = np.float32(0.04)
    t = np.float32(h * h)
    while np.sqrt(t) >= h:
        t = np.nextafter(t, np.float32(0))
    while np.sqrt(t) < h:
        t = np.nextafter(t, np.float32(np.inf))
    base = t.view(np.int32)
    s = (np.arange(-(1 << 20), 1 << 20, dtype=np.int64) + int(base)).astype(np.int32).view(np.float32)
    assert np.array_equal(s < t, np.sqrt(s) < h)
