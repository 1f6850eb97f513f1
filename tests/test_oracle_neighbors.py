"""Pins the neighbour oracle (oracle/neighbors.c) and the starting Gaussians
(oracle.initial_gaussians, src/loadDataset.ts:87-149) before the GPU search is
checked against them: hand-worked lines, the duplicate and few-point cases of the
reference's loop (nearestNeighbors.wgsl:12-44), and a k-d tree in float64."""
import math

import numpy as np
from scipy.spatial import cKDTree

from oracle import oracle as oc

f = np.float32


def test_points_on_a_line_by_hand():
    pts = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0], [7, 0, 0]], np.float32)
    got = oc.mean_neighbor_distance2(pts)
    # squared distances: 0 -> 1, 9, 49; 1 -> 1, 4, 36; 3 -> 4, 9, 16; 7 -> 16, 36, 49
    assert list(got) == [f(59) / f(3), f(41) / f(3), f(29) / f(3), f(101) / f(3)]


def test_duplicates_count_and_the_point_itself_does_not():
    pts = np.array([[1, 2, 3], [1, 2, 3], [1, 2, 5], [4, 2, 3], [9, 9, 9]], np.float32)
    got = oc.mean_neighbor_distance2(pts)
    assert got[0] == f(0 + 4 + 9) / f(3) and got[1] == got[0]


def test_fewer_than_four_points():
    # The reference starts from 3.4e38: with 2 others the third stays 3.4e38, with fewer the sum overflows.
    three = oc.mean_neighbor_distance2(np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]], np.float32))
    assert three[0] == (f(1) + f(4) + f(3.4e38)) / f(3)
    assert np.isinf(oc.mean_neighbor_distance2(np.array([[0, 0, 0], [1, 0, 0]], np.float32))).all()
    assert np.isinf(oc.mean_neighbor_distance2(np.zeros((1, 3), np.float32))).all()


def test_against_a_kd_tree():
    rng = np.random.default_rng(5)
    pts = np.concatenate([rng.normal(0, 1, (3000, 3)), rng.normal(4, 0.01, (1000, 3))]).astype(np.float32)
    got = oc.mean_neighbor_distance2(pts)
    dist, _ = cKDTree(pts.astype(np.float64)).query(pts.astype(np.float64), k=4)
    want = (dist[:, 1:] ** 2).mean(axis=1)
    assert np.abs(got / want - 1).max() < 1e-4
    sample = rng.choice(len(pts), 50, replace=False)
    assert np.array_equal(oc.mean_neighbor_distance2(pts, sample), got[sample])


def test_starting_gaussians_by_hand():
    pts = np.array([[1, 2, 3], [0, 0, 0]], np.float32)
    colors = np.array([[255, 0, 128], [0, 0, 0]], np.uint8)
    params, sh = oc.initial_gaussians(pts, colors, np.array([4.0, 0.0], np.float32))
    c0 = 0.28209479177387814
    assert list(params[0, :4]) == [1, 2, 3, f(math.log(0.1 / 0.9))]
    assert list(params[0, 4:7]) == [f(math.log(2))] * 3
    assert list(params[1, 4:7]) == [f(0.5 * math.log(1e-7))] * 3
    assert list(params[0, 8:12]) == [1, 0, 0, 0] and params[0, 7] == 0 and params[0, 15] == 0
    assert list(params[0, 12:15]) == [f(0.5 / c0), f(-0.5 / c0), f((128 / 255 - 0.5) / c0)]
    assert not sh.any() and sh.shape == (2, 45)
