"""The Morton-ordered, box-culled 3-nearest-neighbour search and the starting
Gaussians on the GPU, bit for bit against the brute-force float32 oracle
(oracle/neighbors.c): ragged sizes around the 64-point chunks, ties, duplicates,
clusters at very different scales, non-finite points, and a million points
checked on a sample of queries."""
import time

import numpy as np
import pytest

import diorama_amd
from oracle import oracle as oc
from tests.gpu_util import context, say

pytestmark = pytest.mark.gpu


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, 1, (n // 2, 3)), rng.uniform(-5, 5, (n // 4, 3)),
             rng.normal(3, 0.001, (n - n // 2 - n // 4, 3))]
    return np.concatenate(parts).astype(np.float32)[rng.permutation(n)]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 63, 64, 65, 4097, 100003])
def test_matches_brute_force(n):
    ctx = context()
    pts = cloud(n, n)
    say(f"neighbors n={n}")
    got = diorama_amd.mean_neighbor_distance2(ctx, pts)
    want = oc.mean_neighbor_distance2(pts)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_ties_duplicates_and_non_finite_points():
    ctx = context()
    g = np.arange(12, dtype=np.float32)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)   # equal distances everywhere
    rng = np.random.default_rng(9)
    dup = np.repeat(rng.normal(0, 1, (50, 3)), 3, axis=0)                    # each point three times
    tiny = rng.normal(100, 1e-6, (200, 3))                                   # a cluster far away, tightly packed
    bad = np.array([[np.nan, 0, 0], [0, np.inf, 0], [-np.inf, 1, 2], [1e38, 1e38, 1e38]])
    pts = np.concatenate([grid, dup, tiny, bad]).astype(np.float32)
    pts = pts[rng.permutation(len(pts))]
    say("neighbors ties/duplicates/non-finite")
    got = diorama_amd.mean_neighbor_distance2(ctx, pts)
    want = oc.mean_neighbor_distance2(pts)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isinf(got[~np.isfinite(pts).all(axis=1)]).all()


def test_a_million_points_on_a_sample():
    ctx = context()
    n = 1_000_000
    pts = cloud(n, 77)
    say("neighbors n=1M")
    t0 = time.perf_counter()
    got = diorama_amd.mean_neighbor_distance2(ctx, pts)
    ms = (time.perf_counter() - t0) * 1e3
    sample = np.random.default_rng(1).choice(n, 3000, replace=False)
    want = oc.mean_neighbor_distance2(pts, sample)
    print(f"neighbors: 1M points in {ms:.1f} ms including transfers")
    assert np.array_equal(got[sample].view(np.uint32), want.view(np.uint32))
    assert np.isfinite(got).all() and (got >= 0).all()


def test_starting_gaussians():
    ctx = context()
    n = 20000
    pts = cloud(n, 5)
    colors = np.random.default_rng(6).integers(0, 256, (n, 3)).astype(np.uint8)
    say("initial gaussians")
    raw = diorama_amd.initial_gaussians(ctx, pts, colors)
    params, sh = oc.initial_gaussians(pts, colors, oc.mean_neighbor_distance2(pts))
    d = np.abs(raw["params"].view(np.int32).astype(np.int64) - params.view(np.int32).astype(np.int64))
    assert d[:, [0, 1, 2, 3, 7, 8, 9, 10, 11, 12, 13, 14, 15]].max() == 0
    assert d[:, 4:7].max() <= 1          # log in double on each side, rounded once
    assert not raw["sh"].any() and raw["sh"].shape == (n, 45)


def test_colmap_reconstruction_to_starting_gaussians():
    """A hand-assembled COLMAP reconstruction read through the C ABI, its points turned
    into starting Gaussians on the GPU (src/loadDataset.ts:87-149), against the oracle's
    reader (oracle/colmap.py), brute-force 3-NN and initial_gaussians."""
    from oracle import colmap
    from tests.test_colmap import scene_files

    cb, ib, pb, _, _ = scene_files(seed=5, photos=6, points=2000)
    got = diorama_amd.parse_colmap(cb, ib, pb)
    want = colmap.parse(cb, ib, pb)
    positions = np.array(want["positions"], np.float32)
    colors = np.array(want["colors"], np.uint8)
    assert np.array_equal(got["positions"], positions) and np.array_equal(got["colors"], colors)
    ctx = context()
    say("colmap to gaussians", len(positions))
    raw = diorama_amd.initial_gaussians(ctx, got["positions"], got["colors"])
    d2 = oc.mean_neighbor_distance2(positions)
    o_params, o_sh = oc.initial_gaussians(positions, colors, d2)
    assert np.array_equal(oc.mean_neighbor_distance2(got["positions"]), d2)
    d = np.abs(raw["params"].view(np.int32).astype(np.int64) - o_params.view(np.int32).astype(np.int64))
    assert d[:, [0, 1, 2, 3, 7, 8, 9, 10, 11, 12, 13, 14, 15]].max() == 0
    assert d[:, 4:7].max() <= 1          # log in double on each side, rounded once
    assert not raw["sh"].any() and raw["sh"].shape == o_sh.shape
