"""Shared helpers for the GPU tests: one context per session, scene builders."""
import math
import os

import numpy as np
import pytest

from diorama_amd import scene

_CTX = None


def context():
    global _CTX
    if _CTX is None:
        import diorama_amd

        _CTX = diorama_amd.Context(0)
    return _CTX


def say(*parts):
    """Names the case before a GPU step, flushed, so a crash still shows it."""
    print("[gpu]", *parts, flush=True)


def random_scene(count, seed, scale=(0.02, 0.4), opacity=(0.2, 0.95)):
    """src/checks/tiles.ts:169-186: Gaussians scattered through a 90-degree view."""
    random = scene.mulberry32(seed)
    between = lambda lo, hi: lo + (hi - lo) * random()
    out = []
    for _ in range(count):
        depth = between(2, 8)
        position = [between(-1.2, 1.2) * depth, between(-0.9, 0.9) * depth, -depth]
        s = [between(*scale), between(*scale), between(*scale)]
        q = [between(-1, 1), between(-1, 1), between(-1, 1), between(-1, 1)]
        color = [random(), random(), random()]
        out.append(position + [between(*opacity)] + s + [0] + q + color + [0])
    return np.array(out, dtype=np.float32)


def config0_scene(seed=0, n=16384):
    """16,384 Gaussians in [-1,1]^3 seen by the 256x256, 60-degree camera at distance 4."""
    params, sh = scene.random_raw_gaussians(n, seed, box=1.0)
    return params, sh, scene.activate(params), scene.small_config_camera()
