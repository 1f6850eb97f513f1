"""TrainingLoss on the GPU: the reference's loss fixture (src/checks/loss.ts:24-89)
and the oracle at the training resolution. Bars: loss within 1e-5, dL/dpixel
within 1e-3 of its largest value."""
import json
import os

import numpy as np
import pytest

import diorama_amd
from oracle import oracle
from tests.gpu_util import context, say

pytestmark = pytest.mark.gpu


def pack_rgba8(rgb):
    rgb = np.asarray(rgb, dtype=np.uint32)
    return rgb[:, 0] | (rgb[:, 1] << 8) | (rgb[:, 2] << 16) | (255 << 24)


def gpu_loss(width, height, rendered, target_rgb):
    ctx = context()
    loss = diorama_amd.TrainingLoss(ctx, width, height)
    r = ctx.upload(np.ascontiguousarray(rendered, dtype=np.float32))
    t = ctx.upload(pack_rgba8(target_rgb))
    loss.encode(r, t)
    value = loss.read()
    grads = loss.pixel_grads.download(np.float32, width * height * 4).reshape(-1, 4)
    shares = loss.pixel_loss.download(np.float32, width * height)
    loss.destroy()
    return value, grads, shares


def test_loss_fixture():
    with open(os.path.join(os.path.dirname(__file__), "golden", "loss.json")) as f:
        data = json.load(f)
    w, h = data["size"]
    rendered = np.zeros((w * h, 4), dtype=np.float32)
    rendered[:, :3] = np.array(data["render"]).reshape(-1, 3)
    target = np.array(data["target"]).reshape(-1, 3)
    say("loss fixture")
    value, grads, _ = gpu_loss(w, h, rendered, target)
    ref = np.array(data["gradient"]).reshape(-1, 3)
    err = float(np.abs(grads[:, :3] - ref).max() / np.abs(ref).max())
    print(f"loss fixture: |loss - ref| = {abs(value - data['loss']):.2e}, gradient err / max {err:.2e}")
    assert abs(value - data["loss"]) < 1e-5
    assert err < 1e-3


@pytest.mark.parametrize("size", [(980, 545), (256, 256), (17, 5)])
def test_loss_against_oracle(size):
    w, h = size
    rng = np.random.default_rng(w)
    target = rng.integers(0, 256, size=(w * h, 3))
    rendered = np.zeros((w * h, 4), dtype=np.float32)
    rendered[:, :3] = np.clip(target / 255 + rng.normal(0, 0.1, size=(w * h, 3)), -0.05, 1.05)
    say("loss", w, h)
    value, grads, shares = gpu_loss(w, h, rendered, target)
    t8 = np.zeros((w * h, 4), dtype=np.uint8)
    t8[:, :3] = target
    o_value, o_grads, o_shares = oracle.loss(w, h, rendered.astype(np.float64), t8)
    err = float(np.abs(grads[:, :3] - o_grads[:, :3]).max() / np.abs(o_grads[:, :3]).max())
    print(f"loss {w}x{h}: |loss - oracle| = {abs(value - o_value):.2e}, gradient err / max {err:.2e}")
    assert abs(value - o_value) < 1e-5
    assert err < 1e-3
    assert np.abs(shares - o_shares).max() < 1e-6
