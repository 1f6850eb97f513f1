"""Pins the MLP oracle (oracle/mlp.py) before the GPU kernels are checked against it:
central finite differences of its loss, the BF16 rounding rule on hand-picked bit
patterns, closed forms (zero weights give log 4), and hand-computed inputs."""
import math

import numpy as np

from oracle import mlp


def test_parameter_count():
    assert mlp.PARAMS == 5252


def test_gradients_match_central_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, size=mlp.PARAMS)
    x = rng.normal(0, 1, size=(24, mlp.INPUTS))
    labels = rng.integers(0, 4, size=24)
    cw = [0.7, 1.3, 2.0, 0.4]
    _, g = mlp.loss_and_grads(w, x, labels, cw)
    picks = np.concatenate([rng.choice(mlp.PARAMS, 60, replace=False), np.arange(mlp.PARAMS - 4, mlp.PARAMS)])
    h = 1e-6
    worst = 0.0
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mlp.loss_and_grads(wp, x, labels, cw)[0] - mlp.loss_and_grads(wm, x, labels, cw)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[k]))
    assert worst < 1e-7 * max(1.0, np.abs(g).max()), worst


def test_zero_weights_give_uniform_loss():
    x = np.random.default_rng(1).normal(size=(5, mlp.INPUTS))
    loss, g = mlp.loss_and_grads(np.zeros(mlp.PARAMS), x, [0, 1, 2, 3, 0])
    assert abs(loss - math.log(4)) < 1e-12
    # only the output biases see a gradient: mean(p - onehot) = 1/4 - share of each class
    assert np.allclose(g[-4:], [0.25 - 0.4, 0.25 - 0.2, 0.25 - 0.2, 0.25 - 0.2])
    assert not g[:-4].any()


def test_bf16_rounds_to_nearest_even():
    bits = np.array([0x3F800000, 0x3F808000, 0x3F818000, 0x3F808001, 0xBF7FFFFF, 0x00000000], dtype=np.uint32)
    want = np.array([0x3F800000, 0x3F800000, 0x3F820000, 0x3F810000, 0xBF800000, 0x00000000], dtype=np.uint32)
    got = mlp.bf16(bits.view(np.float32)).view(np.uint32)
    assert np.array_equal(got, want)


def test_features_by_hand():
    p = np.zeros((1, 16))
    p[0, 3] = 0.0                          # opacity 0.5
    p[0, 4:7] = np.log([0.02, 0.05, 0.01])
    p[0, 12:15] = [1.0, 0.0, -1.0]
    p[0, 15] = 300.0                       # born at step 300
    stats = np.array([[6e-4, 3.0, 4.5, 0.0]])
    extra = np.array([[4e-4, 12.0, 3e-4, -4e-4]])
    f = mlp.features(p, stats, extra, extent=2.0, iteration=1000)[0]
    c = 0.5 + mlp.SH_C0 * np.array([1.0, 0.0, -1.0])
    want = [math.log(1 + 2e-4 / 2e-4), math.log(1 + 2.0), 0.5, math.log(0.05 / 2), math.log(0.01 / 2), math.log(5.0),
            0.299 * c[0] + 0.587 * c[1] + 0.114 * c[2], math.log(4.0), math.log(701.0), math.log(5.5),
            math.log(4.0), 1 - 5e-4 / 6e-4]
    assert np.allclose(f, want, rtol=1e-12, atol=1e-12)
