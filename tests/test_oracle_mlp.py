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


def test_bf16_operand_forward_by_hand():
    """The matrix-core inference's arithmetic (forward(bf16_operands=True)): one input
    of 1 feeds hidden unit 0 of layer 1 with weight 1 + 2^-8 and unit 1 with
    1 + 3 * 2^-8. As BF16 operands (7 mantissa bits) they become 1 (the tie rounds to
    the even neighbour) and 1 + 2^-6 (rounding up); layer 2 passes each to its own
    unit with weight 1, and logit c reads unit c with weight 2^c. The exact forward
    keeps the unrounded values. An input of 1 + 2^-9 standardises to itself and rounds
    to 1 as well. A mean and scale that land it on 2.5 (exact in BF16) give hidden
    values 2.5 (1 + 2^-8) = 2.509765625 and 2.5 (1 + 3 * 2^-8) = 2.529296875, which
    round on BF16's grid of 2^-6 in [2, 4) to 2.515625 and 2.53125."""
    w = np.zeros(mlp.PARAMS)
    w1, b1, w2, b2, w3, b3 = [a for a in mlp.unpack(w)]  # views into w
    w1[0, 0] = 1 + 2.0 ** -8
    w1[0, 1] = 1 + 3 * 2.0 ** -8
    w2[0, 0] = w2[1, 1] = 1.0
    w3[0, 0], w3[1, 1] = 1.0, 2.0
    x = np.zeros((1, mlp.INPUTS))
    x[0, 0] = 1.0
    z, _ = mlp.forward(w, x, bf16_operands=True)
    assert np.array_equal(z[0], [1.0, 2 * (1 + 2.0 ** -6), 0.0, 0.0])
    z, _ = mlp.forward(w, x)
    assert np.array_equal(z[0], [1 + 2.0 ** -8, 2 * (1 + 3 * 2.0 ** -8), 0.0, 0.0])
    x[0, 0] = 1 + 2.0 ** -9
    z, _ = mlp.forward(w, x, bf16_operands=True)
    assert np.array_equal(z[0], [1.0, 2 * (1 + 2.0 ** -6), 0.0, 0.0])
    mean, scale = np.zeros(mlp.INPUTS), np.ones(mlp.INPUTS)
    mean[0], scale[0] = 1 + 2.0 ** -9 - 1.25, 2.0
    z, _ = mlp.forward(w, x, mean, scale, bf16_operands=True)
    assert np.array_equal(z[0], [2.515625, 2 * 2.53125, 0.0, 0.0])


def test_bf16_operand_gradients_by_hand():
    """The matrix-core training's arithmetic (loss_and_grads(bf16_operands=True)) on one
    sample: input 0 is 1; layer 1 passes it to unit 0 with weight 1 + 2^-8 (a BF16
    operand of 1: the tie rounds to even), layers 2 and 3 pass unit 0 on with weight 1
    to logit 0; the label is class 1. So z = (1, 0, 0, 0), dz = softmax(z) - e_1, and
    every delta on the path is dz_0 rounded to BF16 (d2_0 = W3[0,0] bf16(dz_0),
    d1_0 = W2[0,0] bf16(d2_0)). The weight gradients multiply rounded operands
    (dW2[0,0] = h1_0 bf16(d2_0) with h1_0 = 1, not 1 + 2^-8), the bias gradients
    sum the deltas before rounding, and every other entry is zero."""
    w = np.zeros(mlp.PARAMS)
    w1, b1, w2, b2, w3, b3 = mlp.unpack(w)
    w1[0, 0] = 1 + 2.0 ** -8
    w2[0, 0] = 1.0
    w3[0, 0] = 1.0
    x = np.zeros((1, mlp.INPUTS))
    x[0, 0] = 1.0
    loss, g = mlp.loss_and_grads(w, x, [1], bf16_operands=True)
    p = np.exp([1.0, 0.0, 0.0, 0.0])
    p /= p.sum()
    dz = p - np.array([0.0, 1.0, 0.0, 0.0])
    dzb = mlp.bf16(dz.astype(np.float32)).astype(np.float64)
    assert dzb[0] != dz[0] and dzb[1] != dz[1]  # the rounding shows
    assert abs(loss + math.log(p[1])) < 1e-15
    gw1, gb1, gw2, gb2, gw3, gb3 = mlp.unpack(g)
    want = [np.zeros(s_) for s_ in mlp.SHAPES]
    want[0][0, 0] = dzb[0]        # x_0 d1_0
    want[1][0] = dzb[0]           # d1_0 (the rounded d2_0 times W2[0,0] = 1)
    want[2][0, 0] = dzb[0]        # h1_0 bf16(d2_0), h1_0 = 1
    want[3][0] = dzb[0]           # d2_0 = W3[0,0] bf16(dz_0)
    want[4][0, :] = dzb           # h2_0 bf16(dz)
    want[5][:] = dz               # dz, unrounded
    for got, exp in zip((gw1, gb1, gw2, gb2, gw3), want[:5]):
        assert np.array_equal(got, exp)
    assert np.abs(gb3 - dz).max() < 1e-15  # the softmax's float64 rounding only
    # Without the rounding, h1_0 = h2_0 = z_0 = 1 + 2^-8 all the way through.
    _, exact = mlp.loss_and_grads(w, x, [1])
    z0 = 1 + 2.0 ** -8
    p0 = math.exp(z0) / (math.exp(z0) + 3.0)
    assert abs(mlp.unpack(exact)[2][0, 0] - z0 * p0) < 1e-15


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


def test_extra_stats_known_answers():
    """Extra statistics per view (largest NDC gradient length, summed depth, summed
    NDC gradient x and y) on a 200x100 viewport, worked by hand. dL/dmean of
    (0.03, 0.08) is the NDC gradient (3, 4), length 5; its negative sums back to
    the start; a radius-0 splat leaves its row alone."""
    splats = np.zeros((3, 12))
    splats[:, 2] = [2.5, 4.0, 9.0]  # depth
    splats[:, 3] = [7.0, 3.0, 0.0]  # radius
    grads = np.zeros((3, 9))
    grads[0, :2] = [0.03, 0.08]
    grads[1, :2] = [-0.03, -0.08]
    grads[2, :2] = [1.0, 1.0]
    extra = np.array([[2.0, 10.0, 1.0, -1.0], [7.5, 0.0, 0.0, 0.0], [1.0, 2.0, 3.0, 4.0]])
    out = mlp.extra_stats(splats, grads, [200.0, 100.0], extra)
    want = np.array([[5.0, 12.5, 4.0, 3.0],    # max(2, 5); 10 + 2.5; 1 + 3; -1 + 4
                     [7.5, 4.0, -3.0, -4.0],   # max(7.5, 5) keeps 7.5; signs kept in the sums
                     [1.0, 2.0, 3.0, 4.0]])    # culled
    assert np.allclose(out, want, rtol=0, atol=1e-12), out
    # a dropped 0.5 or pixel units would move the first column away from 5
    assert not np.isclose(np.hypot(0.03 * 200, 0.08 * 100), 5.0)
    assert not np.isclose(np.hypot(0.03, 0.08), 5.0)
    # the input array is not modified in place
    assert extra[0, 0] == 2.0


def test_features_read_the_extra_statistics():
    """The direction-variance input is 1 - |summed NDC gradient| / summed length:
    two opposite views cancel to 1, two equal views give 0."""
    params = np.zeros((2, 16))
    params[:, 4:7] = math.log(0.5)
    params[:, 8] = 1.0
    stats = np.array([[10.0, 2.0, 3.0, 0.0], [10.0, 2.0, 3.0, 0.0]])
    extra = np.array([[5.0, 6.0, 0.0, 0.0], [5.0, 6.0, 6.0, 8.0]])
    f = mlp.features(params, stats, extra, extent=1.0, iteration=10)
    assert np.allclose(f[:, 11], [1.0, 0.0])
    assert np.allclose(f[:, 1], math.log1p(5.0 / 2e-4))
    assert np.allclose(f[:, 10], math.log(3.0))  # summed depth 6 over 2 views
