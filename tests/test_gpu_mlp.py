"""The learned densification on the GPU against the float64 oracle (oracle/mlp.py):
BF16 weights, the forward pass, the cross-entropy gradients, training, the 12
per-Gaussian inputs and their statistics; then the trainer's logging, learned and
budgeted densification end to end, and what the trained network does against the
plain rule at an equal Gaussian budget."""
import math

import numpy as np
import pytest

import diorama_amd
from diorama_amd import scene
from oracle import mlp as omlp
from oracle import oracle
from tests.gpu_util import config0_scene, context, say
from tests.test_gpu_batch import orbit_cameras, teacher_images

pytestmark = pytest.mark.gpu


def random_weights(seed, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, size=omlp.PARAMS).astype(np.float32)


def test_weights_round_to_bf16():
    ctx = context()
    net = diorama_amd.DensifyMLP(ctx, seed=3)
    w = random_weights(1)
    net.weights = w
    assert np.array_equal(net.weights.view(np.uint32), omlp.bf16(w).view(np.uint32))
    net.destroy()


def test_forward_and_gradients_against_oracle():
    ctx = context()
    net = diorama_amd.DensifyMLP(ctx)
    w = random_weights(2)
    net.weights = w
    wb = omlp.bf16(w).astype(np.float64)
    rng = np.random.default_rng(4)
    n = 3001
    x = rng.normal(0, 1, size=(n, omlp.INPUTS)).astype(np.float32)
    labels = rng.integers(0, 4, size=n).astype(np.uint32)
    fx, lb = ctx.upload(x), ctx.upload(labels)
    logits = ctx.buffer(n * 16)
    mean = rng.normal(0, 0.5, size=12).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=12).astype(np.float32)
    net.set_input_norm(mean, scale)
    say("mlp forward")
    net.forward(fx, n, logits)
    z = logits.download(np.float32).reshape(n, 4)
    # Inference runs on the matrix cores with BF16 operands: the oracle rounds the
    # same operands. A hidden value whose FP32 sum lies within an FP32 ulp of a BF16
    # tie can round the other way (the sums' order differs), which moves that
    # sample's logits by about a BF16 ulp of one hidden unit times its weights.
    oz, _ = omlp.forward(wb, x, mean, scale, bf16_operands=True)
    per_sample = np.abs(z - oz).max(axis=1) / np.abs(oz).max()
    err = float(per_sample.max())
    exact_share = float((per_sample < 1e-5).mean())
    print(f"mlp forward (BF16 operands): {exact_share:.4f} of samples within 1e-5, worst {err:.2e}")
    assert exact_share > 0.98
    assert err < 5e-3
    say("mlp gradients")
    cw = [0.5, 2.0, 3.0, 1.5]
    net.set_class_weights(cw)
    g, loss = net.gradients(fx, lb, n)
    # Training runs on the matrix cores too: the oracle rounds the same operands.
    oloss, og = omlp.loss_and_grads(wb, x, labels, cw, mean, scale, bf16_operands=True)
    gerr = np.abs(g - og).max() / np.abs(og).max()
    print(f"mlp: loss {loss:.6f} vs {oloss:.6f}, gradient rel err {gerr:.2e}")
    assert abs(loss - oloss) < 1e-5 * max(1.0, oloss)
    assert gerr < 1e-4
    net.destroy()


def test_training_fits_a_separable_labelling():
    ctx = context()
    net = diorama_amd.DensifyMLP(ctx, seed=5)
    rng = np.random.default_rng(6)
    n = 20000
    x = rng.normal(0, 1, size=(n, omlp.INPUTS)).astype(np.float32)
    teacher = rng.normal(0, 1, size=(omlp.INPUTS, 4))
    labels = np.argmax(x @ teacher, axis=1).astype(np.uint32)
    fx, lb = ctx.upload(x), ctx.upload(labels)
    _, start = net.gradients(fx, lb, n)
    say("mlp train")
    net.train(fx, lb, n, steps=400, batch=1024, lr=3e-3, seed=9)
    _, end = net.gradients(fx, lb, n)
    logits = ctx.buffer(n * 16)
    net.forward(fx, n, logits)
    accuracy = float(np.mean(np.argmax(logits.download(np.float32).reshape(n, 4), axis=1) == labels))
    print(f"mlp training: loss {start:.3f} -> {end:.3f}, accuracy {accuracy:.3f}")
    assert end < 0.5 * start and accuracy > 0.9
    net.destroy()


def test_inputs_and_their_statistics_against_oracle():
    params, sh, g, cam = config0_scene(seed=0, n=4096)
    teacher_p, teacher_sh = scene.random_raw_gaussians(4096, 1, box=1.0)
    ctx = context()
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(scene.activate(teacher_p), teacher_sh, 3)
    target = ctx.buffer(256 * 256 * 4)
    r.encode(target, cam, [0, 0, 0])
    image = target.download(np.uint32)
    r.destroy()
    t = diorama_amd.Trainer(ctx, params, sh, [cam], [image], diorama_amd.default_settings(1.0))
    t.set_densify_mode("log")
    t.sh_degree = 3
    extra0 = t.extra_stats.download(np.float32).reshape(-1, 4)
    assert not extra0.any()
    say("learned inputs")
    for _ in range(3):
        t.optimize()
        rend = t.renderer
        splats = rend.splat_buffer.download(np.float32).reshape(-1, 12)[:4096]
        sgrad = rend.gradients["splats"].download(np.float32).reshape(-1, 9)[:4096]
        extra0 = omlp.extra_stats(splats, sgrad, cam["viewport"], extra0)
    extra = t.extra_stats.download(np.float32).reshape(-1, 4)
    scale = np.maximum(np.abs(extra0).max(axis=0), 1e-12)
    assert (np.abs(extra - extra0) / scale).max() < 1e-5
    feats = t.features().download(np.float32, 4096 * 12).reshape(-1, 12)
    p = t.buffers["params"].download(np.float32).reshape(-1, 16)
    stats = t.buffers["stats"].download(np.float32).reshape(-1, 4)
    want = omlp.features(p, stats, extra, 1.0, t.iteration)
    err = np.abs(feats - want).max()
    print(f"learned inputs: statistics rel err {(np.abs(extra - extra0) / scale).max():.2e}, inputs err {err:.2e}")
    assert err < 1e-4
    t.destroy()


def densify_run(ctx, cams, images, params, sh, mode, steps, net=None, budget=0, extent=1.0):
    settings = diorama_amd.default_settings(extent)
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, settings)
    t.set_densify_mode(mode, net, budget=budget, samples_per_densify=20000)
    for _ in range(steps):
        t.step()
    return t


def test_logging_learning_and_budget_end_to_end():
    """Steps 1-1200 logging under the plain rule, then the MLP trained on those
    samples densifies a fresh run; both runs hold the same budget."""
    ctx = context()
    cams = orbit_cameras([(160, 120)] * 8, radius=3.5)
    images = teacher_images(ctx, 8192, cams, seed=11)
    params, sh = scene.random_raw_gaussians(2000, 12, box=1.0)
    say("logging run")
    t = densify_run(ctx, cams, images, params, sh, "log", 1201, budget=6000)
    feats, labels, count = t.samples()
    assert count > 1000
    lab = labels.download(np.uint32, count)
    assert lab.max() <= 3
    assert all(d["after"] <= 6000 for d in t.densifications)
    print(f"logged {count} samples, classes {np.bincount(lab, minlength=4).tolist()}")
    net = diorama_amd.DensifyMLP(ctx, seed=13)
    loss = net.train(feats, labels, count, steps=500, batch=2048, lr=2e-3, seed=3)
    assert math.isfinite(loss)
    t.destroy()
    say("learned run")
    t2 = densify_run(ctx, cams, images, params, sh, "learned", 1201, net=net, budget=6000)
    assert t2.densifications and all(d["after"] <= 6000 for d in t2.densifications)
    raw = t2.export_raw()
    assert np.isfinite(raw["params"]).all()
    t2.destroy()
    net.destroy()


def test_footprint_loss_against_oracle():
    """Mode "log" gathers each Gaussian's footprint terms in the backward pass: the
    sums over its blended pixels of alpha T x the pixel's loss share, and of alpha T
    (oc_footprint), given the GPU's splats, image and pair lists; two steps add up."""
    params, sh, g, cam = config0_scene(seed=2, n=4096)
    teacher_p, teacher_sh = scene.random_raw_gaussians(4096, 3, box=1.0)
    ctx = context()
    image = scene.teacher_images(ctx, [cam], 4096, seed=3, box=1.0, noise=0)[0]
    target = scene.rgba8_channels(image)
    t = diorama_amd.Trainer(ctx, params, sh, [cam], [image], diorama_amd.default_settings(1.0))
    t.set_densify_mode("log")
    t.sh_degree = 3
    want = np.zeros((4096, 2))
    say("footprint")
    for _ in range(2):
        t.optimize()
        r = t.renderer
        r.tiles.after_submit()
        splats = r.splat_buffer.download(np.float32).reshape(-1, 12)[:4096].astype(np.float64)
        color = r.pixel_buffer.download(np.float32, 256 * 256 * 4).reshape(-1, 4).astype(np.float64)
        ranges, pairs = oracle.tile_lists(splats, 256, 256)
        _, _, shares = oracle.loss(256, 256, color, target)
        want += oracle.footprint(splats, ranges, pairs, 256, 256, shares)
    got = t.footprint.download(np.float32, 4096 * 2).reshape(-1, 2)
    err = [float(np.abs(got[:, k] - want[:, k]).max() / np.abs(want[:, k]).max()) for k in range(2)]
    print(f"footprint: err / max {err}; {int((want[:, 1] > 0).sum())} Gaussians blended")
    assert max(err) < 1e-4
    t.destroy()


def test_learned_densification_against_plain_at_a_binding_budget():
    """What the trained network gains: the plain rule (logging samples whose footprint
    loss fell over the next 100 steps) and the MLP trained on them each grow 2,000
    Gaussians under a 3,000 budget that binds (the rule alone grows past it), then
    are scored on held-out views. Like src/checks/training.ts:99-104, the learned run
    must not lose: at most 0.1 dB below the plain rule."""
    ctx = context()
    cams = orbit_cameras([(160, 120)] * 12, radius=3.5)
    held, train = cams[::4], [c for k, c in enumerate(cams) if k % 4]
    images = teacher_images(ctx, 8192, cams, seed=11)
    held_images, train_images = images[::4], [im for k, im in enumerate(images) if k % 4]
    params, sh = scene.random_raw_gaussians(2000, 12, box=1.0)
    steps = 2001
    say("unbudgeted rule")
    free = densify_run(ctx, train, train_images, params, sh, "plain", steps)
    grown = free.count
    free.destroy()
    budget = 3000
    assert grown > budget, f"the rule alone grows to {grown}: the budget would not bind"
    say("plain rule, logging")
    a = densify_run(ctx, train, train_images, params, sh, "log", steps, budget=budget)
    plain = a.evaluate(held, held_images, with_ssim=False)["psnr"]
    feats, labels, count = a.samples()
    lab = labels.download(np.uint32, count)
    x = feats.download(np.float32, count * 12).reshape(count, 12).astype(np.float64)
    net = diorama_amd.DensifyMLP(ctx, seed=13)
    net.set_input_norm(x.mean(axis=0), 1.0 / np.maximum(x.std(axis=0), 1e-6))
    classes = np.bincount(lab, minlength=4)
    net.set_class_weights([min(30.0, math.sqrt(count / (4.0 * max(c, 1)))) for c in classes])
    net.train(feats, labels, count, steps=600, batch=2048, lr=2e-3, seed=3)
    plain_count = a.count
    a.destroy()
    say("learned")
    b = densify_run(ctx, train, train_images, params, sh, "learned", steps, net=net, budget=budget)
    learned = b.evaluate(held, held_images, with_ssim=False)["psnr"]
    learned_count = b.count
    b.destroy()
    net.destroy()
    print(f"binding budget {budget} (the rule alone grows to {grown}): plain {plain:.3f} dB with {plain_count}, "
          f"learned {learned:.3f} dB with {learned_count} Gaussians; delta {learned - plain:+.3f} dB; "
          f"{count} samples {classes.tolist()}")
    assert plain_count <= budget and learned_count <= budget
    assert learned >= plain - 0.1
