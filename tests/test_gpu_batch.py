"""The multi-view batch trainer on the GPU (views of uneven sizes per step, their
gradients summed into one Adam step, the step replayed as a captured graph).

Each batch step is checked against the single-frame path, which the other GPU
tests pin to the oracle: the same views rendered, scored and back-propagated one
at a time by GaussianRenderer + TrainingLoss, their gradients summed, then the
oracle's Adam step and statistics on those sums. Three steps in a row cover the
captured graph's first launch and two replays with the views in other orders."""
import math

import numpy as np
import pytest

import diorama_amd
from diorama_amd import scene
from oracle import oracle
from tests.gpu_util import context, say

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (200, 120), (96, 160)]


def orbit_cameras(sizes, radius=4.0, fov=60.0, phase=0.0):
    cams = []
    for k, (w, h) in enumerate(sizes):
        a = phase + 2 * math.pi * k / len(sizes)
        eye = [radius * math.sin(a), 0.3 * math.cos(3 * a), radius * math.cos(a)]
        view = scene.look_at(eye, [0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        proj = scene.perspective(math.radians(fov), w / h, 0.01, 100.0)
        cams.append(scene.camera(view, proj, w, h))
    return cams


def teacher_images(ctx, n, cams, seed=1):
    p, sh = scene.random_raw_gaussians(n, seed, box=1.0)
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(scene.activate(p), sh, 3)
    out = []
    for cam in cams:
        w, h = int(cam["viewport"][0]), int(cam["viewport"][1])
        target = ctx.buffer(w * h * 4)
        r.encode(target, cam, [0, 0, 0])
        out.append(target.download(np.uint32, w * h))
    r.destroy()
    return out


def single_view(ctx, act, sh, degree, cam, image):
    """One view through the single-frame path: gradients, splats and loss."""
    w, h = int(cam["viewport"][0]), int(cam["viewport"][1])
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(act, sh, degree)
    r.encode(None, cam, [0, 0, 0])
    loss = diorama_amd.TrainingLoss(ctx, w, h)
    loss.encode(r.pixel_buffer, ctx.upload(image))
    value = loss.read()
    r.encode_backward(loss.pixel_grads)
    grads = r.gradients
    n = len(act)
    out = (grads["gaussians"].download(np.float32).reshape(-1, 16)[:n].astype(np.float64),
           grads["sh"].download(np.float32).reshape(-1, 45)[:n].astype(np.float64),
           grads["splats"].download(np.float32).reshape(-1, 9)[:n].astype(np.float64),
           r.splat_buffer.download(np.float32).reshape(-1, 12)[:n].astype(np.float64), value)
    loss.destroy()
    r.destroy()
    return out


def test_batch_steps_match_summed_single_view_steps():
    ctx = context()
    n = 4096
    cams = orbit_cameras(SIZES)
    images = teacher_images(ctx, n, cams)
    params, sh = scene.random_raw_gaussians(n, 0, box=1.0)
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, diorama_amd.default_settings(1.0), batch=3)
    assert t.batch_size == 3
    t.sh_degree = 3
    rng = np.random.default_rng(2)
    moments = rng.normal(size=(n, 122)).astype(np.float32) * 1e-3
    moments[:, 1::2] = np.abs(moments[:, 1::2])
    t.buffers["moments"].upload(moments)
    t.buffers["stats"].upload(np.abs(rng.normal(size=(n, 4))).astype(np.float32))
    for k in range(1, 4):
        p0 = t.buffers["params"].download(np.float32).reshape(n, 16)
        sh0 = t.renderer.sh_buffer.download(np.float32).reshape(n, 45)
        m0 = t.buffers["moments"].download(np.float32).reshape(n, 122)
        st0 = t.buffers["stats"].download(np.float32).reshape(n, 4)
        act0 = t.renderer.gaussian_buffer.download(np.float32).reshape(n, 16)
        gg = np.zeros((n, 16))
        shg = np.zeros((n, 45))
        stats = st0.astype(np.float64)
        losses = []
        say("batch step", k, "single-view sums")
        for cam, image in zip(cams, images):
            g1, s1, sg, splats, value = single_view(ctx, act0, sh0, 3, cam, image)
            gg += g1
            shg += s1
            stats = oracle.accumulate_stats(splats, sg, cam["viewport"], stats)
            losses.append(value)
        say("batch step", k)
        t.optimize()
        p1 = t.buffers["params"].download(np.float32).reshape(n, 16)
        sh1 = t.renderer.sh_buffer.download(np.float32).reshape(n, 45)
        m1 = t.buffers["moments"].download(np.float32).reshape(n, 122)
        st1 = t.buffers["stats"].download(np.float32).reshape(n, 4)
        lr = [1 - 0.9 ** k, 1 - 0.999 ** k, 1.6e-4 * (1e-2) ** (k / 30000), 0.025, 0.005, 0.001, 0.0025, 0.0025 / 20]
        o_p, o_sh, o_m, _ = oracle.adam_step(p0, sh0, m0, gg, shg, lr)
        err_p = np.abs(p1 - o_p).max()
        err_sh = np.abs(sh1 - o_sh).max()
        err_m = np.abs(m1 - o_m).max() / np.abs(o_m).max()
        err_st = np.abs(st1 - stats).max() / np.abs(stats).max()
        loss_err = abs(t.last_loss() - float(np.mean(losses)))
        print(f"batch step {k}: params {err_p:.2e}, sh {err_sh:.2e}, moments rel {err_m:.2e}, "
              f"stats rel {err_st:.2e}, loss {loss_err:.2e}")
        assert err_p < 1e-5 and err_sh < 1e-5 and err_m < 1e-4
        assert err_st < 1e-5
        assert loss_err < 1e-5
    t.destroy()


def test_batch_trainer_densifies_and_keeps_stepping():
    ctx = context()
    n = 2048
    cams = orbit_cameras([(128, 96), (96, 128), (160, 90), (64, 64)], phase=0.4)
    images = teacher_images(ctx, n, cams, seed=3)
    params, sh = scene.random_raw_gaussians(n, 4, box=1.0)
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, diorama_amd.default_settings(1.0), batch=2)
    t.iteration = 590
    stats = np.zeros((n, 4), dtype=np.float32)
    stats[:300, 0] = 1.0  # a large mean view-space gradient: these clone or split at step 600
    stats[:300, 1] = 1.0
    t.buffers["stats"].upload(stats)
    say("batch densify")
    losses = []
    for _ in range(20):  # step 600 densifies; the steps after it run on the new buffers
        t.step()
        losses.append(t.last_loss())
    assert len(t.densifications) == 1
    rec = t.densifications[0]
    assert rec["step"] == 600 and t.count == rec["after"] and rec["cloned"] + rec["split"] >= 300
    assert all(math.isfinite(v) for v in losses)
    raw = t.export_raw()
    assert np.isfinite(raw["params"]).all() and np.isfinite(raw["sh"]).all()
    t.destroy()


def test_batch_training_learns():
    ctx = context()
    cams = orbit_cameras([(160, 120), (96, 96), (200, 100), (120, 160), (64, 96), (128, 128)], radius=3.5)
    images = teacher_images(ctx, 1024, cams, seed=5)
    params, sh = scene.random_raw_gaussians(1024, 6, box=1.0)
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, diorama_amd.default_settings(1.0), batch=3)
    before = t.evaluate(cams, images, with_ssim=False)["psnr"]
    say("batch learning")
    for _ in range(200):
        t.step()
    after = t.evaluate(cams, images, with_ssim=False)["psnr"]
    print(f"batch training: PSNR {before:.2f} -> {after:.2f} dB")
    assert after > before + 3.0
    t.destroy()


@pytest.mark.parametrize("batch", [0, 3])
def test_tuning_changes_speed_not_results(batch):
    """Every tile size (8, 16, 32; the plain trainer), launch form (persistent,
    segmented backward), depth order (all splats, or each tile's list after a stable
    tile sort or after binning from counts; the plain trainer), splat batch and digit
    width gives the same images, and
    gradients equal up to float rounding (the order of float atomics; a segment's
    replay starting from the forward's checkpoint); the autotuner keeps the fastest
    setting. 200x120 has 13 x 8 tiles: the 32-px bins of the last column and row are
    partly outside the frame."""
    ctx = context()
    n = 4096
    cams = orbit_cameras(SIZES) if batch else orbit_cameras([(200, 120)] * 3)
    images = teacher_images(ctx, n, cams, seed=21)
    params, sh = scene.random_raw_gaussians(n, 22, box=1.0)
    outs = {}
    shapes = [(16, 0, 0, 0)] if batch else [(16, 0, 0, 0), (16, 1, 0, 0), (8, 0, 0, 0), (32, 0, 0, 0), (16, 0, 256, 0),
                                            (16, 0, 512, 0), (16, 0, 1024, 0), (16, 0, 0, 1), (16, 1, 0, 1),
                                            (8, 0, 0, 1), (32, 0, 0, 1), (16, 0, 512, 1), (16, 0, 0, 2),
                                            (16, 1, 0, 2), (8, 0, 0, 2), (32, 0, 0, 2), (16, 0, 512, 2)]
    for tile, persistent, segment, per_tile in shapes:
        for bits in (8, 4):
            for splats in (64, 128, 256):
                if tile == 32 and splats == 256:
                    continue
                t = diorama_amd.Trainer(ctx, params, sh, cams, images, diorama_amd.default_settings(1.0), batch=batch)
                t.sh_degree = 3
                t.tuning = {"tile_size": tile, "splat_batch": splats, "digit_bits": bits, "persistent": persistent,
                            "segment": segment, "per_tile_depth": per_tile}
                say("tuning", batch, tile, persistent, segment, per_tile, bits, splats)
                t.optimize()
                p = t.buffers["params"].download(np.float32).reshape(n, 16)
                pix = None if batch else t.renderer.pixel_buffer.download(np.float32, 200 * 120 * 4)
                sg = None if batch else t.renderer.gradients["splats"].download(np.float32).reshape(-1, 9)[:n]
                outs[(tile, persistent, segment, per_tile, bits, splats)] = (p, pix, t.last_loss(), sg)
                t.destroy()
    ref_p, ref_pix, ref_loss, ref_sg = outs[(16, 0, 0, 0, 8, 64)]
    for key, (p, pix, loss, sg) in outs.items():
        if pix is not None:
            assert np.array_equal(pix, ref_pix), key
        assert abs(loss - ref_loss) < 1e-6, key
        assert np.abs(p - ref_p).max() < 1e-5, key
        if not batch:
            assert np.abs(sg - ref_sg).max() <= 1e-5 * np.abs(ref_sg).max(), key
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, diorama_amd.default_settings(1.0), batch=batch)
    say("autotune", batch)
    results = t.autotune(steps=2)
    assert len(results) == (6 if batch else 85)
    best = min(results, key=lambda r: r[1])
    assert t.tuning == best[0]
    print("autotune:", [(r[0]["tile_size"], r[0]["persistent"], r[0]["segment"], r[0]["per_tile_depth"],
                         r[0]["splat_batch"], r[0]["digit_bits"], round(r[1], 3)) for r in results])
    with pytest.raises(diorama_amd.DioramaError):
        t.tuning = {"tile_size": 8 if batch else 12, "splat_batch": 64, "digit_bits": 8}
    with pytest.raises(diorama_amd.DioramaError):
        t.tuning = {"tile_size": 32, "splat_batch": 256, "digit_bits": 8}
    with pytest.raises(diorama_amd.DioramaError):
        t.tuning = {"tile_size": 8, "splat_batch": 64, "digit_bits": 8, "persistent": 1}
    for bad in ({"tile_size": 16, "segment": 128}, {"tile_size": 8, "segment": 256},
                {"tile_size": 16, "persistent": 1, "segment": 256}, {"tile_size": 16, "per_tile_depth": 3}):
        with pytest.raises(diorama_amd.DioramaError):
            t.tuning = dict({"splat_batch": 64, "digit_bits": 8}, **bad)
    if batch:
        for per_tile in (1, 2):
            with pytest.raises(diorama_amd.DioramaError):
                t.tuning = {"tile_size": 16, "splat_batch": 64, "digit_bits": 8, "per_tile_depth": per_tile}
    t.destroy()
