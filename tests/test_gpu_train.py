"""The Trainer on the GPU: Adam, statistics, densification and opacity reset against
the oracle (src/checks/densify.ts:28-185), and end-to-end learning
(src/checks/training.ts:20-120)."""
import math

import numpy as np
import pytest

import diorama_amd
from diorama_amd import scene
from oracle import oracle
from tests.gpu_util import config0_scene, context, say
from tests.test_oracle import densify_case, rotation_columns

pytestmark = pytest.mark.gpu


def render_views(gaussians, cameras, sh=None, degree=0):
    """Photos as rgba8 images rendered by the GPU renderer itself."""
    ctx = context()
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(gaussians, sh, degree)
    images = []
    for cam in cameras:
        w, h = int(cam["viewport"][0]), int(cam["viewport"][1])
        target = ctx.buffer(w * h * 4)
        r.encode(target, cam, [0, 0, 0])
        images.append(target.download(np.uint32, w * h))
    r.destroy()
    return images


@pytest.mark.parametrize("n", [4096, 4133])  # whole 64-Gaussian blocks; a partial last block
def test_adam_and_stats_step_against_oracle(n):
    params, sh, g, cam = config0_scene(seed=0, n=n)
    teacher_p, teacher_sh = scene.random_raw_gaussians(n, 1, box=1.0)
    images = render_views(scene.activate(teacher_p), [cam], teacher_sh, 3)
    ctx = context()
    settings = diorama_amd.default_settings(1.0)
    say("adam step", n)
    t = diorama_amd.Trainer(ctx, params, sh, [cam], images, settings)
    moments = np.random.default_rng(2).normal(size=(n, 122)).astype(np.float32) * 1e-3
    moments[:, 1::2] = np.abs(moments[:, 1::2])
    t.buffers["moments"].upload(moments)
    stats0 = np.abs(np.random.default_rng(3).normal(size=(n, 4))).astype(np.float32)
    t.buffers["stats"].upload(stats0)
    r = t.renderer
    act_before = r.gaussian_buffer.download(np.float32).reshape(-1, 16)
    t.optimize()
    # The step's preprocess backward, statistics and Adam run fused (the per-Gaussian
    # gradients stay on chip), so the expected gradients come from the oracle's
    # preprocess backward of the step's own splat gradients.
    sgrad = r.gradients["splats"].download(np.float32).reshape(-1, 9)
    splats = r.splat_buffer.download(np.float32).reshape(-1, 12)
    gg, shg = oracle.preprocess_backward(act_before, sh, t.sh_degree, cam, splats, sgrad)
    p_after = t.buffers["params"].download(np.float32).reshape(-1, 16)
    sh_after = r.sh_buffer.download(np.float32).reshape(-1, 45)
    m_after = t.buffers["moments"].download(np.float32).reshape(-1, 122)
    stats = t.buffers["stats"].download(np.float32).reshape(-1, 4)
    act = r.gaussian_buffer.download(np.float32).reshape(-1, 16)
    lr = [1 - 0.9, 1 - 0.999, 1.6e-4 * (1e-2) ** (1 / 30000), 0.025, 0.005, 0.001, 0.0025, 0.0025 / 20]
    o_p, o_sh, o_m, o_act = oracle.adam_step(params, sh, moments, gg, shg, lr)
    err_p = np.abs(p_after - o_p).max()
    err_sh = np.abs(sh_after - o_sh).max()
    err_m = np.abs(m_after - o_m).max() / np.abs(o_m).max()
    print(f"adam: params err {err_p:.2e}, sh err {err_sh:.2e}, moments rel err {err_m:.2e}")
    assert err_p < 1e-6 and err_sh < 1e-6 and err_m < 1e-5
    cols = [k for k in range(16) if k not in (7, 15)]
    assert np.allclose(act[:, cols], o_act[:, cols], rtol=1e-5, atol=1e-6)
    o_stats = oracle.accumulate_stats(splats, sgrad, cam["viewport"], stats0)
    assert np.allclose(stats, o_stats, rtol=1e-5, atol=1e-7)
    t.destroy()


@pytest.mark.parametrize("prune_large", [False, True])
def test_densify_against_oracle(prune_large):
    params, sh, moments, stats = densify_case(41 if prune_large else 43)
    ctx = context()
    cam = scene.camera(scene.IDENTITY, scene.IDENTITY, 16, 16)
    image = np.zeros(256, dtype=np.uint32)
    settings = diorama_amd.default_settings(1.0)
    say("densify", prune_large)
    t = diorama_amd.Trainer(ctx, params, sh, [cam], [image], settings)
    t.buffers["moments"].upload(moments)
    t.buffers["stats"].upload(stats)
    t.iteration = 700
    t.densify(prune_large)
    count, o_params, o_sh, o_m, tally = oracle.densify(params, sh, moments, stats,
                                                       [0.0002, 0.01, 0.005, 0.1 if prune_large else 0.0], 700)
    assert t.count == count
    rec = t.densifications[0]
    assert [rec["removed"], rec["kept"], rec["cloned"], rec["split"]] == list(tally)
    raw = t.export_raw()
    got_p = raw["params"]
    child = got_p[:, 7] == 2
    assert np.array_equal(child, o_params[:, 7] == 2)
    err_pos = np.abs(got_p[child, 0:3] - o_params[child, 0:3]).max() if child.any() else 0.0
    print(f"densify: {count} out; split position err {err_pos:.2e}")
    assert err_pos < 1e-4
    assert np.allclose(got_p[:, 3:], o_params[:, 3:], atol=1e-5)
    assert np.allclose(got_p[~child, 0:3], o_params[~child, 0:3], atol=0)
    assert np.array_equal(raw["sh"], o_sh.astype(np.float32))
    m = t.buffers["moments"].download(np.float32).reshape(-1, 122)
    assert np.array_equal(m, o_m.astype(np.float32))
    assert not t.buffers["stats"].download(np.float32).any()
    act = t.renderer.gaussian_buffer.download(np.float32).reshape(-1, 16)
    assert np.allclose(act, scene.activate(got_p) * (np.arange(16) % 8 != 7), rtol=1e-5, atol=1e-6)
    # The opacity reset right after: only opacity and its Adam history change.
    t.reset_opacity()
    p2 = t.buffers["params"].download(np.float32).reshape(-1, 16)
    m2 = t.buffers["moments"].download(np.float32).reshape(-1, 122)
    capped = math.log(0.01 / 0.99)
    assert np.allclose(p2[:, 3], np.minimum(got_p[:, 3], capped), atol=1e-5)
    assert np.all(m2[:, 6:8] == 0)
    keep = np.ones(122, dtype=bool)
    keep[6:8] = False
    assert np.array_equal(m2[:, keep], m[:, keep])
    t.destroy()


def test_densify_removing_everything_is_an_error():
    """src/trainer.ts:295."""
    ctx = context()
    params = np.zeros((8, 16), dtype=np.float32)
    params[:, 3] = -20  # opacity ~0: all pruned
    params[:, 8] = 1
    cam = scene.camera(scene.IDENTITY, scene.IDENTITY, 16, 16)
    t = diorama_amd.Trainer(ctx, params, np.zeros((8, 45), np.float32), [cam], [np.zeros(256, np.uint32)],
                            diorama_amd.default_settings(1.0))
    with pytest.raises(diorama_amd.DioramaError, match="removed every Gaussian"):
        t.densify(False)
    t.destroy()


def test_uneven_views_are_an_error():
    ctx = context()
    cams = [scene.camera(scene.IDENTITY, scene.IDENTITY, 16, 16), scene.camera(scene.IDENTITY, scene.IDENTITY, 8, 8)]
    params = np.zeros((4, 16), dtype=np.float32)
    params[:, 8] = 1
    with pytest.raises(diorama_amd.DioramaError, match="same size"):
        diorama_amd.Trainer(ctx, params, np.zeros((4, 45), np.float32), cams,
                            [np.zeros(256, np.uint32), np.zeros(64, np.uint32)], diorama_amd.default_settings(1.0))


def training_scene():
    """src/checks/training.ts:20-80: a 60-Gaussian truth seen from 8 ring views at 96x96."""
    random = scene.mulberry32(17)
    between = lambda lo, hi: lo + (hi - lo) * random()
    truth = []
    for _ in range(60):
        position = [between(-1, 1), between(-1, 1), between(-1, 1)]
        s = [between(0.08, 0.3), between(0.08, 0.3), between(0.08, 0.3)]
        q = [between(-1, 1), between(-1, 1), between(-1, 1), between(-1, 1)]
        color = [random(), random(), random()]
        truth.append(position + [between(0.6, 0.95)] + s + [0] + q + color + [0])
    truth = np.array(truth, dtype=np.float32)
    cams = []
    for k in range(8):
        a = k / 8 * 2 * math.pi
        eye = [3.5 * math.sin(a), 1.2, 3.5 * math.cos(a)]
        cams.append(scene.camera(scene.look_at(eye, [0, 0, 0], [0, 1, 0]), scene.perspective(math.pi / 3.5, 1, 0.1, 100),
                                 96, 96))
    images = render_views(truth, cams)
    scrambled = np.zeros((60, 16), dtype=np.float32)
    for i in range(60):
        for k in range(3):
            scrambled[i, k] = truth[i, k] + between(-0.15, 0.15)
            scrambled[i, 4 + k] = math.log(truth[i, 4 + k] * between(0.6, 1.6))
            scrambled[i, 12 + k] = (random() - 0.5) / scene.SH_C0
        scrambled[i, 3] = 0
        scrambled[i, 8:12] = [1, 0, 0, 0]
    sparse = np.zeros((12, 16), dtype=np.float32)
    for i in range(12):
        for k in range(3):
            sparse[i, k] = between(-0.8, 0.8)
            sparse[i, 4 + k] = math.log(0.25)
        sparse[i, 3] = math.log(0.1 / 0.9)
        sparse[i, 8:12] = [1, 0, 0, 0]
    return cams, images, scrambled, sparse


def train(params, cams, images, steps=400, **overrides):
    ctx = context()
    settings = diorama_amd.default_settings(3.85)
    settings.update(overrides)
    n = len(params)
    t = diorama_amd.Trainer(ctx, params, np.zeros((n, 45), np.float32), cams, images, settings)
    before = t.evaluate(cams, images, with_ssim=False)["psnr"]
    for _ in range(steps):
        t.step()
    after = t.evaluate(cams, images, with_ssim=False)["psnr"]
    count = t.count
    loss = t.last_loss()
    t.destroy()
    return before, after, count, loss


def test_training_learns_and_densifying_helps():
    cams, images, scrambled, sparse = training_scene()
    say("training fit")
    before, after, _, loss = train(scrambled, cams, images)
    print(f"training: PSNR {before:.1f} -> {after:.1f} dB in 400 steps, last loss {loss:.4f}")
    assert after - before >= 6
    say("training sparse")
    _, plain, _, _ = train(sparse.copy(), cams, images, densify_until=0)
    _, dens, count, _ = train(sparse.copy(), cams, images, densify_from=100, densify_interval=50, densify_until=351)
    print(f"densify: {dens:.1f} dB with {count} Gaussians vs {plain:.1f} dB without")
    assert count > 12 and dens > plain + 1


STAGES = ["preprocess", "depth_sort", "scan", "duplicate", "tile_sort", "ranges", "list_order", "rasterize", "loss",
          "clear_grads", "rasterize_backward", "backward_adam"]


def test_stage_clock_times_plain_steps():
    """The stage clock: one entry per stage of every plain step, each step's own pair
    count, densification steps left out, and the totals start over when re-enabled."""
    params, sh, g, cam = config0_scene(seed=4, n=4096)
    images = render_views(g, [cam], sh, 3)
    ctx = context()
    settings = diorama_amd.default_settings(1.0)
    settings.update(densify_from=0, densify_interval=4, densify_until=10)
    say("stage clock")
    t = diorama_amd.Trainer(ctx, params, sh, [cam], images, settings)
    tiles = t.renderer.tiles
    t.stage_clock(True)
    pairs = 0
    for _ in range(3):  # steps 1-3: plain
        t.step()
        tiles.after_submit()
        pairs += tiles.last_pair_count
    clock = t.stage_totals()
    assert list(clock["ms"]) == STAGES
    assert clock["steps"] == 3 and clock["pairs"] == pairs > 0
    assert all(ms > 0 for ms in clock["ms"].values())
    t.step()  # step 4 densifies: not clocked
    assert t.stage_totals()["steps"] == 3
    t.stage_clock(True)
    assert t.stage_totals()["steps"] == 0 and t.stage_totals()["pairs"] == 0
    t.step()
    assert t.stage_totals()["steps"] == 1
    t.stage_clock(False)
    t.step()
    assert t.stage_totals()["steps"] == 0
    t.destroy()


def fresh_splats(gaussians, sh, degree, cam):
    """The splats a fresh preprocess of these activated Gaussians gives for `cam`."""
    ctx = context()
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(gaussians, sh, degree)
    w, h = int(cam["viewport"][0]), int(cam["viewport"][1])
    target = ctx.buffer(w * h * 4)
    r.encode(target, cam, [0, 0, 0])
    out = r.splat_buffer.download(np.float32).reshape(-1, 12)
    r.destroy()
    return out


def test_prepared_next_frame_equals_a_fresh_preprocess():
    """A plain step's fused backward also runs the next step's preprocess; that step
    renders from exactly what a fresh preprocess of the updated Gaussians gives."""
    params, sh, g, cam = config0_scene(seed=6, n=5000)
    images = render_views(g, [cam], sh, 3)
    ctx = context()
    settings = diorama_amd.default_settings(1.0)
    settings.update(densify_until=0)
    say("prepared frame")
    t = diorama_amd.Trainer(ctx, params, sh, [cam], images, settings)
    t.iteration = 5000
    t.sh_degree = 3
    t.step()  # preprocesses itself, prepares the next frame
    r = t.renderer
    g1 = r.gaussian_buffer.download(np.float32).reshape(-1, 16)
    sh1 = r.sh_buffer.download(np.float32).reshape(-1, 45)
    t.step()  # renders from the prepared frame
    got = r.splat_buffer.download(np.float32).reshape(-1, 12)
    want = fresh_splats(g1, sh1, 3, cam)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    t.destroy()


def test_prepared_frame_is_dropped_after_a_buffer_write():
    """A device write through the C ABI between steps (here new activated positions)
    invalidates the prepared frame: the next step preprocesses what is there now."""
    params, sh, g, cam = config0_scene(seed=7, n=3000)
    images = render_views(g, [cam], sh, 3)
    ctx = context()
    settings = diorama_amd.default_settings(1.0)
    settings.update(densify_until=0)
    say("prepared frame dropped")
    t = diorama_amd.Trainer(ctx, params, sh, [cam], images, settings)
    t.iteration = 5000
    t.sh_degree = 3
    t.step()
    r = t.renderer
    g1 = r.gaussian_buffer.download(np.float32).reshape(-1, 16)
    sh1 = r.sh_buffer.download(np.float32).reshape(-1, 45)
    g1[:, 0] += 0.05
    r.gaussian_buffer.upload(g1)
    t.step()
    got = r.splat_buffer.download(np.float32).reshape(-1, 12)
    want = fresh_splats(g1, sh1, 3, cam)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    t.destroy()


def test_step_gradients_start_from_zero():
    """The rasterizer's backward adds into the splat gradients, which the step's loss
    kernel clears: a second step's gradients equal a fresh backward of its own frame
    (loss and backward through the renderer's own calls, which clear with a fill)."""
    params, sh, g, cam = config0_scene(seed=8, n=4000)
    images = render_views(g, [cam], sh, 3)
    ctx = context()
    settings = diorama_amd.default_settings(1.0)
    settings.update(densify_until=0)
    say("gradients start from zero")
    t = diorama_amd.Trainer(ctx, params, sh, [cam], images, settings)
    t.iteration = 5000
    t.sh_degree = 3
    t.step()
    t.step()
    r = t.renderer
    got = r.gradients["splats"].download(np.float32).reshape(-1, 9)
    w, h = int(cam["viewport"][0]), int(cam["viewport"][1])
    loss = diorama_amd.TrainingLoss(ctx, w, h)
    loss.encode(r.pixel_buffer, ctx.upload(images[0]))
    r.encode_backward(loss.pixel_grads)
    want = r.gradients["splats"].download(np.float32).reshape(-1, 9)
    scale = np.abs(want).max()
    assert scale > 0
    # float atomics add in varying order; a missed clear would add the first step's gradients
    assert np.abs(got - want).max() <= 1e-4 * scale
    loss.destroy()
    t.destroy()


def reference_view_order(views, steps):
    """nextView (src/trainer.ts:395-404) with mulberry32(1): every view once, in a
    Fisher-Yates shuffled order popped from the end, then reshuffle."""
    random = scene.mulberry32(1)
    queue, out = [], []
    for _ in range(steps):
        if not queue:
            queue = list(range(views))
            for i in range(len(queue) - 1, 0, -1):
                j = int(math.floor(random() * (i + 1)))
                queue[i], queue[j] = queue[j], queue[i]
        out.append(queue.pop())
    return out


def test_schedule_view_order_sh_degree_and_opacity_reset():
    """The Trainer's schedule against src/trainer.ts:192-252: the view order of a
    mulberry32(1) Fisher-Yates replay, the SH degree one higher at steps 1000, 2000
    and 3000, and the opacity reset at step 3000 (while densifying). Around the SH
    step at 1000 every step's splats equal a fresh preprocess of the Gaussians the
    step started from, at that step's degree, for that step's view: the frame the
    previous step prepared ahead was drawn for the right view and degree. A handle
    to the renderer's splats taken before those steps still names its splats."""
    ctx = context()
    n = 600
    params, sh = scene.random_raw_gaussians(n, 3, box=1.0)
    cams = []
    for k in range(5):
        a = 2 * math.pi * k / 5
        cams.append(scene.camera(scene.look_at([3 * math.sin(a), 0.4, 3 * math.cos(a)], [0, 0, 0], [0, 1, 0]),
                                 scene.perspective(math.pi / 3, 1.0, 0.01, 100.0), 48, 48))
    teacher, teacher_sh = scene.random_raw_gaussians(n, 4, box=1.0)
    images = render_views(scene.activate(teacher), cams, teacher_sh, 3)
    settings = diorama_amd.default_settings(100.0)
    # Densify without cloning, splitting or pruning (only the cadence is under test),
    # every 300 steps so that none falls in 997-1003: a densification hands the
    # renderer new buffers, as adoptGaussians does in the reference.
    settings.update(densify_grad_threshold=1e9, min_opacity=0.0, densify_interval=300)
    say("schedule")
    t = diorama_amd.Trainer(ctx, params, sh, cams, images, settings)
    r = t.renderer
    held = None
    order = reference_view_order(len(cams), 3001)
    cap = math.log(0.01 / 0.99)
    for step in range(1, 3001):
        check = 997 <= step <= 1003
        if step == 997:
            held = r.splat_buffer
        if check:
            g0 = r.gaussian_buffer.download(np.float32).reshape(-1, 16)[:n]
            sh0 = r.sh_buffer.download(np.float32).reshape(-1, 45)[:n]
        if step == 3000:
            before = t.buffers["params"].download(np.float32).reshape(-1, 16)[:, 3]
            assert (before > cap + 0.5).any()
        t.step()
        assert t.iteration == step
        assert t.last_view == order[step - 1], f"step {step}"
        assert t.sh_degree == min(3, step // 1000), f"step {step}"
        if check:
            got = held.download(np.float32).reshape(-1, 12)[:n]
            want = fresh_splats(g0, sh0, t.sh_degree, cams[t.last_view])
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"step {step}"
    assert len(t.densifications) == len([k for k in range(501, 3001) if k % 300 == 0])
    p = t.buffers["params"].download(np.float32).reshape(-1, 16)
    m = t.buffers["moments"].download(np.float32).reshape(-1, 122)
    assert (p[:, 3] <= np.float32(cap) + 1e-6).all()
    assert not m[:, 6:8].any()
    t.destroy()
