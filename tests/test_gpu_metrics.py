"""Held-out PSNR and SSIM scored on the device (metrics.hip behind Trainer.evaluate
and Trainer.measure_psnr) against the reference's CPU metrics (src/metrics.ts:5-72,
src/trainer.ts:371-386): renders clamped and rounded to 8 bits, PSNR over the rgb
channels in float32 values, SSIM by the float64 oracle (oc_ssim). The device sums
the 8-bit errors exactly and takes k / 255 in double where the reference holds
float32(k / 255), so both metrics differ from the formula as written only by that
rounding: bars 1e-5 dB for PSNR and 1e-7 for SSIM."""
import math

import numpy as np
import pytest

import diorama_amd
from oracle import oracle as oc
from tests.gpu_util import config0_scene, context, say
from tests.test_gpu_batch import orbit_cameras

pytestmark = pytest.mark.gpu


def reference_metrics(pixels, photo_u32, width, height):
    """src/trainer.ts:378-382 and src/metrics.ts:5-15 as written: Float32Array values."""
    x = pixels.reshape(-1, 4).astype(np.float64)
    rendered = (np.floor(np.clip(x, 0, 1) * 255 + 0.5) / 255).astype(np.float32)   # Math.round: half up
    photo = (photo_u32.view(np.uint8).reshape(-1, 4) / 255).astype(np.float32)
    d = rendered[:, :3].astype(np.float64) - photo[:, :3].astype(np.float64)
    se = float((d * d).sum())
    psnr = 10 * math.log10(1 / (se / (3 * width * height))) if se > 0 else math.inf
    ssim = oc.ssim(width, height, rendered.astype(np.float64), photo.astype(np.float64))
    return psnr, ssim


def trainer_and_views(sizes, seed):
    params, sh, _, _ = config0_scene(seed=seed, n=3000)
    cams = orbit_cameras(sizes, radius=3.5)
    rng = np.random.default_rng(seed)
    images = [rng.integers(0, 2**32, w * h, dtype=np.uint32) for w, h in sizes]
    ctx = context()
    # trained on the first view (training views share one size); held-out views of any size
    t = diorama_amd.Trainer(ctx, params, sh, cams[:1], images[:1], diorama_amd.default_settings(1.0))
    t.sh_degree = 3
    return ctx, t, cams, images


def render_pixels(t, ctx, cam):
    w, h = (int(v) for v in cam["viewport"])
    r = t.renderer
    out = ctx.buffer(w * h * 4)
    r.encode(out, cam, [0, 0, 0])
    return r.pixel_buffer.download(np.float32, w * h * 4)


def test_evaluate_matches_the_reference_metrics():
    sizes = [(160, 120), (37, 23), (200, 150)]
    ctx, t, cams, images = trainer_and_views(sizes, 3)
    for _ in range(5):
        t.step()
    say("evaluate on device")
    got = t.evaluate(cams, images, with_ssim=True)
    want_p, want_s = [], []
    for cam, img, (w, h) in zip(cams, images, sizes):
        p, s = reference_metrics(render_pixels(t, ctx, cam), img, w, h)
        want_p.append(p)
        want_s.append(s)
    print(f"evaluate: psnr {got['psnr']:.9f} vs {np.mean(want_p):.9f}, ssim {got['ssim']:.12f} vs {np.mean(want_s):.12f}")
    assert abs(got["psnr"] - np.mean(want_p)) < 1e-5
    assert abs(got["ssim"] - np.mean(want_s)) < 1e-7
    assert math.isnan(t.evaluate(cams, images, with_ssim=False)["ssim"])
    t.destroy()


def test_perfect_photo_and_device_resident_photos():
    sizes = [(64, 48), (48, 64)]
    ctx, t, cams, images = trainer_and_views(sizes, 4)
    say("metrics on its own render")
    photos = []
    for cam, (w, h) in zip(cams, sizes):
        q = np.floor(np.clip(render_pixels(t, ctx, cam).reshape(-1, 4), 0, 1) * 255 + 0.5).astype(np.uint32)
        photos.append(q[:, 0] | (q[:, 1] << 8) | (q[:, 2] << 16) | (255 << 24))
    perfect = t.evaluate(cams, photos, with_ssim=True)
    assert perfect["psnr"] == math.inf and abs(perfect["ssim"] - 1.0) < 1e-12
    on_host = t.evaluate(cams, images, with_ssim=False)["psnr"]
    buffers = [ctx.upload(img) for img in images]
    say("measure_psnr")
    assert t.measure_psnr(cams, buffers) == on_host
    t.destroy()
