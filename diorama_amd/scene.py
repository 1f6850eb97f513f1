"""Host-side scene helpers: camera matrices, the seeded RNG, buffer packing and the
synthetic scenes the tests and the benchmark run on.

Camera math follows src/mat4.ts (column-major 4x4, WebGPU 0..1 clip depth):
`perspective` src/mat4.ts:39-48, `look_at` src/mat4.ts:51-61. The RNG is
mulberry32 (src/random.ts:3-12). Layouts are those of src/gaussians.ts:32-70.
"""
import math

import numpy as np

FLOATS_PER_GAUSSIAN = 16
FLOATS_PER_SPLAT = 12
SH_REST_FLOATS = 45
SH_C0 = 0.28209479177387814


def mulberry32(seed):
    """The reference's seeded generator: returns a function giving floats in [0, 1)."""
    state = [seed & 0xFFFFFFFF]

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def next_value():
        state[0] = (state[0] + 0x6D2B79F5) & 0xFFFFFFFF
        t = state[0]
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return next_value


def perspective(fov_y, aspect, near, far):
    f = 1.0 / math.tan(fov_y / 2)
    depth = 1.0 / (near - far)
    return [f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, far * depth, -1, 0, 0, near * far * depth, 0]


def _normalize(v):
    length = math.sqrt(sum(c * c for c in v))
    return [c / length for c in v]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def look_at(eye, target, up):
    z = _normalize([eye[k] - target[k] for k in range(3)])
    x = _normalize(_cross(up, z))
    y = _cross(z, x)
    dot = lambda a, b: sum(a[k] * b[k] for k in range(3))
    return [x[0], y[0], z[0], 0, x[1], y[1], z[1], 0, x[2], y[2], z[2], 0,
            -dot(x, eye), -dot(y, eye), -dot(z, eye), 1]


IDENTITY = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]


def camera(view, proj, width, height):
    return {"view": [float(v) for v in view], "proj": [float(v) for v in proj], "viewport": [float(width), float(height)]}


def pack_gaussians(position, scale, rotation, color, opacity):
    """Activated Gaussians as (N, 16) float32, laid out like packGaussians() (src/gaussians.ts:38-46)."""
    n = len(position)
    out = np.zeros((n, 16), dtype=np.float32)
    out[:, 0:3] = position
    out[:, 3] = opacity
    out[:, 4:7] = scale
    out[:, 8:12] = rotation
    out[:, 12:15] = color
    return out


def activate(params):
    """Raw parameters (N, 16) to activated Gaussians (src/trainer.ts:449-461), float32."""
    p = np.asarray(params, dtype=np.float32).reshape(-1, 16)
    out = np.zeros_like(p)
    out[:, 0:3] = p[:, 0:3]
    out[:, 4:7] = np.exp(p[:, 4:7])
    out[:, 12:15] = np.float32(0.5) + np.float32(SH_C0) * p[:, 12:15]
    out[:, 3] = 1.0 / (1.0 + np.exp(-p[:, 3]))
    out[:, 8:12] = p[:, 8:12]
    return out


def random_raw_gaussians(n, seed, box=1.0):
    """Raw training parameters of the synthetic scenes: means uniform in [-box, box]^3,
    log-scales N(-4, 0.5), normalised N(0,1) quaternions, opacity logits N(0,1), the
    SH DC colour N(0.5, 0.3) as an activated colour (so raw = (c - 0.5) / SH_C0), and
    the degree 1-3 bands N(0, 0.05). Returns (params (n,16), sh (n,45)) float32."""
    rng = np.random.default_rng(seed)
    params = np.zeros((n, 16), dtype=np.float32)
    params[:, 0:3] = rng.uniform(-box, box, size=(n, 3))
    params[:, 3] = rng.normal(0.0, 1.0, size=n)
    params[:, 4:7] = rng.normal(-4.0, 0.5, size=(n, 3))
    q = rng.normal(0.0, 1.0, size=(n, 4))
    params[:, 8:12] = q / np.linalg.norm(q, axis=1, keepdims=True)
    color = rng.normal(0.5, 0.3, size=(n, 3))
    params[:, 12:15] = (color - 0.5) / SH_C0
    sh = rng.normal(0.0, 0.05, size=(n, 45)).astype(np.float32)
    return params, sh


def small_config_camera():
    """The 256x256, 60-degree camera at distance 4 looking at the origin."""
    view = look_at([0.0, 0.0, 4.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    proj = perspective(math.radians(60), 1.0, 0.01, 100.0)
    return camera(view, proj, 256, 256)


def ring_cameras(count, width=980, height=545, fov_deg=50.0, radius=6.0, seed=3):
    """Poses on a jittered ring (radius 6 +- 0.5, height N(0, 0.3)) looking at the origin."""
    rng = np.random.default_rng(seed)
    cams = []
    for k in range(count):
        angle = 2 * math.pi * k / count
        r = radius + rng.uniform(-0.5, 0.5)
        eye = [r * math.cos(angle), float(rng.normal(0.0, 0.3)), r * math.sin(angle)]
        view = look_at(eye, [0.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        proj = perspective(math.radians(fov_deg), width / height, 0.01, 100.0)
        cams.append(camera(view, proj, width, height))
    return cams


def split_views(cams, held_out=33, train=263):
    """The reference's split (src/loadDataset.ts:37-38): every 8th view held out."""
    held = [k for k in range(len(cams)) if k % 8 == 7][:held_out]
    skip = set(held)
    return [c for k, c in enumerate(cams) if k not in skip][:train], [cams[k] for k in held]


def teacher_images(ctx, cams, n, seed=7, box=5.0, noise=0.01, noise_seed=7):
    """rgba8 targets (u32 per pixel): a `seed` teacher of n Gaussians rendered on the
    GPU for each camera, plus N(0, noise) pixel noise, rounded to 8 bits."""
    import diorama_amd

    params, sh = random_raw_gaussians(n, seed, box=box)
    r = diorama_amd.GaussianRenderer(ctx)
    r.set_gaussians(activate(params), sh, 3)
    rng = np.random.default_rng(noise_seed)
    images = []
    for cam in cams:
        r.encode(None, cam, [0, 0, 0])
        px = int(cam["viewport"][0]) * int(cam["viewport"][1])
        color = r.pixel_buffer.download(np.float32, px * 4).reshape(-1, 4)[:, :3]
        if noise:
            color = color + rng.standard_normal(color.shape, dtype=np.float32) * np.float32(noise)
        q = np.rint(np.clip(color, 0, 1) * 255).astype(np.uint32)
        images.append(q[:, 0] | (q[:, 1] << 8) | (q[:, 2] << 16) | np.uint32(255 << 24))
    r.after_submit()
    r.destroy()
    return images


def rgba8_channels(image):
    """(pixels, 4) uint8 channels of u32 rgba8 words."""
    out = np.zeros((len(image), 4), dtype=np.uint8)
    for c in range(4):
        out[:, c] = (np.asarray(image, dtype=np.uint32) >> (8 * c)) & 255
    return out


def eye_of(cam):
    """Camera centre of a view matrix [R | t]: -R^T t."""
    v = cam["view"]
    return [-(v[c * 4] * v[12] + v[c * 4 + 1] * v[13] + v[c * 4 + 2] * v[14]) for c in range(3)]


def scene_extent(cams):
    """1.1 x the largest camera distance from the origin (src/loadDataset.ts:39-43)."""
    return 1.1 * max(math.sqrt(sum(v * v for v in eye_of(c))) for c in cams)


# Scan/sort microbenchmark inputs (shared bit-for-bit by the tests and the bench).

def tile_touch_counts(n, seed):
    """u32 tile-touch counts ~ Geometric(mean 6) capped at 256."""
    rng = np.random.default_rng(seed)
    return np.minimum(rng.geometric(1 / 6, size=n), 256).astype(np.uint32)


def zipf_tiles(n, seed, tiles=62 * 35, s=1.1):
    """Tile ids ~ Zipf(s) over `tiles` (the 62x35 tiles of 980x545), ranks shuffled."""
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, tiles + 1) ** s
    p /= p.sum()
    perm = rng.permutation(tiles)
    return perm[rng.choice(tiles, size=n, p=p)].astype(np.uint32)


def tile_depth_keys(n, seed):
    """64-bit (low = float32 bits of depth ~ U[0.2, 50], high = Zipf tile) keys as
    interleaved u32 words, plus the two words separately."""
    rng = np.random.default_rng(seed)
    hi = zipf_tiles(n, seed)
    lo = rng.uniform(0.2, 50.0, size=n).astype(np.float32).view(np.uint32)
    keys = np.empty(2 * n, dtype=np.uint32)
    keys[0::2] = lo
    keys[1::2] = hi
    return keys, lo, hi
