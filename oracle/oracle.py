"""ctypes binding of the CPU oracle (liboracle.so, built from oracle.c by the Makefile here).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product package never imports it.
All arrays are float64 numpy arrays in the reference's buffer layouts; see oracle.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compiles liboracle.so in place (gcc + OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        c_long, c_int, vp = ctypes.c_long, ctypes.c_int, ctypes.c_void_p
        L.oc_preprocess.argtypes = [c_long, _d, vp, c_int, _d, _d, _d, _d]
        L.oc_tile_lists.argtypes = [c_long, _d, c_int, c_int, _u32, vp, c_long]
        L.oc_tile_lists.restype = c_long
        L.oc_tile_lists_capped.argtypes = [c_long, _d, c_int, c_int, c_long, _u32, _u32]
        L.oc_tile_lists_capped.restype = c_long
        L.oc_rasterize.argtypes = [_d, _u32, _u32, c_int, c_int, _d, _d, _u32]
        L.oc_rasterize_margin.argtypes = [_d, _u32, _u32, c_int, c_int, _d, _d, _u32, _d]
        L.oc_rasterize_backward.argtypes = [c_long, _d, _u32, _u32, c_int, c_int, _d, _d, _u32, _d, _d]
        L.oc_footprint.argtypes = [c_long, _d, _u32, _u32, c_int, c_int, _d, _d]
        L.oc_preprocess_backward.argtypes = [c_long, _d, _d, c_int, _d, _d, _d, _d, _d, _d, _d]
        L.oc_loss.argtypes = [c_int, c_int, _d, _u8, vp, vp]
        L.oc_loss.restype = ctypes.c_double
        L.oc_ssim.argtypes = [c_int, c_int, _d, _d]
        L.oc_ssim.restype = ctypes.c_double
        L.oc_activate.argtypes = [c_long, _d, _d]
        L.oc_accumulate_stats.argtypes = [c_long, _d, _d, _d, _d]
        L.oc_adam_step.argtypes = [c_long, _d, _d, _d, _d, _d, _d, vp]
        L.oc_reset_opacity.argtypes = [c_long, _d, _d, vp]
        L.oc_pcg.argtypes = [ctypes.c_uint32]
        L.oc_pcg.restype = ctypes.c_uint32
        L.oc_densify.argtypes = [c_long, _d, _d, _d, _d, _d, ctypes.c_uint32, _d, _d, _d, _u32]
        L.oc_densify.restype = c_long
        L.oc_train_step.argtypes = [c_long, _d, _d, _d, _d, _d, c_int, _d, _d, c_int, c_int, _d, _u8, _d,
                                    c_int, ctypes.POINTER(c_long)]
        L.oc_train_step.restype = ctypes.c_double
        L.oc_threads.restype = c_int
        _f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.oc_mean_neighbor_distance2.argtypes = [c_long, _f32, c_long, vp, _f32]
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def camera_arrays(camera):
    return _f64(camera["view"]), _f64(camera["proj"]), _f64(camera["viewport"])


def preprocess(gaussians, camera, sh=None, degree=0):
    g = _f64(gaussians).reshape(-1)
    n = g.size // 16
    view, proj, vp = camera_arrays(camera)
    out = np.zeros(n * 12)
    shp = None if sh is None else _f64(sh).reshape(-1)
    lib().oc_preprocess(n, g, _ptr(shp), int(degree), view, proj, vp, out)
    return out.reshape(n, 12)


def tile_lists(splats, width, height):
    """(ranges[T, 2] uint32, sorted pair splats uint32)."""
    s = _f64(splats).reshape(-1)
    n = s.size // 12
    tiles = ((width + 15) // 16) * ((height + 15) // 16)
    ranges = np.zeros(2 * tiles, dtype=np.uint32)
    total = lib().oc_tile_lists(n, s, width, height, ranges, None, 0)
    pairs = np.zeros(max(total, 1), dtype=np.uint32)
    lib().oc_tile_lists(n, s, width, height, ranges, _ptr(pairs), total)
    return ranges.reshape(tiles, 2), pairs[:total]


def tile_lists_capped(splats, width, height, capacity):
    """(ranges[T, 2] uint32, kept pair splats uint32, full pair count) of a frame whose
    pair buffers hold `capacity` pairs: the nearest pairs are kept (oc_tile_lists_capped)."""
    s = _f64(splats).reshape(-1)
    n = s.size // 12
    tiles = ((width + 15) // 16) * ((height + 15) // 16)
    ranges = np.zeros(2 * tiles, dtype=np.uint32)
    pairs = np.zeros(max(capacity, 1), dtype=np.uint32)
    total = lib().oc_tile_lists_capped(n, s, width, height, capacity, ranges, pairs)
    return ranges.reshape(tiles, 2), pairs[:min(total, capacity)], total


def rasterize(splats, ranges, pairs, width, height, background, margin=False):
    """(pixel_color[X,4], pixel_last[X]) and, with margin=True, each pixel's distance
    to the nearest blend discontinuity (see oc_rasterize_margin)."""
    s = _f64(splats).reshape(-1)
    color = np.zeros(width * height * 4)
    last = np.zeros(width * height, dtype=np.uint32)
    near = np.zeros(width * height)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32)
    if pr.size == 0:
        pr = np.zeros(1, dtype=np.uint32)
    lib().oc_rasterize_margin(s, np.ascontiguousarray(ranges, dtype=np.uint32).reshape(-1), pr, width, height,
                              _f64(background), color, last, near)
    if margin:
        return color.reshape(-1, 4), last, near
    return color.reshape(-1, 4), last


def footprint(splats, ranges, pairs, width, height, pixel_loss):
    """Per splat (sum of w x pixel_loss, sum of w) over the pixels it blends into, w
    its blend weight alpha T (oc_footprint)."""
    s = _f64(splats).reshape(-1)
    n = s.size // 12
    out = np.zeros(n * 2)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32)
    if pr.size == 0:
        pr = np.zeros(1, dtype=np.uint32)
    lib().oc_footprint(n, s, np.ascontiguousarray(ranges, dtype=np.uint32).reshape(-1), pr, width, height,
                       _f64(pixel_loss).reshape(-1), out)
    return out.reshape(n, 2)


def rasterize_backward(splats, ranges, pairs, width, height, background, pixel_color, pixel_last, pixel_grads):
    s = _f64(splats).reshape(-1)
    n = s.size // 12
    out = np.zeros(n * 9)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32)
    if pr.size == 0:
        pr = np.zeros(1, dtype=np.uint32)
    lib().oc_rasterize_backward(n, s, np.ascontiguousarray(ranges, dtype=np.uint32).reshape(-1), pr, width, height,
                                _f64(background), _f64(pixel_color).reshape(-1),
                                np.ascontiguousarray(pixel_last, dtype=np.uint32), _f64(pixel_grads).reshape(-1), out)
    return out.reshape(n, 9)


def preprocess_backward(gaussians, sh, degree, camera, splats, splat_grads):
    g = _f64(gaussians).reshape(-1)
    n = g.size // 16
    view, proj, vp = camera_arrays(camera)
    gg = np.zeros(n * 16)
    sg = np.zeros(n * 45)
    shp = np.zeros(n * 45) if sh is None else _f64(sh).reshape(-1)
    lib().oc_preprocess_backward(n, g, shp, int(degree), view, proj, vp, _f64(splats).reshape(-1),
                                 _f64(splat_grads).reshape(-1), gg, sg)
    return gg.reshape(n, 16), sg.reshape(n, 45)


def render(gaussians, camera, background, sh=None, degree=0):
    """Forward render: (splats, ranges, pairs, pixel_color[X,4], pixel_last)."""
    w, h = int(camera["viewport"][0]), int(camera["viewport"][1])
    splats = preprocess(gaussians, camera, sh, degree)
    ranges, pairs = tile_lists(splats, w, h)
    color, last = rasterize(splats, ranges, pairs, w, h, background)
    return splats, ranges, pairs, color, last


def backward(gaussians, camera, background, pixel_grads, sh=None, degree=0):
    """Gradients of sum(pixel_grads * image) w.r.t. the activated Gaussians and SH."""
    w, h = int(camera["viewport"][0]), int(camera["viewport"][1])
    splats, ranges, pairs, color, last = render(gaussians, camera, background, sh, degree)
    sgrad = rasterize_backward(splats, ranges, pairs, w, h, background, color, last, pixel_grads)
    gg, shg = preprocess_backward(gaussians, sh, degree, camera, splats, sgrad)
    return gg, shg, sgrad


def loss(width, height, rendered, target_u8):
    """(loss, pixel_grads[X,4], pixel_loss[X])."""
    n = width * height
    grads = np.zeros(n * 4)
    shares = np.zeros(n)
    value = lib().oc_loss(width, height, _f64(rendered).reshape(-1),
                          np.ascontiguousarray(target_u8, dtype=np.uint8).reshape(-1), _ptr(grads), _ptr(shares))
    return value, grads.reshape(n, 4), shares


def ssim(width, height, a, b):
    return lib().oc_ssim(width, height, _f64(a).reshape(-1), _f64(b).reshape(-1))


def activate(params):
    p = _f64(params).reshape(-1)
    out = np.zeros_like(p)
    lib().oc_activate(p.size // 16, p, out)
    return out.reshape(-1, 16)


def accumulate_stats(splats, splat_grads, viewport, stats):
    s = _f64(stats).reshape(-1).copy()
    n = s.size // 4
    lib().oc_accumulate_stats(n, _f64(splats).reshape(-1), _f64(splat_grads).reshape(-1), _f64(viewport), s)
    return s.reshape(n, 4)


def adam_step(params, sh, moments, gaussian_grads, sh_grads, lr):
    """Returns (params, sh, moments, activated gaussians) after one update."""
    p = _f64(params).reshape(-1).copy()
    s = _f64(sh).reshape(-1).copy()
    m = _f64(moments).reshape(-1).copy()
    n = p.size // 16
    g = np.zeros(n * 16)
    lib().oc_adam_step(n, p, s, m, _f64(gaussian_grads).reshape(-1), _f64(sh_grads).reshape(-1), _f64(lr), _ptr(g))
    return p.reshape(n, 16), s.reshape(n, 45), m.reshape(n, 122), g.reshape(n, 16)


def reset_opacity(params, moments):
    p = _f64(params).reshape(-1).copy()
    m = _f64(moments).reshape(-1).copy()
    n = p.size // 16
    g = np.zeros(n * 16)
    lib().oc_reset_opacity(n, p, m, _ptr(g))
    return p.reshape(n, 16), m.reshape(n, 122), g.reshape(n, 16)


def pcg(v):
    return int(lib().oc_pcg(ctypes.c_uint32(v & 0xFFFFFFFF)))


def densify(params, sh, moments, stats, settings, seed):
    """(count, new params, new sh, new moments, tally[drop, keep, clone, split])."""
    p = _f64(params).reshape(-1)
    n = p.size // 16
    np_ = np.zeros(2 * n * 16 + 16)
    nsh = np.zeros(2 * n * 45 + 45)
    nm = np.zeros(2 * n * 122 + 122)
    tally = np.zeros(4, dtype=np.uint32)
    count = lib().oc_densify(n, p, _f64(sh).reshape(-1), _f64(moments).reshape(-1), _f64(stats).reshape(-1),
                             _f64(settings), seed, np_, nsh, nm, tally)
    return (count, np_[:count * 16].reshape(count, 16), nsh[:count * 45].reshape(count, 45),
            nm[:count * 122].reshape(count, 122), tally)


def train_step(params, sh, moments, stats, gaussians, degree, camera, background, target_u8, lr, adam=True):
    """One optimize() step in place on float64 arrays; returns (loss, pair count)."""
    view, proj, vp = camera_arrays(camera)
    n = params.size // 16
    pc = ctypes.c_long(0)
    loss_value = lib().oc_train_step(n, params.reshape(-1), sh.reshape(-1), moments.reshape(-1), stats.reshape(-1),
                                     gaussians.reshape(-1), int(degree), view, proj, int(vp[0]), int(vp[1]),
                                     _f64(background), np.ascontiguousarray(target_u8, dtype=np.uint8).reshape(-1),
                                     _f64(lr), 1 if adam else 0, ctypes.byref(pc))
    return loss_value, pc.value


def threads():
    return lib().oc_threads()


def mean_neighbor_distance2(points, queries=None):
    """Brute-force float32 mean squared distance to the 3 nearest other points
    (src/shaders/nearestNeighbors.wgsl:12-44), for every point or for `queries`."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    n = pts.size // 3
    if queries is None:
        out = np.empty(n, np.float32)
        lib().oc_mean_neighbor_distance2(n, pts, n, None, out)
        return out
    q = np.ascontiguousarray(queries, dtype=np.int64)
    out = np.empty(q.size, np.float32)
    lib().oc_mean_neighbor_distance2(n, pts, q.size, q.ctypes.data_as(ctypes.c_void_p), out)
    return out


def initial_gaussians(points, colors, distance2):
    """src/loadDataset.ts:93-110: raw params (n x 16 float32) and zero SH (n x 45),
    the arithmetic in float64 rounded once."""
    p = np.asarray(points, np.float32).reshape(-1, 3)
    c = np.asarray(colors, np.uint8).reshape(-1, 3)
    d2 = np.asarray(distance2, np.float32).astype(np.float64)
    n = len(p)
    params = np.zeros((n, 16), np.float32)
    params[:, 0:3] = p
    params[:, 3] = np.float32(np.log(0.1 / 0.9))
    with np.errstate(over="ignore", invalid="ignore"):
        params[:, 4:7] = (0.5 * np.log(np.maximum(d2, 1e-7)))[:, None].astype(np.float32)
    params[:, 8] = 1.0
    params[:, 12:15] = ((c / 255.0 - 0.5) / 0.28209479177387814).astype(np.float32)
    return params, np.zeros((n, 45), np.float32)
