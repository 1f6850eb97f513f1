"""Python host mirror of the reference's operator classes over the C ABI
(include/diorama.h, built as libdiorama_hip.so next to this file).

Classes keep the names and methods of src/prefixSum.ts, src/radixSort.ts,
src/tileRasterizer.ts, src/gaussianRenderer.ts, src/trainingLoss.ts and
src/trainer.ts; buffers are `Buffer` objects wrapping device allocations. The HIP
library is required: importing works without it, but constructing a context
raises if the library or a GPU is missing (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiorama_hip.so")

_lib = None


# Status codes of include/diorama.h.
DIO_OK, DIO_ERR_INVALID, DIO_ERR_INTERNAL, DIO_ERR_HIP, DIO_ERR_MEMORY = 0, -1, -2, -3, -4


class DioramaError(RuntimeError):
    pass


class _Camera(ctypes.Structure):
    _fields_ = [("view", ctypes.c_float * 16), ("proj", ctypes.c_float * 16), ("viewport", ctypes.c_float * 2)]


class _Tuning(ctypes.Structure):
    _fields_ = [("tile_size", ctypes.c_uint32), ("splat_batch", ctypes.c_uint32), ("digit_bits", ctypes.c_uint32),
                ("persistent", ctypes.c_uint32), ("segment", ctypes.c_uint32),
                ("per_tile_depth", ctypes.c_uint32)]


class _TuningResult(ctypes.Structure):
    _fields_ = [("tuning", _Tuning), ("ms_per_step", ctypes.c_double)]


class _ColmapCamera(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("fx", ctypes.c_double), ("fy", ctypes.c_double), ("position", ctypes.c_double * 3),
                ("right", ctypes.c_double * 3), ("down", ctypes.c_double * 3), ("forward", ctypes.c_double * 3)]


class _Settings(ctypes.Structure):
    _fields_ = [
        ("extent", ctypes.c_float),
        ("position_lr_steps", ctypes.c_uint32),
        ("position_lr", ctypes.c_float * 2),
        ("opacity_lr", ctypes.c_float),
        ("scale_lr", ctypes.c_float),
        ("rotation_lr", ctypes.c_float),
        ("color_lr", ctypes.c_float),
        ("sh_lr", ctypes.c_float),
        ("sh_degree_interval", ctypes.c_uint32),
        ("densify_from", ctypes.c_uint32),
        ("densify_until", ctypes.c_uint32),
        ("densify_interval", ctypes.c_uint32),
        ("densify_grad_threshold", ctypes.c_float),
        ("percent_dense", ctypes.c_float),
        ("min_opacity", ctypes.c_float),
        ("opacity_reset_interval", ctypes.c_uint32),
        ("background", ctypes.c_float * 3),
    ]


class _Densification(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("step", "before", "after", "kept", "cloned", "split", "removed")]


# Every exported symbol of include/diorama.h with (restype, argtypes).
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_F3 = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "dio_last_error": (ctypes.c_char_p, []),
    "dio_version": (ctypes.c_char_p, []),
    "dio_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "dio_ctx_synchronize": (_I, [_P]),
    "dio_ctx_destroy": (None, [_P]),
    "dio_ctx_timestamp": (_I, [_P, _U32]),
    "dio_ctx_elapsed": (_I, [_P, _U32, _U32, ctypes.POINTER(ctypes.c_double)]),
    "dio_buf_create": (_I, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "dio_buf_upload": (_I, [_P, ctypes.c_size_t, _P, ctypes.c_size_t]),
    "dio_buf_download": (_I, [_P, ctypes.c_size_t, _P, ctypes.c_size_t]),
    "dio_buf_clear": (_I, [_P]),
    "dio_buf_copy": (_I, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.c_size_t]),
    "dio_buf_size": (ctypes.c_size_t, [_P]),
    "dio_buf_destroy": (None, [_P]),
    "dio_prefix_sum_create": (_I, [_P, _P, _U32, _P, ctypes.POINTER(_P)]),
    "dio_prefix_sum_total": (_P, [_P]),
    "dio_prefix_sum_encode": (_I, [_P]),
    "dio_prefix_sum_destroy": (None, [_P]),
    "dio_prefix_sum_encode_static": (_I, [_P, _P, _U32, _P, _P]),
    "dio_radix_sort_create": (_I, [_P, _U32, _U32, _P, ctypes.POINTER(_P)]),
    "dio_radix_sort_keys": (_P, [_P]),
    "dio_radix_sort_values": (_P, [_P]),
    "dio_radix_sort_capacity": (_U32, [_P]),
    "dio_radix_sort_encode": (_I, [_P]),
    "dio_radix_sort_destroy": (None, [_P]),
    "dio_tile_rasterizer_create": (_I, [_P, _P, _U32, _U32, ctypes.POINTER(_P)]),
    "dio_tile_rasterizer_encode": (_I, [_P, _P, _U32, _U32, _F3]),
    "dio_tile_rasterizer_encode_backward": (_I, [_P, _U32, _U32, _P, _P]),
    "dio_tile_rasterizer_after_submit": (_I, [_P]),
    "dio_tile_rasterizer_last_pair_count": (_U32, [_P]),
    "dio_tile_rasterizer_pair_capacity": (_U32, [_P]),
    "dio_tile_rasterizer_set_tuning": (_I, [_P, ctypes.POINTER(_Tuning)]),
    "dio_tile_rasterizer_pair_splats": (_P, [_P]),
    "dio_tile_rasterizer_ranges": (_P, [_P]),
    "dio_tile_rasterizer_pixel_color": (_P, [_P]),
    "dio_tile_rasterizer_pixel_last": (_P, [_P]),
    "dio_tile_rasterizer_destroy": (None, [_P]),
    "dio_renderer_create": (_I, [_P, _U32, ctypes.POINTER(_P)]),
    "dio_renderer_set_gaussians": (_I, [_P, _P, _U32, _P, _U32]),
    "dio_renderer_adopt_gaussians": (_I, [_P, _U32, _P, _P, _U32]),
    "dio_renderer_set_sh_degree": (_I, [_P, _U32]),
    "dio_renderer_encode": (_I, [_P, _P, ctypes.POINTER(_Camera), _F3]),
    "dio_renderer_encode_preprocess": (_I, [_P, ctypes.POINTER(_Camera)]),
    "dio_renderer_encode_tiles": (_I, [_P, _P, _F3]),
    "dio_renderer_encode_draw": (_I, [_P, _P, _P, _U32, _U32, _U32, _F3]),
    "dio_renderer_encode_backward": (_I, [_P, _P]),
    "dio_renderer_after_submit": (_I, [_P]),
    "dio_renderer_count": (_U32, [_P]),
    "dio_renderer_splat_buffer": (_P, [_P]),
    "dio_renderer_gaussian_buffer": (_P, [_P]),
    "dio_renderer_sh_buffer": (_P, [_P]),
    "dio_renderer_tiles": (_P, [_P]),
    "dio_renderer_pixel_buffer": (_P, [_P]),
    "dio_renderer_pixel_last_buffer": (_P, [_P]),
    "dio_renderer_gradient_gaussians": (_P, [_P]),
    "dio_renderer_gradient_sh": (_P, [_P]),
    "dio_renderer_gradient_splats": (_P, [_P]),
    "dio_renderer_destroy": (None, [_P]),
    "dio_training_loss_create": (_I, [_P, _U32, _U32, ctypes.POINTER(_P)]),
    "dio_training_loss_encode": (_I, [_P, _P, _P]),
    "dio_training_loss_read": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "dio_training_loss_pixel_grads": (_P, [_P]),
    "dio_training_loss_pixel_loss": (_P, [_P]),
    "dio_training_loss_destroy": (None, [_P]),
    "dio_default_settings": (None, [ctypes.c_float, ctypes.POINTER(_Settings)]),
    "dio_trainer_create": (_I, [_P, _P, _P, _U32, ctypes.POINTER(_Camera), ctypes.POINTER(_P), _U32,
                                ctypes.POINTER(_Settings), ctypes.POINTER(_P)]),
    "dio_trainer_create_batched": (_I, [_P, _P, _P, _U32, ctypes.POINTER(_Camera), ctypes.POINTER(_P), _U32, _U32,
                                        ctypes.POINTER(_Settings), ctypes.POINTER(_P)]),
    "dio_trainer_batch_size": (_U32, [_P]),
    "dio_mlp_create": (_I, [_P, _U32, ctypes.POINTER(_P)]),
    "dio_mlp_get_weights": (_I, [_P, ctypes.POINTER(ctypes.c_float)]),
    "dio_mlp_set_weights": (_I, [_P, ctypes.POINTER(ctypes.c_float)]),
    "dio_mlp_forward": (_I, [_P, _P, _U32, _P]),
    "dio_mlp_gradients": (_I, [_P, _P, _P, _U32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]),
    "dio_mlp_train": (_I, [_P, _P, _P, _U32, _U32, _U32, ctypes.c_float, _U32, ctypes.POINTER(ctypes.c_double)]),
    "dio_mlp_set_class_weights": (_I, [_P, ctypes.POINTER(ctypes.c_float)]),
    "dio_mlp_set_input_norm": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "dio_mlp_destroy": (None, [_P]),
    "dio_trainer_set_densify_mode": (_I, [_P, _U32, _P, _U32, ctypes.c_float, _U32]),
    "dio_trainer_set_tuning": (_I, [_P, ctypes.POINTER(_Tuning)]),
    "dio_trainer_get_tuning": (_I, [_P, ctypes.POINTER(_Tuning)]),
    "dio_trainer_autotune": (_I, [_P, _U32, ctypes.POINTER(_TuningResult), _U32, ctypes.POINTER(_U32)]),
    "dio_trainer_samples": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_U32)]),
    "dio_trainer_features": (_I, [_P, ctypes.POINTER(_P)]),
    "dio_trainer_extra_stats": (_P, [_P]),
    "dio_trainer_footprint": (_P, [_P]),
    "dio_trainer_batch_frames": (_I, [_P, ctypes.POINTER(_U32), _U32, ctypes.POINTER(_U32), ctypes.POINTER(_P),
                                     ctypes.POINTER(_U32), ctypes.POINTER(_P)]),
    "dio_trainer_step": (_I, [_P]),
    "dio_trainer_optimize": (_I, [_P]),
    "dio_trainer_maintain": (_I, [_P]),
    "dio_trainer_densify": (_I, [_P, _I]),
    "dio_trainer_reset_opacity": (_I, [_P]),
    "dio_trainer_export_raw": (_I, [_P, _P, _P]),
    "dio_trainer_last_loss": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "dio_trainer_evaluate": (_I, [_P, ctypes.POINTER(_Camera), ctypes.POINTER(_P), _U32, _I,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "dio_trainer_count": (_U32, [_P]),
    "dio_trainer_iteration": (_U32, [_P]),
    "dio_trainer_set_iteration": (None, [_P, _U32]),
    "dio_trainer_sh_degree": (_U32, [_P]),
    "dio_trainer_last_view": (_U32, [_P]),
    "dio_trainer_set_sh_degree": (_I, [_P, _U32]),
    "dio_trainer_densification_count": (_U32, [_P]),
    "dio_trainer_densification": (_I, [_P, _U32, ctypes.POINTER(_Densification)]),
    "dio_trainer_params": (_P, [_P]),
    "dio_trainer_moments": (_P, [_P]),
    "dio_trainer_stats": (_P, [_P]),
    "dio_trainer_renderer": (_P, [_P]),
    "dio_trainer_destroy": (None, [_P]),
    "dio_trainer_time_stage": (_I, [_P, ctypes.c_char_p, _U32, ctypes.POINTER(ctypes.c_double)]),
    "dio_trainer_stage_clock": (_I, [_P, ctypes.c_int]),
    "dio_trainer_stage_totals": (_I, [_P, ctypes.POINTER(ctypes.c_double), _U32, ctypes.POINTER(_U32),
                                      ctypes.POINTER(_U32), ctypes.POINTER(ctypes.c_uint64)]),
    "dio_trainer_stage_name": (ctypes.c_char_p, [_U32]),
    "dio_trainer_measure_psnr": (_I, [_P, ctypes.POINTER(_Camera), ctypes.POINTER(_P), _U32,
                                     ctypes.POINTER(ctypes.c_double)]),
    "dio_colmap_parse": (_I, [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.c_size_t,
                              ctypes.POINTER(_ColmapCamera), _U32, ctypes.POINTER(_U32), _P, _P, _U32,
                              ctypes.POINTER(_U32)]),
    "dio_colmap_training_camera": (_I, [ctypes.POINTER(_ColmapCamera), _U32, _U32, ctypes.POINTER(_Camera)]),
    "dio_mean_neighbor_distance2": (_I, [_P, _P, _U32, _P]),
    "dio_initial_gaussians": (_I, [_P, _P, _P, _U32, _P, _P]),
    "dio_ply_write": (_I, [_P, _P, _U32, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "dio_ply_parse": (_I, [_P, _P, ctypes.c_size_t, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_U32),
                           ctypes.POINTER(_U32)]),
}


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DioramaError(f"{LIB_PATH} is missing: build it with `python -m diorama_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise DioramaError(lib().dio_last_error().decode())


def _floats(values, n):
    arr = (ctypes.c_float * n)(*[float(v) for v in values])
    return arr


def _camera(cam):
    c = _Camera()
    c.view[:] = [float(v) for v in cam["view"]]
    c.proj[:] = [float(v) for v in cam["proj"]]
    c.viewport[:] = [float(v) for v in cam["viewport"]]
    return c


class Context:
    """One device and its in-order HIP stream (replaces initGpu, src/gpu.ts:8-47)."""

    def __init__(self, device=0):
        h = _P()
        _check(lib().dio_ctx_create(device, ctypes.byref(h)))
        self.handle = h

    def synchronize(self):
        _check(lib().dio_ctx_synchronize(self.handle))

    def buffer(self, nbytes):
        return Buffer.create(self, nbytes)

    def timestamp(self, slot):
        """Records timestamp `slot` on the stream (a timestamp query)."""
        _check(lib().dio_ctx_timestamp(self.handle, slot))

    def elapsed_ms(self, from_slot, to_slot):
        """Milliseconds between two recorded timestamps; waits for the later one."""
        ms = ctypes.c_double()
        _check(lib().dio_ctx_elapsed(self.handle, from_slot, to_slot, ctypes.byref(ms)))
        return ms.value

    def upload(self, array):
        a = np.ascontiguousarray(array)
        b = Buffer.create(self, a.nbytes)
        b.upload(a)
        return b

    def close(self):
        if self.handle:
            lib().dio_ctx_destroy(self.handle)
            self.handle = None


def _size(value, what):
    """A byte offset or size for the C ABI: a non-negative integer (ctypes would turn
    -16 into 2**64 - 16 without complaint)."""
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)) or value < 0:
        raise ValueError(f"{what} must be a non-negative integer, got {value!r}")
    return int(value)


class Buffer:
    """A device allocation. `owned` buffers are freed with the object."""

    def __init__(self, handle, owned, keep=None):
        self.handle = handle
        self.owned = owned
        self._keep = keep

    @classmethod
    def create(cls, ctx, nbytes):
        h = _P()
        _check(lib().dio_buf_create(ctx.handle, _size(nbytes, "size"), ctypes.byref(h)))
        return cls(h, True)

    @property
    def nbytes(self):
        return lib().dio_buf_size(self.handle)

    def upload(self, array, offset=0):
        a = np.ascontiguousarray(array)
        _check(lib().dio_buf_upload(self.handle, _size(offset, "offset"), a.ctypes.data_as(_P), a.nbytes))

    def download(self, dtype=np.float32, count=None, offset=0):
        itemsize = np.dtype(dtype).itemsize
        offset = _size(offset, "offset")
        if count is None:
            count = max(self.nbytes - offset, 0) // itemsize
        out = np.empty(_size(count, "count"), dtype=dtype)
        _check(lib().dio_buf_download(self.handle, offset, out.ctypes.data_as(_P), out.nbytes))
        return out

    def clear(self):
        _check(lib().dio_buf_clear(self.handle))

    def copy_from(self, src, nbytes=None, dst_offset=0, src_offset=0):
        """encoder.copyBufferToBuffer on this buffer's stream."""
        src_offset, dst_offset = _size(src_offset, "src_offset"), _size(dst_offset, "dst_offset")
        n = max(src.nbytes - src_offset, 0) if nbytes is None else _size(nbytes, "nbytes")
        _check(lib().dio_buf_copy(self.handle, dst_offset, src.handle, src_offset, n))

    def destroy(self):
        if self.owned and self.handle:
            lib().dio_buf_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _view(handle, keep):
    return None if not handle else Buffer(ctypes.c_void_p(handle), False, keep)


class PrefixSum:
    """src/prefixSum.ts:10-103: exclusive scan of `data` in place."""

    def __init__(self, ctx, data, capacity, count=None):
        h = _P()
        _check(lib().dio_prefix_sum_create(ctx.handle, data.handle, capacity, count.handle if count else None,
                                           ctypes.byref(h)))
        self.handle, self._keep = h, (data, count)

    @property
    def total(self):
        return _view(lib().dio_prefix_sum_total(self.handle), self)

    def encode(self):
        _check(lib().dio_prefix_sum_encode(self.handle))

    @staticmethod
    def encode_static(ctx, data, capacity, total, count=None):
        """The allocation-free scan (no reference counterpart): `data` scanned in place,
        the grand total into element 0 of `total`, with the library's static workspace."""
        _check(lib().dio_prefix_sum_encode_static(ctx.handle, data.handle, capacity, count.handle if count else None,
                                                  total.handle))

    def destroy(self):
        if self.handle:
            lib().dio_prefix_sum_destroy(self.handle)
            self.handle = None


class RadixSort:
    """src/radixSort.ts:27-142: stable sort of u32 or (low, high) u64 keys with u32 values."""

    def __init__(self, ctx, capacity, high_bits=0, count=None):
        h = _P()
        _check(lib().dio_radix_sort_create(ctx.handle, capacity, high_bits, count.handle if count else None,
                                           ctypes.byref(h)))
        self.handle, self._keep = h, count
        self.capacity = capacity

    @property
    def keys(self):
        return _view(lib().dio_radix_sort_keys(self.handle), self)

    @property
    def values(self):
        return _view(lib().dio_radix_sort_values(self.handle), self)

    def encode(self):
        _check(lib().dio_radix_sort_encode(self.handle))

    def destroy(self):
        if self.handle:
            lib().dio_radix_sort_destroy(self.handle)
            self.handle = None


class TileRasterizer:
    """src/tileRasterizer.ts:29-317 (for splats already on the device)."""

    def __init__(self, ctx=None, splats=None, count=0, initial_pair_capacity=0, handle=None, keep=None):
        if handle is not None:
            self.handle, self._keep, self.owned = handle, keep, False
            return
        h = _P()
        _check(lib().dio_tile_rasterizer_create(ctx.handle, splats.handle, count, initial_pair_capacity,
                                                ctypes.byref(h)))
        self.handle, self._keep, self.owned = h, splats, True

    def encode(self, target, viewport, background):
        _check(lib().dio_tile_rasterizer_encode(self.handle, target.handle if target else None, int(viewport[0]),
                                                int(viewport[1]), _floats(background, 3)))

    def encode_backward(self, viewport, pixel_grads, splat_grads):
        _check(lib().dio_tile_rasterizer_encode_backward(self.handle, int(viewport[0]), int(viewport[1]),
                                                         pixel_grads.handle, splat_grads.handle))

    def after_submit(self):
        _check(lib().dio_tile_rasterizer_after_submit(self.handle))

    def set_tuning(self, value):
        """Launch settings (dio_tuning keys; missing ones take their defaults)."""
        k = _Tuning(value.get("tile_size", 16), value.get("splat_batch", 64), value.get("digit_bits", 8),
                    value.get("persistent", 0), value.get("segment", 0), value.get("per_tile_depth", 0))
        _check(lib().dio_tile_rasterizer_set_tuning(self.handle, ctypes.byref(k)))

    @property
    def last_pair_count(self):
        return lib().dio_tile_rasterizer_last_pair_count(self.handle)

    @property
    def pair_capacity(self):
        return lib().dio_tile_rasterizer_pair_capacity(self.handle)

    @property
    def debug_buffers(self):
        return {"pairSplats": _view(lib().dio_tile_rasterizer_pair_splats(self.handle), self),
                "ranges": _view(lib().dio_tile_rasterizer_ranges(self.handle), self)}

    @property
    def pixel_color(self):
        return _view(lib().dio_tile_rasterizer_pixel_color(self.handle), self)

    @property
    def pixel_last(self):
        return _view(lib().dio_tile_rasterizer_pixel_last(self.handle), self)

    def destroy(self):
        if self.owned and self.handle:
            lib().dio_tile_rasterizer_destroy(self.handle)
        self.handle = None


class GaussianRenderer:
    """src/gaussianRenderer.ts:55-311."""

    def __init__(self, ctx, initial_pair_capacity=0, handle=None, keep=None):
        self.ctx = ctx
        if handle is not None:
            self.handle, self.owned, self._keep = handle, False, keep
            return
        h = _P()
        _check(lib().dio_renderer_create(ctx.handle, initial_pair_capacity, ctypes.byref(h)))
        self.handle, self.owned = h, True

    def set_gaussians(self, data, sh=None, degree=0):
        g = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        n = g.size // 16
        s = None if sh is None else np.ascontiguousarray(sh, dtype=np.float32).reshape(-1)
        if s is not None and s.size != n * 45:
            raise DioramaError(f"Expected {n * 45} spherical-harmonic floats, got {s.size}.")
        _check(lib().dio_renderer_set_gaussians(self.handle, g.ctypes.data_as(_P), n,
                                                None if s is None else s.ctypes.data_as(_P), degree))

    def adopt_gaussians(self, count, gaussians, sh, degree):
        """adoptGaussians (src/gaussianRenderer.ts:118-122): the renderer takes ownership
        of the two device buffers and destroys the old scene."""
        _check(lib().dio_renderer_adopt_gaussians(self.handle, count, gaussians.handle, sh.handle, degree))
        gaussians.owned = sh.owned = False

    def set_sh_degree(self, degree):
        _check(lib().dio_renderer_set_sh_degree(self.handle, degree))

    def encode(self, target, camera, background):
        c = _camera(camera)
        _check(lib().dio_renderer_encode(self.handle, target.handle if target else None, ctypes.byref(c),
                                         _floats(background, 3)))

    def encode_preprocess(self, camera):
        c = _camera(camera)
        _check(lib().dio_renderer_encode_preprocess(self.handle, ctypes.byref(c)))

    def encode_tiles(self, target, background):
        _check(lib().dio_renderer_encode_tiles(self.handle, target.handle if target else None, _floats(background, 3)))

    def encode_draw(self, target, splats, count, viewport, background):
        _check(lib().dio_renderer_encode_draw(self.handle, target.handle if target else None, splats.handle, count,
                                              int(viewport[0]), int(viewport[1]), _floats(background, 3)))

    def encode_backward(self, pixel_grads):
        _check(lib().dio_renderer_encode_backward(self.handle, pixel_grads.handle))

    def after_submit(self):
        _check(lib().dio_renderer_after_submit(self.handle))

    @property
    def count(self):
        return lib().dio_renderer_count(self.handle)

    @property
    def splat_buffer(self):
        return _view(lib().dio_renderer_splat_buffer(self.handle), self)

    @property
    def gaussian_buffer(self):
        return _view(lib().dio_renderer_gaussian_buffer(self.handle), self)

    @property
    def sh_buffer(self):
        return _view(lib().dio_renderer_sh_buffer(self.handle), self)

    @property
    def tiles(self):
        h = lib().dio_renderer_tiles(self.handle)
        return None if not h else TileRasterizer(handle=ctypes.c_void_p(h), keep=self)

    @property
    def pixel_buffer(self):
        return _view(lib().dio_renderer_pixel_buffer(self.handle), self)

    @property
    def pixel_last_buffer(self):
        return _view(lib().dio_renderer_pixel_last_buffer(self.handle), self)

    @property
    def gradients(self):
        g = lib().dio_renderer_gradient_gaussians(self.handle)
        if not g:
            return None
        return {"gaussians": _view(g, self), "sh": _view(lib().dio_renderer_gradient_sh(self.handle), self),
                "splats": _view(lib().dio_renderer_gradient_splats(self.handle), self)}

    def destroy(self):
        if self.owned and self.handle:
            lib().dio_renderer_destroy(self.handle)
        self.handle = None


class TrainingLoss:
    """src/trainingLoss.ts:16-130."""

    def __init__(self, ctx, width, height):
        h = _P()
        _check(lib().dio_training_loss_create(ctx.handle, width, height, ctypes.byref(h)))
        self.handle, self.width, self.height = h, width, height

    def encode(self, rendered, target):
        _check(lib().dio_training_loss_encode(self.handle, rendered.handle, target.handle))

    def read(self):
        v = ctypes.c_double()
        _check(lib().dio_training_loss_read(self.handle, ctypes.byref(v)))
        return v.value

    @property
    def pixel_grads(self):
        return _view(lib().dio_training_loss_pixel_grads(self.handle), self)

    @property
    def pixel_loss(self):
        return _view(lib().dio_training_loss_pixel_loss(self.handle), self)

    def destroy(self):
        if self.handle:
            lib().dio_training_loss_destroy(self.handle)
            self.handle = None


def default_settings(extent):
    """DEFAULT_SETTINGS (src/trainer.ts:65-82) with the scene extent, as a dict."""
    s = _Settings()
    lib().dio_default_settings(float(extent), ctypes.byref(s))
    out = {}
    for name, _ in _Settings._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def _settings(d):
    s = _Settings()
    for name, _ in _Settings._fields_:
        v = d[name]
        if hasattr(v, "__len__"):
            getattr(s, name)[:] = [float(x) for x in v]
        else:
            setattr(s, name, v)
    return s


def _images(images):
    arrs = [np.ascontiguousarray(im, dtype=np.uint32).reshape(-1) for im in images]
    ptrs = (_P * len(arrs))(*[a.ctypes.data_as(_P) for a in arrs])
    return arrs, ptrs


MLP_PARAMS = 5252
MLP_INPUTS = 12
MLP_CLASSES = ("keep", "clone", "split", "prune")


class DensifyMLP:
    """The learned densification proposal: 12 -> 64 -> 64 -> 4, BF16 weights, FP32
    accumulation (no counterpart in the reference)."""

    def __init__(self, ctx, seed=1):
        h = _P()
        _check(lib().dio_mlp_create(ctx.handle, seed, ctypes.byref(h)))
        self.handle, self.ctx = h, ctx

    @property
    def weights(self):
        out = np.empty(MLP_PARAMS, dtype=np.float32)
        _check(lib().dio_mlp_get_weights(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    @weights.setter
    def weights(self, w):
        a = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        if a.size != MLP_PARAMS:
            raise ValueError(f"the MLP has {MLP_PARAMS} parameters")
        _check(lib().dio_mlp_set_weights(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def forward(self, features, count, logits):
        _check(lib().dio_mlp_forward(self.handle, features.handle, count, logits.handle))

    def gradients(self, features, labels, count):
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().dio_mlp_gradients(self.handle, features.handle, labels.handle, count,
                                       g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)))
        return g, loss.value

    def set_input_norm(self, mean, scale):
        m = np.ascontiguousarray(mean, dtype=np.float32).reshape(MLP_INPUTS)
        sc = np.ascontiguousarray(scale, dtype=np.float32).reshape(MLP_INPUTS)
        _check(lib().dio_mlp_set_input_norm(self.handle, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            sc.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def set_class_weights(self, weights):
        a = np.ascontiguousarray(weights, dtype=np.float32).reshape(4)
        _check(lib().dio_mlp_set_class_weights(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def train(self, features, labels, count, steps, batch=4096, lr=1e-3, seed=1):
        loss = ctypes.c_double()
        _check(lib().dio_mlp_train(self.handle, features.handle, labels.handle, count, steps, batch, lr, seed,
                                   ctypes.byref(loss)))
        return loss.value

    def destroy(self):
        if self.handle:
            lib().dio_mlp_destroy(self.handle)
            self.handle = None


class Trainer:
    """src/trainer.ts:103-446: one training view per step, plain 3DGS schedule.

    With `batch` > 0 it is a multi-view batch trainer: `batch` views of any sizes
    per step, their gradients summed into one Adam step (one captured HIP graph)."""

    def __init__(self, ctx, params, sh, cameras, images, settings, batch=0):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        s = np.ascontiguousarray(sh, dtype=np.float32).reshape(-1)
        n = p.size // 16
        cams = (_Camera * len(cameras))(*[_camera(c) for c in cameras])
        arrs, ptrs = _images(images)
        st = _settings(settings)
        h = _P()
        if batch:
            _check(lib().dio_trainer_create_batched(ctx.handle, p.ctypes.data_as(_P), s.ctypes.data_as(_P), n, cams,
                                                    ptrs, len(cameras), batch, ctypes.byref(st), ctypes.byref(h)))
        else:
            _check(lib().dio_trainer_create(ctx.handle, p.ctypes.data_as(_P), s.ctypes.data_as(_P), n, cams, ptrs,
                                            len(cameras), ctypes.byref(st), ctypes.byref(h)))
        self.handle, self.ctx = h, ctx

    @property
    def batch_size(self):
        return lib().dio_trainer_batch_size(self.handle)

    def set_densify_mode(self, mode, mlp=None, budget=0, confidence=0.5, samples_per_densify=80000):
        """mode 0 plain rule, 1 plain rule + sample logging, 2 learned (see dio_trainer_set_densify_mode)."""
        modes = {"plain": 0, "log": 1, "learned": 2}
        m = modes.get(mode, mode)
        _check(lib().dio_trainer_set_densify_mode(self.handle, m, mlp.handle if mlp else None, budget, confidence,
                                                  samples_per_densify))

    @property
    def tuning(self):
        k = _Tuning()
        _check(lib().dio_trainer_get_tuning(self.handle, ctypes.byref(k)))
        return {"tile_size": k.tile_size, "splat_batch": k.splat_batch, "digit_bits": k.digit_bits,
                "persistent": k.persistent, "segment": k.segment, "per_tile_depth": k.per_tile_depth}

    @tuning.setter
    def tuning(self, value):
        k = _Tuning(value.get("tile_size", 16), value.get("splat_batch", 64), value.get("digit_bits", 8),
                    value.get("persistent", 0), value.get("segment", 0), value.get("per_tile_depth", 0))
        _check(lib().dio_trainer_set_tuning(self.handle, ctypes.byref(k)))

    def autotune(self, steps=3):
        """Times every candidate setting over `steps` training steps, keeps the fastest;
        returns [(tuning, ms per step)]."""
        out = (_TuningResult * 128)()
        n = _U32()
        _check(lib().dio_trainer_autotune(self.handle, steps, out, 128, ctypes.byref(n)))
        return [({"tile_size": r.tuning.tile_size, "splat_batch": r.tuning.splat_batch,
                  "digit_bits": r.tuning.digit_bits, "persistent": r.tuning.persistent,
                  "segment": r.tuning.segment, "per_tile_depth": r.tuning.per_tile_depth}, r.ms_per_step)
                for r in out[:min(n.value, 128)]]

    def samples(self):
        f, l, n = _P(), _P(), _U32()
        _check(lib().dio_trainer_samples(self.handle, ctypes.byref(f), ctypes.byref(l), ctypes.byref(n)))
        return _view(f.value, self), _view(l.value, self), n.value

    def features(self):
        f = _P()
        _check(lib().dio_trainer_features(self.handle, ctypes.byref(f)))
        return _view(f.value, self)

    @property
    def extra_stats(self):
        return _view(lib().dio_trainer_extra_stats(self.handle), self)

    @property
    def footprint(self):
        """Mode "log": per Gaussian (sum of blend weight x loss share, sum of blend weight)."""
        return _view(lib().dio_trainer_footprint(self.handle), self)

    def batch_frames(self, cameras):
        """A batch trainer's last step, downloaded: per slot (training view index,
        splats (count, 12) float32, pixel colours (w*h, 4) float32); `cameras` are
        the trainer's training cameras, for the image sizes."""
        views = (_U32 * 64)()
        count, stride, buf, pix = _U32(), _U32(), _P(), _P()
        _check(lib().dio_trainer_batch_frames(self.handle, views, 64, ctypes.byref(count), ctypes.byref(buf),
                                              ctypes.byref(stride), ctypes.byref(pix)))
        n = self.count
        splats = _view(buf.value, self).download(np.float32, count.value * stride.value * 12)
        splats = splats.reshape(count.value, stride.value, 12)[:, :n]
        sizes = [int(cameras[views[k]]["viewport"][0]) * int(cameras[views[k]]["viewport"][1])
                 for k in range(count.value)]
        colors = _view(pix.value, self).download(np.float32, 4 * sum(sizes)).reshape(-1, 4)
        starts = np.concatenate([[0], np.cumsum(sizes)])
        return [(int(views[k]), splats[k], colors[starts[k]:starts[k + 1]]) for k in range(count.value)]

    def step(self):
        _check(lib().dio_trainer_step(self.handle))

    def optimize(self):
        _check(lib().dio_trainer_optimize(self.handle))

    def maintain(self):
        _check(lib().dio_trainer_maintain(self.handle))

    def densify(self, prune_large):
        _check(lib().dio_trainer_densify(self.handle, 1 if prune_large else 0))

    def reset_opacity(self):
        _check(lib().dio_trainer_reset_opacity(self.handle))

    def export_raw(self):
        n = self.count
        p = np.empty(n * 16, dtype=np.float32)
        s = np.empty(n * 45, dtype=np.float32)
        _check(lib().dio_trainer_export_raw(self.handle, p.ctypes.data_as(_P), s.ctypes.data_as(_P)))
        return {"params": p.reshape(n, 16), "sh": s.reshape(n, 45)}

    def last_loss(self):
        v = ctypes.c_double()
        _check(lib().dio_trainer_last_loss(self.handle, ctypes.byref(v)))
        return v.value

    def evaluate(self, cameras, images, with_ssim=True):
        cams = (_Camera * len(cameras))(*[_camera(c) for c in cameras])
        arrs, ptrs = _images(images)
        p, s = ctypes.c_double(), ctypes.c_double()
        _check(lib().dio_trainer_evaluate(self.handle, cams, ptrs, len(cameras), 1 if with_ssim else 0,
                                          ctypes.byref(p), ctypes.byref(s)))
        return {"psnr": p.value, "ssim": s.value}

    def measure_psnr(self, cameras, images):
        """measurePsnr (src/insideView.ts:233-267): the mean PSNR of views whose photos
        are device Buffers (rgba8)."""
        cams = (_Camera * len(cameras))(*[_camera(c) for c in cameras])
        ptrs = (_P * len(images))(*[b.handle for b in images])
        p = ctypes.c_double()
        _check(lib().dio_trainer_measure_psnr(self.handle, cams, ptrs, len(cameras), ctypes.byref(p)))
        return p.value

    def time_stage(self, stage, repeats=10):
        v = ctypes.c_double()
        _check(lib().dio_trainer_time_stage(self.handle, stage.encode(), repeats, ctypes.byref(v)))
        return v.value

    def stage_clock(self, enable=True):
        """Starts (or stops) summing HIP-event times per stage over the plain steps
        that follow; either way the totals start from zero."""
        _check(lib().dio_trainer_stage_clock(self.handle, 1 if enable else 0))

    def stage_totals(self):
        """Since the clock started: {"ms": milliseconds summed per stage name, "steps":
        steps clocked, "pairs": their summed (splat, tile) pair counts}."""
        stages, steps, pairs = _U32(), _U32(), ctypes.c_uint64()
        _check(lib().dio_trainer_stage_totals(self.handle, None, 0, ctypes.byref(stages), None, None))
        ms = (ctypes.c_double * stages.value)()
        _check(lib().dio_trainer_stage_totals(self.handle, ms, stages.value, None, ctypes.byref(steps),
                                              ctypes.byref(pairs)))
        names = [lib().dio_trainer_stage_name(k).decode() for k in range(stages.value)]
        return {"ms": dict(zip(names, ms)), "steps": steps.value, "pairs": pairs.value}

    @property
    def count(self):
        return lib().dio_trainer_count(self.handle)

    @property
    def iteration(self):
        return lib().dio_trainer_iteration(self.handle)

    @iteration.setter
    def iteration(self, value):
        lib().dio_trainer_set_iteration(self.handle, int(value))

    @property
    def sh_degree(self):
        return lib().dio_trainer_sh_degree(self.handle)

    @sh_degree.setter
    def sh_degree(self, value):
        _check(lib().dio_trainer_set_sh_degree(self.handle, int(value)))

    @property
    def last_view(self):
        """The training view the last plain step used."""
        return lib().dio_trainer_last_view(self.handle)

    @property
    def densifications(self):
        out = []
        for k in range(lib().dio_trainer_densification_count(self.handle)):
            d = _Densification()
            _check(lib().dio_trainer_densification(self.handle, k, ctypes.byref(d)))
            out.append({name: getattr(d, name) for name, _ in _Densification._fields_})
        return out

    @property
    def buffers(self):
        return {"params": _view(lib().dio_trainer_params(self.handle), self),
                "moments": _view(lib().dio_trainer_moments(self.handle), self),
                "stats": _view(lib().dio_trainer_stats(self.handle), self)}

    def write_ply(self):
        """writeGaussianPly(trainer.exportRaw()) without the host round trip: the
        .ply packed from the trainer's buffers in device memory."""
        raw = {"params": self.buffers["params"], "sh": self.renderer.sh_buffer}
        return write_gaussian_ply(raw, count=self.count)

    @property
    def renderer(self):
        return GaussianRenderer(self.ctx, handle=ctypes.c_void_p(lib().dio_trainer_renderer(self.handle)), keep=self)

    def destroy(self):
        if self.handle:
            lib().dio_trainer_destroy(self.handle)
            self.handle = None


def write_gaussian_ply(raw, ctx=None, count=None):
    """writeGaussianPly (src/plyWriter.ts:14-43): the .ply file of raw Gaussians as bytes.

    raw: {"params": count x 16, "sh": count x 45} as host arrays (uploaded on `ctx`)
    or device Buffers (then `count` may be given; the body is packed on the device)."""
    params, sh = raw["params"], raw["sh"]
    keep = []
    if not isinstance(params, Buffer):
        if ctx is None:
            raise DioramaError("write_gaussian_ply needs a Context for host arrays")
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        s = np.ascontiguousarray(sh, dtype=np.float32).reshape(-1)
        count = p.size // 16 if count is None else count
        if s.size < count * 45:
            raise DioramaError("sh holds fewer than 45 floats per Gaussian")
        params, sh = ctx.upload(p), ctx.upload(s)
        keep = [params, sh]
    elif count is None:
        count = params.nbytes // 64
    length = ctypes.c_size_t()
    _check(lib().dio_ply_write(None, None, count, None, 0, ctypes.byref(length)))
    out = ctypes.create_string_buffer(length.value)
    _check(lib().dio_ply_write(params.handle, sh.handle, count, out, length.value, ctypes.byref(length)))
    del keep
    return out.raw


def parse_gaussian_ply(data, ctx, on_device=False):
    """parseGaussianPly (src/plyLoader.ts:21-98): {"count", "gaussians" (count x 16,
    activated), "sh_rest" (count x 45), "sh_degree"}; host arrays, or device Buffers
    (ready for GaussianRenderer.adopt_gaussians) with on_device."""
    buf = bytes(data)
    g, s, n, d = _P(), _P(), _U32(), _U32()
    _check(lib().dio_ply_parse(ctx.handle if ctx else None, buf, len(buf), ctypes.byref(g), ctypes.byref(s),
                               ctypes.byref(n), ctypes.byref(d)))
    gb, sb = Buffer(g, True), Buffer(s, True)
    out = {"count": n.value, "sh_degree": d.value}
    if on_device:
        out.update(gaussians=gb, sh_rest=sb)
    else:
        out.update(gaussians=gb.download(np.float32, n.value * 16).reshape(-1, 16),
                   sh_rest=sb.download(np.float32, n.value * 45).reshape(-1, 45))
        gb.destroy()
        sb.destroy()
    return out


def mean_neighbor_distance2(ctx, positions):
    """meanNeighborDistance2 (src/loadDataset.ts:126-149): per point the mean squared
    distance to its 3 nearest other points, computed on the device."""
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    n = p.size // 3
    pts = ctx.upload(p)
    out = ctx.buffer(max(n, 1) * 4)
    _check(lib().dio_mean_neighbor_distance2(ctx.handle, pts.handle, n, out.handle))
    return out.download(np.float32, n)


def initial_gaussians(ctx, positions, colors):
    """initialGaussians (src/loadDataset.ts:87-110): {"params": n x 16, "sh": n x 45}
    raw Gaussians for structure-from-motion points and their 0-255 colours."""
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    c = np.ascontiguousarray(colors, dtype=np.uint8).reshape(-1)
    n = p.size // 3
    if c.size < 3 * n:
        raise DioramaError("colors hold fewer than 3 bytes per point")
    pts, cols = ctx.upload(p), ctx.upload(c)
    params, sh = ctx.buffer(max(n, 1) * 64), ctx.buffer(max(n, 1) * 180)
    _check(lib().dio_initial_gaussians(ctx.handle, pts.handle, cols.handle, n, params.handle, sh.handle))
    return {"params": params.download(np.float32, n * 16).reshape(n, 16),
            "sh": sh.download(np.float32, n * 45).reshape(n, 45)}


def _colmap_camera(c):
    out = _ColmapCamera()
    out.name = c["name"].encode()
    out.width, out.height = int(c["width"]), int(c["height"])
    out.fx, out.fy = float(c["fx"]), float(c["fy"])
    for key in ("position", "right", "down", "forward"):
        getattr(out, key)[:] = [float(v) for v in c[key]]
    return out


def parse_colmap(cameras_bin, images_bin, points_bin):
    """parseColmap (src/colmap.ts:18-129): {"cameras": [TrainingCamera dicts sorted by
    name], "positions": n x 3 float32, "colors": n x 3 uint8}."""
    files = [bytes(cameras_bin), bytes(images_bin), bytes(points_bin)]
    nc, npts = _U32(), _U32()
    args = []
    for f in files:
        args += [f, len(f)]
    _check(lib().dio_colmap_parse(*args, None, 0, ctypes.byref(nc), None, None, 0, ctypes.byref(npts)))
    cams = (_ColmapCamera * max(nc.value, 1))()
    positions = np.empty((npts.value, 3), np.float32)
    colors = np.empty((npts.value, 3), np.uint8)
    _check(lib().dio_colmap_parse(*args, cams, nc.value, ctypes.byref(nc), positions.ctypes.data_as(_P),
                                  colors.ctypes.data_as(_P), npts.value, ctypes.byref(npts)))
    cameras = [{"name": c.name.decode(), "width": c.width, "height": c.height, "fx": c.fx, "fy": c.fy,
                "position": list(c.position), "right": list(c.right), "down": list(c.down),
                "forward": list(c.forward)} for c in cams[:nc.value]]
    return {"cameras": cameras, "positions": positions, "colors": colors}


def training_camera_data(camera, viewport):
    """trainingCameraData (src/loadDataset.ts:80-86): the renderer's camera for a
    TrainingCamera rendered at viewport (width, height)."""
    out = _Camera()
    _check(lib().dio_colmap_training_camera(ctypes.byref(_colmap_camera(camera)), int(viewport[0]),
                                            int(viewport[1]), ctypes.byref(out)))
    return {"view": list(out.view), "proj": list(out.proj), "viewport": list(out.viewport)}


def split_dataset(cameras, test_every=8):
    """loadDataset's split and extent (src/loadDataset.ts:10-43): every 8th photo by
    name is held out; extent = 1.1 x the farthest training camera from their mean."""
    train = [c for i, c in enumerate(cameras) if i % test_every != 0]
    test = [c for i, c in enumerate(cameras) if i % test_every == 0]
    centers = np.array([c["position"] for c in train], np.float64).reshape(-1, 3)
    extent = 1.1 * float(np.linalg.norm(centers - centers.mean(axis=0), axis=1).max()) if len(centers) else 0.0
    return {"train": train, "test": test, "extent": extent}
