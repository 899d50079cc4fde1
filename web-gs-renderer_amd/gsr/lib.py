"""ctypes binding of the renderer's C ABI (include/gsr.h).

The Python side is the test and benchmark harness: it mirrors what the
reference's Renderer does per frame (src/renderer/renderer.js:481-508) but
every computation runs in libgsr.so on the GPU.  Loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "build", "libgsr.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "gsr.h")

GSR_OK = 0
GSR_ERR_ARGUMENT = 1
GSR_ERR_DEVICE = 2
GSR_ERR_ALLOC = 3
GSR_ERR_NOT_FOUND = 4
GSR_ERR_CAPACITY = 5
GSR_ERR_KERNEL = 6
GSR_ERR_STATE = 7

KEY_AUTO, KEY_32, KEY_64 = 0, 4, 8


class Camera(ctypes.Structure):
    _fields_ = [("view", ctypes.c_float * 16), ("view_proj", ctypes.c_float * 16), ("eye", ctypes.c_float * 3),
                ("count", ctypes.c_uint32), ("focal", ctypes.c_float * 2), ("tan_fov", ctypes.c_float * 2),
                ("texture_size", ctypes.c_float * 2), ("pad", ctypes.c_uint32 * 2)]

    @staticmethod
    def from_uniform(u: np.ndarray) -> "Camera":
        b = np.ascontiguousarray(u, dtype=np.float32).tobytes()
        if len(b) != 176:
            raise ValueError("uniform block must be 44 float32 slots (176 bytes)")
        return Camera.from_buffer_copy(b)


assert ctypes.sizeof(Camera) == 176


class Options(ctypes.Structure):
    _fields_ = [("tile_size", ctypes.c_uint32), ("key_bytes", ctypes.c_uint32),
                ("instance_capacity", ctypes.c_uint64), ("timing", ctypes.c_uint32),
                ("save_state", ctypes.c_uint32), ("raster_waves_per_cu", ctypes.c_uint32),
                ("sort_blocks_per_cu", ctypes.c_uint32), ("sort_digit_bits", ctypes.c_uint32)]


class TimingRow(ctypes.Structure):
    _fields_ = [("label", ctypes.c_char * 32), ("ms", ctypes.c_double)]


class GsrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[gsr {code}] {message}")
        self.code = code


_lib = None
_VP = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_INT = ctypes.c_int


def header_functions() -> list:
    """Names of every function declared in include/gsr.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gsr_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libgsr.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.gsr_create.argtypes = [ctypes.POINTER(_VP), _INT]
    L.gsr_destroy.argtypes = [_VP]
    L.gsr_last_error.argtypes = [_VP]
    L.gsr_synchronize.argtypes = [_VP]
    L.gsr_stream.argtypes = [_VP]
    L.gsr_scene_upload.argtypes = [_VP, ctypes.c_char_p, _VP, _U32]
    L.gsr_scene_select.argtypes = [_VP, ctypes.c_char_p]
    L.gsr_scene_free.argtypes = [_VP, ctypes.c_char_p]
    L.gsr_scene_params.argtypes = [_VP, ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.POINTER(_U32)]
    L.gsr_render.argtypes = [_VP, ctypes.POINTER(Camera), ctypes.POINTER(Options), _VP, _VP]
    L.gsr_forward.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), ctypes.POINTER(Options), _VP]
    L.gsr_last_instance_count.argtypes = [_VP, ctypes.POINTER(_U64)]
    L.gsr_get_timings.argtypes = [_VP, ctypes.POINTER(TimingRow), _INT, ctypes.POINTER(_INT)]
    L.gsr_to_rgba8.argtypes = [_VP, _VP, _U32, _U32, _INT, _VP]
    L.gsr_preprocess.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), _U32, _VP, _VP, _VP, _VP, _VP]
    L.gsr_scan_offsets.argtypes = [_VP, _VP, _U32, _VP, _VP]
    L.gsr_emit.argtypes = [_VP, _VP, _VP, _U32, _U32, _U32, _U64, _U32, _VP, _VP]
    L.gsr_sort_pairs.argtypes = [_VP, _U32, _VP, _VP, _VP, _VP, _U64, _U32, _U32, ctypes.POINTER(_INT)]
    L.gsr_sort_pairs_geometry.argtypes = [_VP, _U32, _VP, _VP, _VP, _VP, _U64, _U32, _U32, _U32, _U32,
                                          ctypes.POINTER(_INT)]
    L.gsr_sort_pairs_noalloc.argtypes = [_VP, _U32, _VP, _VP, _VP, _VP, _U64, _U32, _U32, ctypes.POINTER(_INT)]
    L.gsr_tile_ranges.argtypes = [_VP, _VP, _U32, _U64, _U32, _VP]
    L.gsr_rasterize.argtypes = [_VP, _U32, _U32, _U32, _U32, _U64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
    L.gsr_backward.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), _VP, _VP, _VP]
    L.gsr_render_batch.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), _U32, ctypes.POINTER(Options), _INT,
                                   _VP, ctypes.POINTER(_U64)]
    L.gsr_batch_instance_count.argtypes = [_VP, ctypes.POINTER(_U64)]
    L.gsr_batch_capacity.argtypes = [_VP, ctypes.POINTER(_U64)]
    L.gsr_backward_batch.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), _U32, _VP, _VP, _VP]
    L.gsr_autotune_batch.argtypes = [_VP, _VP, _U32, ctypes.POINTER(Camera), _U32, ctypes.POINTER(Options), _VP,
                                     ctypes.POINTER(ctypes.c_double)]
    special = {"gsr_destroy": None, "gsr_last_error": ctypes.c_char_p, "gsr_stream": _VP}
    for name in header_functions():
        getattr(L, name).restype = special.get(name, ctypes.c_int)
    _lib = L
    return L


def last_error(ctx=None) -> str:
    return load().gsr_last_error(ctx).decode(errors="replace")


def options(tile_size=16, key_bytes=KEY_AUTO, capacity=0, timing=False, save_state=False, raster_waves_per_cu=0,
            sort_blocks_per_cu=0, sort_digit_bits=0) -> Options:
    return Options(tile_size, key_bytes, capacity, int(timing), 1 if save_state else 0, raster_waves_per_cu,
                   sort_blocks_per_cu, sort_digit_bits)


def _ptr(x):
    """Device address of a torch tensor, or an int address."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


class Context:
    """One renderer context: one HIP stream, resident scenes, frame buffers."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = _VP()
        rc = self.L.gsr_create(ctypes.byref(h), device)
        if rc != GSR_OK:
            raise GsrError(rc, last_error(None))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.gsr_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GSR_OK:
            raise GsrError(rc, last_error(self.h))

    @property
    def stream(self) -> int:
        return self.L.gsr_stream(self.h) or 0

    def synchronize(self):
        self._check(self.L.gsr_synchronize(self.h))

    # scenes -------------------------------------------------------------
    def scene_upload(self, name: str, packed: np.ndarray, count: int):
        packed = np.ascontiguousarray(packed, dtype=np.float32)
        if packed.size != count * 59:
            raise ValueError("packed array must hold 59 floats per splat")
        self._check(self.L.gsr_scene_upload(self.h, name.encode(), packed.ctypes.data, count))

    def scene_select(self, name: str):
        self._check(self.L.gsr_scene_select(self.h, name.encode()))

    def scene_free(self, name: str):
        self._check(self.L.gsr_scene_free(self.h, name.encode()))

    def scene_params(self, name: str):
        p = _VP()
        n = _U32()
        self._check(self.L.gsr_scene_params(self.h, name.encode(), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    # frames -------------------------------------------------------------
    def render(self, uniform: np.ndarray, opts: Options | None = None, out_dev=None, host=True) -> np.ndarray | None:
        cam = Camera.from_uniform(uniform)
        W, H = int(cam.texture_size[0]), int(cam.texture_size[1])
        out = np.empty((3, H, W), np.float32) if host else None
        self._check(self.L.gsr_render(self.h, ctypes.byref(cam), ctypes.byref(opts) if opts else None,
                                      out.ctypes.data if host else None, _ptr(out_dev)))
        return out

    def forward(self, params_dev, count: int, uniform: np.ndarray, opts: Options | None, out_dev):
        cam = Camera.from_uniform(uniform)
        self._check(self.L.gsr_forward(self.h, _ptr(params_dev), count, ctypes.byref(cam),
                                       ctypes.byref(opts) if opts else None, _ptr(out_dev)))

    def backward(self, params_dev, count: int, uniform: np.ndarray, dl_drgb_dev, grad_dev, grad2d_dev=None):
        cam = Camera.from_uniform(uniform)
        self._check(self.L.gsr_backward(self.h, _ptr(params_dev), count, ctypes.byref(cam), _ptr(dl_drgb_dev),
                                        _ptr(grad_dev), _ptr(grad2d_dev)))

    def render_batch(self, params_dev, count: int, uniforms: list, opts: Options | None, out_dev, use_graph=False):
        cams = (Camera * len(uniforms))(*[Camera.from_uniform(u) for u in uniforms])
        offs = (_U64 * (len(uniforms) + 1))()
        self._check(self.L.gsr_render_batch(self.h, _ptr(params_dev), count, cams, len(uniforms),
                                            ctypes.byref(opts) if opts else None, 1 if use_graph else 0,
                                            _ptr(out_dev), offs))
        return list(offs)

    def backward_batch(self, params_dev, count: int, uniforms: list, dl_drgb_dev, grad_dev, grad2d_dev=None):
        """Gradients of the last render_batch (save_state) summed over its views."""
        cams = (Camera * len(uniforms))(*[Camera.from_uniform(u) for u in uniforms])
        self._check(self.L.gsr_backward_batch(self.h, _ptr(params_dev), count, cams, len(uniforms),
                                              _ptr(dl_drgb_dev), _ptr(grad_dev), _ptr(grad2d_dev)))

    def batch_layout(self, count: int, uniforms: list, opts: Options | None = None):
        """Float offsets of each view's image in a batch output (views + 1 entries)."""
        cams = (Camera * len(uniforms))(*[Camera.from_uniform(u) for u in uniforms])
        offs = (_U64 * (len(uniforms) + 1))()
        self._check(self.L.gsr_render_batch(self.h, 1, count, cams, len(uniforms),
                                            ctypes.byref(opts) if opts else None, 0, None, offs))
        return list(offs)

    def batch_instance_count(self) -> int:
        v = _U64()
        self._check(self.L.gsr_batch_instance_count(self.h, ctypes.byref(v)))
        return v.value

    def batch_capacity(self) -> int:
        v = _U64()
        self._check(self.L.gsr_batch_capacity(self.h, ctypes.byref(v)))
        return v.value

    def autotune_batch(self, params_dev, count: int, uniforms: list, opts: Options, out_dev):
        cams = (Camera * len(uniforms))(*[Camera.from_uniform(u) for u in uniforms])
        best = ctypes.c_double()
        self._check(self.L.gsr_autotune_batch(self.h, _ptr(params_dev), count, cams, len(uniforms), ctypes.byref(opts),
                                              _ptr(out_dev), ctypes.byref(best)))
        return opts, best.value

    def instance_count(self) -> int:
        v = _U64()
        self._check(self.L.gsr_last_instance_count(self.h, ctypes.byref(v)))
        return v.value

    def timings(self) -> dict:
        n = _INT()
        self._check(self.L.gsr_get_timings(self.h, None, 0, ctypes.byref(n)))
        rows = (TimingRow * max(n.value, 1))()
        self._check(self.L.gsr_get_timings(self.h, rows, n.value, ctypes.byref(n)))
        return {rows[i].label.decode(): rows[i].ms for i in range(n.value)}

    def to_rgba8(self, rgb_dev, width: int, height: int, flip_y=False) -> np.ndarray:
        out = np.empty((height, width, 4), np.uint8)
        self._check(self.L.gsr_to_rgba8(self.h, _ptr(rgb_dev), width, height, 1 if flip_y else 0, out.ctypes.data))
        return out

    # stage entry points (device arrays) ------------------------------------
    def preprocess(self, params_dev, count, uniform, tile, tiles, meta, conic, xy, rgb):
        cam = Camera.from_uniform(uniform)
        self._check(self.L.gsr_preprocess(self.h, _ptr(params_dev), count, ctypes.byref(cam), tile, _ptr(tiles),
                                          _ptr(meta), _ptr(conic), _ptr(xy), _ptr(rgb)))

    def scan_offsets(self, counts, n, offsets, total=None):
        self._check(self.L.gsr_scan_offsets(self.h, _ptr(counts), n, _ptr(offsets), _ptr(total)))

    def emit(self, offsets, meta, count, width, tile, capacity, key_bytes, keys, vals):
        self._check(self.L.gsr_emit(self.h, _ptr(offsets), _ptr(meta), count, width, tile, capacity, key_bytes,
                                    _ptr(keys), _ptr(vals)))

    def sort_pairs(self, key_bytes, k0, v0, k1, v1, n, begin_bit, end_bit, blocks_per_cu=None,
                   digit_bits=None) -> int:
        idx = _INT()
        if blocks_per_cu is None and digit_bits is None:
            self._check(self.L.gsr_sort_pairs(self.h, key_bytes, _ptr(k0), _ptr(v0), _ptr(k1), _ptr(v1), n,
                                              begin_bit, end_bit, ctypes.byref(idx)))
        else:
            self._check(self.L.gsr_sort_pairs_geometry(self.h, key_bytes, _ptr(k0), _ptr(v0), _ptr(k1), _ptr(v1), n,
                                                       begin_bit, end_bit, blocks_per_cu or 1, digit_bits or 8,
                                                       ctypes.byref(idx)))
        return idx.value

    def sort_pairs_noalloc(self, key_bytes, k0, v0, k1, v1, n, begin_bit, end_bit) -> int:
        """gsr_sort_pairs without device allocation (fixed per-GPU scratch)."""
        idx = _INT()
        self._check(self.L.gsr_sort_pairs_noalloc(self.h, key_bytes, _ptr(k0), _ptr(v0), _ptr(k1), _ptr(v1), n,
                                                  begin_bit, end_bit, ctypes.byref(idx)))
        return idx.value

    def tile_ranges(self, keys, key_bytes, n, num_tiles, ranges):
        self._check(self.L.gsr_tile_ranges(self.h, _ptr(keys), key_bytes, n, num_tiles, _ptr(ranges)))

    def rasterize(self, width, height, tile, count, instances, ranges, vals, conic, xy, rgb, out, tfinal=None,
                  last=None):
        self._check(self.L.gsr_rasterize(self.h, width, height, tile, count, instances, _ptr(ranges), _ptr(vals),
                                         _ptr(conic), _ptr(xy), _ptr(rgb), _ptr(out), _ptr(tfinal), _ptr(last)))
