"""ctypes access to the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the reference for checks and CPU timing.  The renderer never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-B"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_exp_f32.restype = ctypes.c_float
        L.orc_exp_f32.argtypes = [ctypes.c_float]
        L.orc_exp_f32_array.argtypes = [_f32p, _f32p, ctypes.c_int64]
        L.orc_exp_blend_f32_array.argtypes = [_f32p, _f32p, ctypes.c_int64]
        L.orc_blend_alpha_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int64]
        L.orc_power_f32_array.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64]
        L.orc_transmit_f32_array.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int64]
        L.orc_exp_blend_shift_mismatches.restype = ctypes.c_int64
        L.orc_exp_blend_shift_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_preprocess_f32.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32,
                                         _u32p, _u32p, _f32p, _f32p, _f32p]
        L.orc_preprocess_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32,
                                         _u32p, _u32p, _f64p, _f64p, _f64p]
        L.orc_scan.restype = ctypes.c_uint64
        L.orc_scan.argtypes = [_u32p, ctypes.c_uint32, _u32p]
        L.orc_emit.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint64, _u64p, _u32p]
        L.orc_sort_stable.argtypes = [_u64p, _u32p, ctypes.c_uint64]
        L.orc_sort_radix.argtypes = [_u64p, _u32p, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_tile_ranges.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, _u32p]
        L.orc_raster_f32.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                     _f32p, _f32p, _f32p, _f32p, _f32p, _u32p]
        L.orc_raster_f64.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                     _f64p, _f64p, _f64p, _f64p, _f64p, _u32p]
        L.orc_render_f32.restype = ctypes.c_int
        L.orc_render_f32.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32, ctypes.c_uint64,
                                     _f32p, _f32p, _u32p, _u64p, _f64p, ctypes.c_int, ctypes.c_void_p]
        L.orc_render_f64.restype = ctypes.c_int
        L.orc_render_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32, ctypes.c_uint64,
                                     _f64p, _f64p, _u32p, _u64p, ctypes.c_void_p]
        L.orc_backward_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32, _f64p, ctypes.c_int,
                                       _f64p, _f64p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_param_mass_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, _f64p, _f64p]
        L.orc_fd_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32, _f64p, _u32p, ctypes.c_uint32,
                                 ctypes.c_float, _f64p, np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]
        L.orc_fd_window_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32, _f64p, _u32p,
                                        ctypes.c_uint32, ctypes.c_float, _f64p,
                                        np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]
        L.orc_raster_backward_f32.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _f32p,
                                              _f32p, _f32p, _f32p, _u32p, _f32p, ctypes.c_uint32, _f64p]
        L.orc_preprocess_backward_f64.argtypes = [_f32p, ctypes.c_uint32, _f32p, _f64p, _f64p]
        _lib = L
    return _lib


def exp_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().orc_exp_f32_array(x.ravel(), y.ravel(), x.size)
    return y


def exp_blend_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().orc_exp_blend_f32_array(x.ravel(), y.ravel(), x.size)
    return y


def blend_alpha_f32(power, op):
    """The FP32 raster's alpha for each (power, opacity) pair, 0 where the pixel skips the splat."""
    p, o = np.broadcast_arrays(np.asarray(power, np.float32), np.asarray(op, np.float32))
    p = np.ascontiguousarray(p, np.float32)
    o = np.ascontiguousarray(o, np.float32)
    a = np.empty_like(p)
    lib().orc_blend_alpha_f32(p.ravel(), o.ravel(), a.ravel(), p.size)
    return a


def power_f32(A, B, C, dx, dy):
    """The FP32 blend's power of a splat with conic (A, B, C) at pixel offset
    (dx, dy) = (splat - pixel), elementwise, in the kernels' order."""
    arrs = [np.ascontiguousarray(a, np.float32).ravel() for a in np.broadcast_arrays(
        *[np.asarray(v, np.float32) for v in (A, B, C, dx, dy)])]
    out = np.empty_like(arrs[0])
    lib().orc_power_f32_array(*arrs, out, out.size)
    return out


def transmit_f32(T, alpha):
    """The FP32 blend's transmittance after blending alpha, elementwise."""
    t, a = [np.ascontiguousarray(v, np.float32).ravel() for v in np.broadcast_arrays(
        np.asarray(T, np.float32), np.asarray(alpha, np.float32))]
    out = np.empty_like(t)
    lib().orc_transmit_f32_array(t, a, out, out.size)
    return out


def exp_blend_shift_mismatches(mag_lo, mag_hi):
    """How many powers p in [-mag_hi, -mag_lo] (every float) give a different
    exp_blend result when 2^f is scaled by adding bits(t) << 23 to its bits
    (the GPU kernels' form) instead of by ldexp."""
    lo = int(np.float32(mag_lo).view(np.uint32))
    hi = int(np.float32(mag_hi).view(np.uint32))
    return int(lib().orc_exp_blend_shift_mismatches(lo, hi))


def preprocess(packed, n, uni, tile=16, fp64=False):
    dt = np.float64 if fp64 else np.float32
    tiles = np.zeros(n, np.uint32)
    meta = np.zeros(4 * n, np.uint32)
    conic = np.zeros(4 * n, dt)
    xy = np.zeros(2 * n, dt)
    rgb = np.zeros(4 * n, dt)
    f = lib().orc_preprocess_f64 if fp64 else lib().orc_preprocess_f32
    f(np.ascontiguousarray(packed, np.float32), n, np.ascontiguousarray(uni, np.float32), tile,
      tiles, meta, conic, xy, rgb)
    return dict(tiles=tiles, meta=meta.reshape(n, 4), conic=conic.reshape(n, 4),
                xy=xy.reshape(n, 2), rgb=rgb.reshape(n, 4))


def scan(counts):
    counts = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros_like(counts)
    total = lib().orc_scan(counts, counts.size, out)
    return out, int(total)


def emit(offsets, meta, n, width, tile, cap):
    offsets = np.ascontiguousarray(offsets, np.uint32)
    m = min(int(offsets[n]) if offsets.size > n else 0, cap)
    keys = np.zeros(max(m, 1), np.uint64)
    vals = np.zeros(max(m, 1), np.uint32)
    lib().orc_emit(offsets, np.ascontiguousarray(meta, np.uint32).ravel(), n, width, tile, cap, keys, vals)
    return keys[:m], vals[:m]


def sort_stable(keys, vals):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_sort_stable(k, v, k.size)
    return k, v


def sort_radix(keys, vals, end_bit):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_sort_radix(k, v, k.size, end_bit)
    return k, v


def tile_ranges(keys, num_tiles):
    k = np.ascontiguousarray(keys, np.uint64)
    r = np.zeros(2 * num_tiles, np.uint32)
    lib().orc_tile_ranges(k, k.size, num_tiles, r)
    return r


def raster(width, height, tile, ranges, vals, conic, xy, rgb, fp64=False):
    dt = np.float64 if fp64 else np.float32
    out = np.zeros(3 * width * height, dt)
    tf = np.zeros(width * height, dt)
    last = np.zeros(width * height, np.uint32)
    f = lib().orc_raster_f64 if fp64 else lib().orc_raster_f32
    f(width, height, tile, np.ascontiguousarray(ranges, np.uint32), np.ascontiguousarray(vals, np.uint32),
      np.ascontiguousarray(conic, dt).ravel(), np.ascontiguousarray(xy, dt).ravel(),
      np.ascontiguousarray(rgb, dt).ravel(), out, tf, last)
    return out.reshape(3, height, width), tf.reshape(height, width), last.reshape(height, width)


def render(packed, n, uni, tile=16, cap=0, fp64=False, use_radix=False, signature=False, evals=None):
    """Whole-frame oracle.  Returns (rgb[3,H,W], T_final[H,W], last[H,W], instances, stage_ms);
    with signature=True (FP64 only) the last item is instead the hash of every
    discrete decision of the frame."""
    W, H = int(uni[40]), int(uni[41])
    dt = np.float64 if fp64 else np.float32
    out = np.zeros(3 * W * H, dt)
    tf = np.zeros(W * H, dt)
    last = np.zeros(W * H, np.uint32)
    inst = np.zeros(1, np.uint64)
    ms = np.zeros(6, np.float64)
    uni = np.ascontiguousarray(uni, np.float32)
    packed = np.ascontiguousarray(packed, np.float32)
    if fp64:
        sig = np.zeros(1, np.uint64)
        lib().orc_render_f64(packed, n, uni, tile, cap, out, tf, last, inst, sig.ctypes.data if signature else None)
        if signature:
            ms = int(sig[0])
    else:
        ev = np.zeros(1, np.uint64)
        lib().orc_render_f32(packed, n, uni, tile, cap, out, tf, last, inst, ms, 1 if use_radix else 0, ev.ctypes.data)
        if evals is not None:
            evals.append(int(ev[0]))
    return (out.reshape(3, H, W), tf.reshape(H, W), last.reshape(H, W), int(inst[0]), ms)


def backward(packed, n, uni, dLdC, tile=16, decisions_f32=True, want_image=False, want_mass=False):
    """FP64 adjoint.  Returns (grad planar [59,N], grad2d [N,12], image [3,H,W] or None), plus
    with want_mass=True (grad2d_mass [N,12], param_mass [59,N]): the summed absolute values of
    the terms each gradient adds up (the scale its rounding error is relative to)."""
    W, H = int(uni[40]), int(uni[41])
    packed = np.ascontiguousarray(packed, np.float32)
    uni = np.ascontiguousarray(uni, np.float32)
    grad = np.zeros(59 * n, np.float64)
    g2 = np.zeros(12 * n, np.float64)
    img = np.zeros(3 * W * H, np.float64) if want_image else None
    m2 = np.zeros(12 * n, np.float64) if want_mass else None
    lib().orc_backward_f64(packed, n, uni, tile, np.ascontiguousarray(dLdC, np.float64).ravel(),
                           1 if decisions_f32 else 0, grad, g2, img.ctypes.data if want_image else None,
                           m2.ctypes.data if want_mass else None)
    out = (grad.reshape(59, n), g2.reshape(n, 12), (img.reshape(3, H, W) if want_image else None))
    if want_mass:
        pm = np.zeros(59 * n, np.float64)
        lib().orc_param_mass_f64(packed, n, uni, m2, pm)
        out = out + (m2.reshape(n, 12), pm.reshape(59, n))
    return out


def preprocess_backward(packed, n, uni, grad2d):
    grad = np.zeros(59 * n, np.float64)
    lib().orc_preprocess_backward_f64(np.ascontiguousarray(packed, np.float32), n, np.ascontiguousarray(uni, np.float32),
                                      np.ascontiguousarray(grad2d, np.float64).ravel(), grad)
    return grad.reshape(59, n)


def finite_differences(packed, n, uni, dLdC, splats, eps=1e-4, tile=16):
    """FP64 central differences of sum(dLdC * image) for every parameter of the
    given splats.  Returns (fd [59, k], same_branch [59, k] bool)."""
    splats = np.ascontiguousarray(splats, np.uint32)
    k = splats.size
    fd = np.zeros(k * 59, np.float64)
    same = np.zeros(k * 59, np.uint8)
    lib().orc_fd_f64(np.ascontiguousarray(packed, np.float32), n, np.ascontiguousarray(uni, np.float32), tile,
                     np.ascontiguousarray(dLdC, np.float64).ravel(), splats, k, eps, fd, same)
    return fd.reshape(k, 59).T, same.reshape(k, 59).T.astype(bool)


def finite_differences_window(packed, n, uni, dLdC, splats, eps=1e-4, tile=16):
    """finite_differences with each perturbed frame rendered only over the pixels of the
    splat's tiles (base or perturbed) where either perturbed splat can blend: the same
    central differences at a cost set by the splat's footprint, not the frame.
    Returns (fd [59, k], same_branch [59, k])."""
    splats = np.ascontiguousarray(splats, np.uint32)
    k = splats.size
    fd = np.zeros(k * 59, np.float64)
    same = np.zeros(k * 59, np.uint8)
    lib().orc_fd_window_f64(np.ascontiguousarray(packed, np.float32), n, np.ascontiguousarray(uni, np.float32), tile,
                            np.ascontiguousarray(dLdC, np.float64).ravel(), splats, k, eps, fd, same)
    return fd.reshape(k, 59).T, same.reshape(k, 59).T.astype(bool)


def raster_backward_f32(W, H, tile, ranges, vals, conic, xy, rgb, tfinal, last, dLdC, n):
    """FP32 replay of the GPU raster adjoint (diagnostic twin); returns grad2d [n, 12] (double sums)."""
    g2 = np.zeros(12 * n, np.float64)
    lib().orc_raster_backward_f32(W, H, tile, np.ascontiguousarray(ranges, np.uint32), np.ascontiguousarray(vals, np.uint32),
                                  np.ascontiguousarray(conic, np.float32).ravel(), np.ascontiguousarray(xy, np.float32).ravel(),
                                  np.ascontiguousarray(rgb, np.float32).ravel(), np.ascontiguousarray(tfinal, np.float32).ravel(),
                                  np.ascontiguousarray(last, np.uint32).ravel(), np.ascontiguousarray(dLdC, np.float32).ravel(),
                                  n, g2)
    return g2.reshape(n, 12)
