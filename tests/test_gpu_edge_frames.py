"""Edge frames through the raster, the whole frame and the backward on the GPU.

The edge frame (tests/edge_scenes.py) holds the preprocess edge cases (near
plane, behind the camera, off screen, huge and tiny scales, degenerate
quaternions, saturated opacities) plus near needle splats whose centres lie
thousands of pixels off screen while their squares cover every tile, and
huge near-degenerate covariances.  Every pixel meets the far needles at powers
near -1e7, where raseter.wgsl:84-92 gives exp(power) = 0 and skips them.
Checks:
* gsr_rasterize on the oracle's stage inputs (tiles 8/16/32): image, T_final
  and last contributor bit-exact;
* gsr_render: image bit-exact, instance count exact, and the frame without
  the far needles renders to the same image bit for bit;
* gsr_backward: gradients against the FP32 twin element-wise and against the
  FP64 adjoint norm-wise (tiles 8/16/32)."""
import sys

import numpy as np
import pytest

from conftest import empty_dev, to_dev, to_host
from edge_scenes import edge_frame, powers_in_rect
from gsr import lib, scenes
from test_gpu_backward import G2, GROUPS

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _far_needles(orc, sc, cam, u):
    pre = orc.preprocess(sc.packed(), sc.count, u)
    return [g for g in range(4096, sc.count)
            if pre["tiles"][g] > 0 and np.nanmax(powers_in_rect(pre, g, cam.width, cam.height)) < -8.7e6]


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_edge_frame_raster_stage_bit_exact(torch_cuda, ctx, orc, tile):
    sc, cam = edge_frame()
    u = cam.uniform(sc.count)
    W, H, n = cam.width, cam.height, sc.count
    pre = orc.preprocess(sc.packed(), n, u, tile=tile)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, tile, 1 << 40)
    ks, vs = orc.sort_stable(keys, vals)
    TX, TY = (W + tile - 1) // tile, (H + tile - 1) // tile
    ranges = orc.tile_ranges(ks, TX * TY)
    _say(f"edge raster tile={tile} instances={total}")
    img_o, T_o, last_o = orc.raster(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    d = {k: to_dev(torch_cuda, np.ascontiguousarray(pre[k]).ravel()) for k in ("conic", "xy", "rgb")}
    out = empty_dev(torch_cuda, 3 * W * H, np.float32)
    T = empty_dev(torch_cuda, W * H, np.float32)
    last = empty_dev(torch_cuda, W * H, np.uint32)
    d_r, d_v = to_dev(torch_cuda, ranges), to_dev(torch_cuda, vs)
    torch_cuda.cuda.synchronize()
    ctx.rasterize(W, H, tile, n, vs.size, d_r, d_v, d["conic"], d["xy"], d["rgb"], out, T, last)
    img = to_host(out, np.float32).reshape(3, H, W)
    assert np.array_equal(img, img_o), np.abs(img - img_o).max()
    assert np.array_equal(to_host(T, np.float32).reshape(H, W), T_o)
    assert np.array_equal(to_host(last, np.uint32).reshape(H, W), last_o)


def test_edge_frame_render_bit_exact(torch_cuda, ctx, orc):
    sc, cam = edge_frame()
    u = cam.uniform(sc.count)
    far = _far_needles(orc, sc, cam, u)
    assert len(far) == 15
    _say(f"edge frame render n={sc.count}")
    ctx.scene_upload("edge", sc.packed(), sc.count)
    ctx.scene_select("edge")
    img = ctx.render(u)
    I = ctx.instance_count()
    ref, _, _, Iref, _ = orc.render(sc.packed(), sc.count, u, use_radix=True)
    assert I == Iref
    assert np.array_equal(img, ref), np.abs(img - ref).max()
    keep = np.setdiff1d(np.arange(sc.count), far)
    sub = scenes.Scene(pos=sc.pos[keep], lscale=sc.lscale[keep], rot=sc.rot[keep], dc=sc.dc[keep],
                       opacity=sc.opacity[keep], rest=sc.rest[keep])
    ctx.scene_upload("edge_sub", sub.packed(), sub.count)
    ctx.scene_select("edge_sub")
    img_sub = ctx.render(cam.uniform(sub.count))
    assert ctx.instance_count() == I - 15 * 19 * 13
    assert np.array_equal(img, img_sub)
    ctx.scene_free("edge")
    ctx.scene_free("edge_sub")


def _edge_grads(torch, ctx, sc, u, dLdC, tile):
    params = to_dev(torch, sc.planar())
    out = torch.empty(3 * int(u[40]) * int(u[41]), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    # tile 8: the huge splats alone take 19,000 of the frame's 489,060 instances, above 64 per splat
    ctx.forward(params, sc.count, u, lib.options(tile_size=tile, save_state=True, capacity=1 << 20), out)
    d = to_dev(torch, np.ascontiguousarray(dLdC, np.float32).ravel())
    grad = torch.empty_like(params)
    g2 = torch.empty(sc.count * 12, dtype=torch.float32, device="cuda")
    ctx.backward(params, sc.count, u, d, grad, g2)
    ctx.synchronize()
    return (to_host(grad, np.float32).reshape(59, sc.count).astype(np.float64),
            to_host(g2, np.float32).reshape(sc.count, 12).astype(np.float64),
            to_host(out, np.float32))


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_edge_frame_backward(torch_cuda, ctx, orc, tile):
    """Gradients of the edge frame.  The FP32 twin pins the kernel on every
    element (1e-4 of the mass, as in tests/test_gpu_backward.py).  Against the
    FP64 adjoint the check is norm-wise within 1e-3 per group over the splats
    whose power FP32 resolves (power_conditioning below 1e4): for the diagonal
    needles (2,000 px off screen, conditioning 3e6) FP32 evaluates alpha about
    18% away from FP64, so the FP32 frame's gradient is not the FP64 frame's
    there, and the pixels they cover carry that into their neighbours'
    colour gradients, which is why no element-wise FP64 bound is asserted."""
    from edge_scenes import power_conditioning
    sc, cam = edge_frame()
    u = cam.uniform(sc.count)
    W, H, n = cam.width, cam.height, sc.count
    rng = np.random.default_rng(5)
    dLdC = rng.normal(size=(3, H, W))
    _say(f"edge frame backward tile={tile}")
    grad, g2, img = _edge_grads(torch_cuda, ctx, sc, u, dLdC, tile)
    ref_img, _, _, _, _ = orc.render(sc.packed(), n, u, tile=tile, cap=1 << 20)
    assert np.array_equal(img.reshape(ref_img.shape), ref_img)
    assert np.isfinite(grad).all() and np.isfinite(g2).all()
    ref, ref2, _, m2, pm = orc.backward(sc.packed(), n, u, dLdC, tile=tile, decisions_f32=True, want_mass=True)
    # the FP32 twin on the same forward
    pre = orc.preprocess(sc.packed(), n, u, tile=tile)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, tile, 1 << 40)
    ks, vs = orc.sort_stable(keys, vals)
    TX, TY = (W + tile - 1) // tile, (H + tile - 1) // tile
    ranges = orc.tile_ranges(ks, TX * TY)
    _, T, last = orc.raster(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    twin = orc.raster_backward_f32(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"], T, last,
                                   np.asarray(dLdC, np.float32), n)
    err = np.abs(g2[:, :9] - twin[:, :9]) / np.maximum(1e-4 * m2[:, :9], 1e-30)
    _say(f"twin: error / (1e-4 mass) max {err.max():.3f}")
    assert err.max() <= 1.0
    ok = power_conditioning(pre, W, H, tile) < 1e4
    _say(f"splats resolved in FP32: {ok.sum()} of {n}")
    assert ok.sum() >= n - 20
    for gname, sl in G2.items():
        a, b = g2[ok][:, sl], ref2[ok][:, sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        _say(f"grad2d/{gname}: norm rel {rel:.2e}")
        assert rel < 1e-3, (gname, rel)
    for gname, sl in GROUPS.items():
        a, b = grad[sl][:, ok], ref[sl][:, ok]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        _say(f"params/{gname}: norm rel {rel:.2e}")
        assert rel < 1e-3, (gname, rel)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_raster_never_culls_an_indefinite_conic(torch_cuda, ctx, orc, tile):
    """A conic with A < 0 and C < 0 (what a covariance whose FP32 determinant
    rounds below zero gives: preprocess.wgsl:103-110 has no guard) is
    indefinite, and pixels off its axes still meet q in (0, q_max].  Splat 0:
    conic (-0.04, 0.11, -0.04), opacity 0.9, centre (0.5, 0.5): pixel (8, 8)
    has q = 7.875 (alpha 0.018) and is blended, but along each edge of the
    block [8, 15]^2 q is concave, so its stationary point is the edge's
    maximum (13.3 > q_max = 11.19): a cull that takes the stationary points
    for the minima would skip the whole block.  The raster must match the
    oracle (which culls nothing) bit for bit.  Splat 1 is an ordinary splat
    behind it."""
    W = H = 32
    conic = np.array([[-0.04, 0.11, -0.04, 0.9], [0.05, 0.01, 0.04, 0.6]], np.float32)
    xy = np.array([[0.5, 0.5], [20.0, 12.0]], np.float32)
    rgb = np.array([[1.0, 0.5, 0.25, 0.0], [0.2, 0.7, 0.3, 0.0]], np.float32)
    TX = (W + tile - 1) // tile
    tiles = TX * TX
    # every tile holds both splats in depth order (0 in front)
    vs = np.tile(np.array([0, 1], np.uint32), tiles)
    ranges = np.stack([np.arange(tiles) * 2, np.arange(tiles) * 2 + 2], 1).astype(np.uint32).ravel()
    img_o, T_o, last_o = orc.raster(W, H, tile, ranges, vs, conic, xy, rgb)
    assert last_o[8, 8] >= 1 and img_o[0, 8:16, 8:16].max() > 0.01   # the block is blended
    d = {k: to_dev(torch_cuda, np.ascontiguousarray(v).ravel()) for k, v in (("conic", conic), ("xy", xy), ("rgb", rgb))}
    out = empty_dev(torch_cuda, 3 * W * H, np.float32)
    T = empty_dev(torch_cuda, W * H, np.float32)
    last = empty_dev(torch_cuda, W * H, np.uint32)
    d_r, d_v = to_dev(torch_cuda, ranges), to_dev(torch_cuda, vs)
    torch_cuda.cuda.synchronize()
    _say(f"indefinite conic raster tile={tile}")
    ctx.rasterize(W, H, tile, 2, vs.size, d_r, d_v, d["conic"], d["xy"], d["rgb"], out, T, last)
    img = to_host(out, np.float32).reshape(3, H, W)
    assert np.array_equal(img, img_o)
    assert np.array_equal(to_host(T, np.float32).reshape(H, W), T_o)
    assert np.array_equal(to_host(last, np.uint32).reshape(H, W), last_o)


def test_indefinite_conic_frame(torch_cuda, ctx, orc):
    """The edge frame plus a pancake whose FP32 conic comes out indefinite
    (A < 0, C < 0; tests/edge_scenes.py): the whole frame bit-exact against
    the FP32 oracle, and its backward (tile 16) against the FP32 twin on
    every element."""
    from edge_scenes import indefinite_frame
    sc, cam = indefinite_frame()
    u = cam.uniform(sc.count)
    W, H, n, tile = cam.width, cam.height, sc.count, 16
    pre = orc.preprocess(sc.packed(), n, u, tile=tile)
    assert pre["conic"][n - 1, 0] < 0 and pre["conic"][n - 1, 2] < 0 and pre["tiles"][n - 1] > 0
    _say(f"indefinite conic frame n={n}")
    ctx.scene_upload("indef", sc.packed(), n)
    ctx.scene_select("indef")
    img = ctx.render(u)
    ref, _, _, Iref, _ = orc.render(sc.packed(), n, u, use_radix=True)
    assert ctx.instance_count() == Iref
    assert np.array_equal(img, ref), np.abs(img - ref).max()
    ctx.scene_free("indef")
    dLdC = np.random.default_rng(9).normal(size=(3, H, W))
    grad, g2, _ = _edge_grads(torch_cuda, ctx, sc, u, dLdC, tile)
    assert np.isfinite(grad).all() and np.isfinite(g2).all()
    _, _, _, m2, _ = orc.backward(sc.packed(), n, u, dLdC, tile=tile, decisions_f32=True, want_mass=True)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, tile, 1 << 40)
    ks, vs = orc.sort_stable(keys, vals)
    ranges = orc.tile_ranges(ks, ((W + tile - 1) // tile) * ((H + tile - 1) // tile))
    _, T, last = orc.raster(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    twin = orc.raster_backward_f32(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"], T, last,
                                   np.asarray(dLdC, np.float32), n)
    assert np.abs(twin[n - 1, :9]).max() > 0   # the pancake is blended and differentiated
    err = np.abs(g2[:, :9] - twin[:, :9]) / np.maximum(1e-4 * m2[:, :9], 1e-30)
    _say(f"twin: error / (1e-4 mass) max {err.max():.3f}")
    assert err.max() <= 1.0
