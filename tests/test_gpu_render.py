"""Whole-frame forward through the C ABI against the whole-frame oracle.

The frame (preprocess -> scan -> emit -> sort -> ranges -> raster on the GPU)
must reproduce the oracle's FP32 image: within the stated 1e-4 max-abs and, by
construction, bit for bit; the instance count must match exactly.  The FP64
oracle's error is reported."""
import sys

import numpy as np
import pytest

from conftest import empty_dev, to_dev, to_host
from gsr import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_render_small_matches_oracle(torch_cuda, ctx, orc):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    ctx.scene_upload("small", sc.packed(), sc.count)
    ctx.scene_select("small")
    img = ctx.render(u, lib.options(timing=True))
    ref, Tref, lref, I, _ = orc.render(sc.packed(), sc.count, u)
    ref64, _, _, _, _ = orc.render(sc.packed(), sc.count, u, fp64=True)
    assert ctx.instance_count() == I == 84916
    assert np.abs(img - ref).max() <= 1e-4
    assert np.array_equal(img, ref)
    err64 = np.abs(img.astype(np.float64) - ref64)
    _say(f"small: FP64 oracle error max {err64.max():.3e} mean {err64.mean():.3e}")
    rows = ctx.timings()
    assert {"preprocess", "offset scan", "emit", "sort", "tile ranges", "raster", "TOTAL"} <= set(rows)
    assert rows["TOTAL"] > 0
    # timing mode 2: events around the raster only, same image
    img2 = ctx.render(u, lib.options(timing=2))
    assert np.array_equal(img2, img)
    rows2 = ctx.timings()
    assert set(rows2) == {"raster", "TOTAL"}
    assert rows2["raster"] > 0


def test_forward_from_device_params(torch_cuda, ctx, orc):
    sc, cam = scenes.workload_small()
    cam.width, cam.height = 333, 217
    u = cam.uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar())
    out = empty_dev(torch_cuda, 3 * 333 * 217, np.float32)
    torch_cuda.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(save_state=True), out)
    ctx.synchronize()
    ref, _, _, I, _ = orc.render(sc.packed(), sc.count, u)
    assert np.array_equal(to_host(out, np.float32).reshape(3, 217, 333), ref)


@pytest.mark.parametrize("tile", [8, 32])
def test_render_other_tile_sizes(torch_cuda, ctx, orc, tile):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    ctx.scene_upload("small", sc.packed(), sc.count)
    ctx.scene_select("small")
    img = ctx.render(u, lib.options(tile_size=tile))
    ref, _, _, I, _ = orc.render(sc.packed(), sc.count, u, tile=tile)
    assert np.array_equal(img, ref)


def test_render_64bit_keys_equal_32bit(torch_cuda, ctx):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    ctx.scene_upload("small", sc.packed(), sc.count)
    a = ctx.render(u, lib.options(key_bytes=lib.KEY_32))
    b = ctx.render(u, lib.options(key_bytes=lib.KEY_64))
    assert np.array_equal(a, b)


def test_render_1m_1080p_matches_oracle(torch_cuda, ctx, orc):
    sc, cam = scenes.workload_1m_1080p()
    u = cam.uniform(sc.count)
    _say("render 1M splats 1920x1080")
    ctx.scene_upload("big", sc.packed(), sc.count)
    ctx.scene_select("big")
    img = ctx.render(u)
    I = ctx.instance_count()
    ref, _, _, Iref, _ = orc.render(sc.packed(), sc.count, u, use_radix=True)
    _say(f"1M: instances gpu {I} oracle {Iref}; max abs {np.abs(img - ref).max():.3e}")
    assert I == Iref
    assert np.abs(img - ref).max() <= 1e-4
    assert np.array_equal(img, ref)
    ctx.scene_free("big")


@pytest.mark.parametrize("key_bytes", [lib.KEY_32, lib.KEY_64])
def test_sort_geometries_give_the_same_frame(torch_cuda, ctx, orc, key_bytes):
    """Onesweep with 1, 2 and 4 blocks per CU (1024-, 512-, 256-thread blocks,
    different tile sizes and look-back chains) and 8- to 11-bit digits (4 or
    3 passes) must give the same frame."""
    sc = scenes.clustered_scene(300_000, seed=12)
    cam = scenes.Camera((0.5, -0.3, -7.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 55.0, 1280, 720)
    u = cam.uniform(sc.count)
    ctx.scene_upload("geo", sc.packed(), sc.count)
    ctx.scene_select("geo")
    imgs = {}
    for digits in (8, 9, 10, 11):
        for blocks in (1, 2, 4):
            _say(f"sort geometry: {blocks} blocks/CU, {digits}-bit digits, key bytes {key_bytes}")
            imgs[blocks, digits] = ctx.render(u, lib.options(key_bytes=key_bytes, sort_blocks_per_cu=blocks,
                                                             sort_digit_bits=digits))
            assert np.array_equal(imgs[blocks, digits], imgs[1, 8]), (blocks, digits)
    imgs[1] = imgs[1, 8]
    ref, _, _, Iref, _ = orc.render(sc.packed(), sc.count, u, use_radix=True)
    assert ctx.instance_count() == Iref
    assert np.array_equal(imgs[1], ref)
    for bad in (lib.options(sort_blocks_per_cu=3), lib.options(sort_digit_bits=12)):
        with pytest.raises(lib.GsrError) as e:
            ctx.render(u, bad)
        assert e.value.code == lib.GSR_ERR_ARGUMENT
    ctx.scene_free("geo")


def test_capacity_overflow_is_reported(torch_cuda, ctx):
    sc, cam = scenes.workload_small()
    ctx.scene_upload("small", sc.packed(), sc.count)
    ctx.scene_select("small")
    with pytest.raises(lib.GsrError) as e:
        ctx.render(cam.uniform(sc.count), lib.options(capacity=1000))
    assert e.value.code == lib.GSR_ERR_CAPACITY
    assert "84916" in str(e.value)
    img = ctx.render(cam.uniform(sc.count))     # the context keeps working
    assert np.isfinite(img).all()


def test_scene_lifecycle(torch_cuda, ctx):
    sc, cam = scenes.workload_small()
    ctx.scene_upload("a", sc.packed(), sc.count)
    ctx.scene_upload("b", sc.packed(), sc.count)
    ctx.scene_select("b")
    with pytest.raises(lib.GsrError) as e:
        ctx.scene_select("nope")
    assert e.value.code == lib.GSR_ERR_NOT_FOUND
    ctx.scene_free("a")
    with pytest.raises(lib.GsrError):
        ctx.scene_free("a")
    p, n = ctx.scene_params("b")
    assert p and n == sc.count


def test_rgba8_conversion(torch_cuda, ctx):
    W, H = 40, 30
    rgb = np.random.default_rng(0).uniform(-0.5, 1.5, size=(3, H, W)).astype(np.float32)
    d = to_dev(torch_cuda, rgb.ravel())
    torch_cuda.cuda.synchronize()
    out = ctx.to_rgba8(d, W, H, flip_y=False)
    ref = np.rint(np.clip(rgb, 0, 1) * 255).astype(np.uint8).transpose(1, 2, 0)
    assert np.array_equal(out[..., :3], ref) and (out[..., 3] == 255).all()
    flipped = ctx.to_rgba8(d, W, H, flip_y=True)
    assert np.array_equal(flipped, out[::-1])


def test_empty_scene_renders_black(torch_cuda, ctx):
    """Zero splats: a black frame through every entry point (the reference
    renders an empty scene as background), T_final 1 and no last contributor
    for the backward, which has no gradients to produce."""
    cam = scenes.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 200, 120)
    u = cam.uniform(0)
    ctx.scene_upload("empty", np.zeros(0, np.float32), 0)
    ctx.scene_select("empty")
    img = ctx.render(u, lib.options())
    assert img.shape == (3, 120, 200) and not img.any()
    assert ctx.instance_count() == 0
    out = torch_cuda.full((3 * 120 * 200,), 7.0, device="cuda")
    torch_cuda.cuda.synchronize()
    ctx.forward(None, 0, u, lib.options(save_state=True), out)
    ctx.backward(None, 0, u, torch_cuda.zeros(3 * 120 * 200, device="cuda"), None)
    ctx.synchronize()
    assert not out.any().item()
    us = [u, scenes.Camera((1.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, 64, 48).uniform(0)]
    offs = ctx.batch_layout(0, us)
    bout = torch_cuda.full((offs[-1],), 7.0, device="cuda")
    torch_cuda.cuda.synchronize()
    ctx.render_batch(None, 0, us, lib.options(save_state=True), bout)
    ctx.backward_batch(None, 0, us, torch_cuda.zeros(offs[-1], device="cuda"), None)
    ctx.synchronize()
    assert not bout.any().item() and ctx.batch_instance_count() == 0
    ctx.scene_free("empty")


def test_render_4k_8px_tiles_17bit_tile_ids(torch_cuda, ctx, orc):
    """3840x2160 at 8-pixel tiles: 129,600 tiles, so tile ids need 17 bits and
    the frame sorts 64-bit (tile << 16 | depth16) keys on its own (33 bits).
    Image bit-exact and instance count exact against the FP32 oracle, which
    builds the same 64-bit keys (emit.wgsl:62)."""
    sc = scenes.clustered_scene(200_000, seed=4)
    cam = scenes.Camera((0.0, 0.0, -8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, 3840, 2160)
    u = cam.uniform(sc.count)
    _say("render 200k splats 3840x2160 tile 8")
    ctx.scene_upload("k4", sc.packed(), sc.count)
    ctx.scene_select("k4")
    img = ctx.render(u, lib.options(tile_size=8))
    ref, _, _, I, _ = orc.render(sc.packed(), sc.count, u, tile=8)
    assert ctx.instance_count() == I
    assert np.array_equal(img, ref), np.abs(img - ref).max()
    ctx.scene_free("k4")


def _tile_loads(orc, sc, u, tile=16):
    """Instances per tile of the oracle's frame (preprocess -> scan -> emit -> sort -> ranges)."""
    W, H = int(u[40]), int(u[41])
    tx, ty = (W + tile - 1) // tile, (H + tile - 1) // tile
    pre = orc.preprocess(sc.packed(), sc.count, u, tile)
    offs, total = orc.scan(pre["tiles"])
    offsets = np.append(offs, np.uint32(total))
    keys, vals = orc.emit(offsets, pre["meta"], sc.count, W, tile, total)
    keys, vals = orc.sort_stable(keys, vals)
    r = orc.tile_ranges(keys, tx * ty).reshape(-1, 2)
    return (r[:, 1] - r[:, 0]).astype(np.int64)


def test_render_deep_tiles_matches_oracle(torch_cuda, ctx, orc):
    """24,000 splats around the centre of a 160x128 frame: 4 central tiles
    hold more than 16384 instances each (up to 23,997), 32 between 1025 and
    16384, the rest at most 1024.  The image, and the instance count, must
    equal the oracle's frame."""
    sc = scenes.uniform_scene(24000, seed=5, extent=0.15, scale_mu=-2.5, scale_sigma=0.8)
    cam = scenes.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 160, 128)
    u = cam.uniform(sc.count)
    loads = _tile_loads(orc, sc, u)
    _say(f"deep tiles: max {loads.max()}, tiles above 16384: {(loads > 16384).sum()}, "
         f"in (1024, 16384]: {((loads > 1024) & (loads <= 16384)).sum()}, at most 1024: {(loads <= 1024).sum()}")
    assert loads.max() > 16384
    assert ((loads > 1024) & (loads <= 16384)).any() and ((loads > 1) & (loads <= 1024)).any()
    ctx.scene_upload("deep", sc.packed(), sc.count)
    ctx.scene_select("deep")
    img = ctx.render(u)
    ref, _, _, I, _ = orc.render(sc.packed(), sc.count, u)
    assert ctx.instance_count() == I
    assert np.array_equal(img, ref)
