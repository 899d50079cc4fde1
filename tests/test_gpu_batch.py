"""Multi-view batches: one view|tile|depth sort for every camera, a persistent
(view, tile) raster queue, hipGraph replays and the autotuner.

A batch view must be bit-identical to the plain single-view render of the
same camera at the same tile size: both run the same per-splat arithmetic,
and ties in (tile, depth16) resolve by splat index in both (emit order).
Graph replays are checked against eager renders, with eager work and
synchronisations between replays, including replays with new cameras."""
import sys

import numpy as np
import pytest

from conftest import to_dev, to_host
from gsr import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _plain(torch, ctx, params, n, u, tile):
    W, H = int(u[40]), int(u[41])
    out = torch.empty(3 * W * H, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.forward(params, n, u, lib.options(tile_size=tile), out)
    ctx.synchronize()
    return to_host(out, np.float32).reshape(3, H, W)


def _batch(torch, ctx, params, n, us, opts, use_graph=False, out=None):
    offs = ctx.batch_layout(n, us, opts)
    if out is None:
        out = torch.full((offs[-1],), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.render_batch(params, n, us, opts, out, use_graph=use_graph)
    host = to_host(out, np.float32)
    imgs = []
    for v, u in enumerate(us):
        W, H = int(u[40]), int(u[41])
        imgs.append(host[offs[v]:offs[v] + 3 * W * H].reshape(3, H, W))
    return imgs, out


def _scene_and_cams(n=40_000, views=8, seed=5):
    sc = scenes.clustered_scene(n, seed=seed)
    cams = scenes.orbit_cameras(views, seed=seed)
    return sc, [c.uniform(n) for c in cams]


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batch_views_equal_plain_renders(torch_cuda, ctx, tile):
    sc, us = _scene_and_cams()
    params = to_dev(torch_cuda, sc.planar())
    _say(f"batch tile={tile} views={len(us)}")
    imgs, _ = _batch(torch_cuda, ctx, params, sc.count, us, lib.options(tile_size=tile))
    for v, u in enumerate(us):
        ref = _plain(torch_cuda, ctx, params, sc.count, u, tile)
        assert np.array_equal(imgs[v], ref), (v, np.abs(imgs[v] - ref).max())
    # every view's instances add up to the batch's
    total = 0
    for u in us:
        _plain(torch_cuda, ctx, params, sc.count, u, tile)
        total += ctx.instance_count()
    assert ctx.batch_instance_count() == total


@pytest.mark.parametrize("blocks,digits", [(2, 8), (4, 8), (1, 9), (2, 10), (1, 11), (4, 11)])
def test_batch_sort_geometries_agree(torch_cuda, ctx, blocks, digits):
    """64-bit view|tile|depth keys through onesweep at 1, 2 and 4 blocks per CU
    and 8- to 11-bit digits give the batch of the default (1 block per CU,
    8-bit digits); the saved state for the backward follows the sort's
    result buffer."""
    sc, us = _scene_and_cams(n=60_000, views=12, seed=13)
    params = to_dev(torch_cuda, sc.planar())
    want, _ = _batch(torch_cuda, ctx, params, sc.count, us, lib.options(tile_size=16, key_bytes=lib.KEY_64))
    got, _ = _batch(torch_cuda, ctx, params, sc.count, us,
                    lib.options(tile_size=16, key_bytes=lib.KEY_64, sort_blocks_per_cu=blocks,
                                sort_digit_bits=digits))
    for v in range(len(us)):
        assert np.array_equal(got[v], want[v]), v


def test_batch_matches_oracle_for_one_view(torch_cuda, ctx, orc):
    sc, us = _scene_and_cams(n=20_000, views=4, seed=8)
    params = to_dev(torch_cuda, sc.planar())
    imgs, _ = _batch(torch_cuda, ctx, params, sc.count, us, lib.options(tile_size=16))
    for v in (0, 3):
        ref, _, _, _, _ = orc.render(sc.packed(), sc.count, us[v])
        assert np.array_equal(imgs[v], ref)


def test_graph_replay_equals_eager(torch_cuda, ctx):
    sc, us = _scene_and_cams(n=30_000, views=6, seed=9)
    params = to_dev(torch_cuda, sc.planar())
    opts = lib.options(tile_size=16)
    eager, _ = _batch(torch_cuda, ctx, params, sc.count, us, opts)
    offs = ctx.batch_layout(sc.count, us, opts)
    out = torch_cuda.empty(offs[-1], dtype=torch_cuda.float32, device="cuda")
    for rep in range(4):   # first call captures, the rest replay
        out.fill_(-1.0)
        imgs, _ = _batch(torch_cuda, ctx, params, sc.count, us, opts, use_graph=True, out=out)
        for v in range(len(us)):
            assert np.array_equal(imgs[v], eager[v]), (rep, v)
        _plain(torch_cuda, ctx, params, sc.count, us[0], 16)     # eager work between replays
    # new cameras with the same resolutions replay the same graph
    rng = np.random.default_rng(1)
    moved = []
    for u in us:
        eye = tuple(float(v) for v in rng.uniform(-1, 1, 3) + np.array([0.0, 0.0, -7.0]))
        cam = scenes.Camera(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 55.0, int(u[40]), int(u[41]))
        moved.append(cam.uniform(sc.count))
    want, _ = _batch(torch_cuda, ctx, params, sc.count, moved, opts)
    out.fill_(-1.0)
    got, _ = _batch(torch_cuda, ctx, params, sc.count, moved, opts, use_graph=True, out=out)
    for v in range(len(us)):
        assert np.array_equal(got[v], want[v]), v


def test_autotuner_picks_a_valid_configuration(torch_cuda, ctx):
    sc, us = _scene_and_cams(n=20_000, views=4, seed=10)
    params = to_dev(torch_cuda, sc.planar())
    offs = ctx.batch_layout(sc.count, us, lib.options(tile_size=8))
    scratch = torch_cuda.empty(offs[-1], dtype=torch_cuda.float32, device="cuda")
    opts, best = ctx.autotune_batch(params, sc.count, us, lib.options(), scratch)
    _say(f"autotuned: tile {opts.tile_size} waves {opts.raster_waves_per_cu} blocks {opts.sort_blocks_per_cu} "
         f"{best:.3f} ms")
    assert opts.tile_size in (8, 16, 32) and opts.raster_waves_per_cu in (8, 16, 32)
    assert opts.sort_blocks_per_cu in (1, 2, 4) and opts.sort_digit_bits in (8, 9, 10, 11) and best > 0
    imgs, _ = _batch(torch_cuda, ctx, params, sc.count, us, opts)
    ref = _plain(torch_cuda, ctx, params, sc.count, us[1], opts.tile_size)
    assert np.array_equal(imgs[1], ref)


def test_batch_capacity(torch_cuda, ctx):
    sc, us = _scene_and_cams(n=20_000, views=3, seed=11)
    params = to_dev(torch_cuda, sc.planar())
    offs = ctx.batch_layout(sc.count, us, lib.options())
    out = torch_cuda.empty(offs[-1], dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(lib.GsrError) as e:
        ctx.render_batch(params, sc.count, us, lib.options(capacity=500), out)
    assert e.value.code == lib.GSR_ERR_CAPACITY
    ctx.render_batch(params, sc.count, us, lib.options(), out)   # planner sizes it
    assert ctx.batch_instance_count() > 500


def test_batch_64_views_500k(torch_cuda, ctx):
    sc, cams = scenes.workload_batch64()
    us = [c.uniform(sc.count) for c in cams]
    params = to_dev(torch_cuda, sc.planar())
    _say("batch 64 views x 500k splats")
    imgs, _ = _batch(torch_cuda, ctx, params, sc.count, us, lib.options(tile_size=16), use_graph=True)
    _say(f"instances {ctx.batch_instance_count()}")
    for v in (0, 17, 63):
        ref = _plain(torch_cuda, ctx, params, sc.count, us[v], 16)
        assert np.array_equal(imgs[v], ref), v
    assert all(np.isfinite(i).all() for i in imgs)
