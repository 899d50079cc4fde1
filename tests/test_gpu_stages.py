"""Stage-wise parity of the HIP kernels against the oracle, through the C ABI.

Each stage receives the oracle's inputs, so integer outputs (counts, rects,
depth keys, offsets, keys/values, ranges) must be bit-exact, and the FP32
outputs must be bit-identical as well: kernels and oracle execute the same
FP32 operation sequence (tolerance kept at the stated 1e-4 for colour)."""
import sys

import numpy as np
import pytest

from conftest import empty_dev, to_dev, to_host
from edge_scenes import edge_scene as _edge_scene
from gsr import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pre_gpu(torch, ctx, sc, u, tile=16):
    n = sc.count
    params = to_dev(torch, sc.planar())
    tiles = empty_dev(torch, n, np.uint32)
    meta = empty_dev(torch, 4 * n, np.uint32)
    conic = empty_dev(torch, 4 * n, np.float32)
    xy = empty_dev(torch, 2 * n, np.float32)
    rgb = empty_dev(torch, 4 * n, np.float32)
    torch.cuda.synchronize()
    ctx.preprocess(params, n, u, tile, tiles, meta, conic, xy, rgb)
    return dict(tiles=to_host(tiles, np.uint32)[:n], meta=to_host(meta, np.uint32)[:4 * n].reshape(n, 4),
                conic=to_host(conic, np.float32)[:4 * n].reshape(n, 4), xy=to_host(xy, np.float32)[:2 * n].reshape(n, 2),
                rgb=to_host(rgb, np.float32)[:4 * n].reshape(n, 4))


@pytest.mark.parametrize("which,tile", [("small", 16), ("edge", 16), ("edge", 8), ("edge", 32)])
def test_preprocess_bit_exact(torch_cuda, ctx, orc, which, tile):
    if which == "small":
        sc, cam = scenes.workload_small()
    else:
        sc = _edge_scene()
        cam = scenes.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 300, 200)
    u = cam.uniform(sc.count)
    _say(f"preprocess {which} tile={tile} n={sc.count}")
    g = _pre_gpu(torch_cuda, ctx, sc, u, tile)
    o = orc.preprocess(sc.packed(), sc.count, u, tile=tile)
    assert np.array_equal(g["tiles"], o["tiles"])
    assert np.array_equal(g["meta"], o["meta"])
    for k in ("conic", "xy", "rgb"):
        assert np.array_equal(g[k], o[k]), (k, np.abs(g[k] - o[k]).max())


@pytest.mark.parametrize("n", [0, 1, 100, 4095, 4096, 4097, 1_000_000, 3_000_001])
def test_scan_exact(torch_cuda, ctx, orc, n):
    _say(f"scan n={n}")
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 64, size=n).astype(np.uint32)
    d_in = to_dev(torch_cuda, counts if n else np.zeros(1, np.uint32))
    d_out = empty_dev(torch_cuda, n + 1, np.uint32)
    d_tot = empty_dev(torch_cuda, 1, np.uint32)
    torch_cuda.cuda.synchronize()
    ctx.scan_offsets(d_in, n, d_out, d_tot)
    out = to_host(d_out, np.uint32)[:n + 1]
    ref, total = orc.scan(np.concatenate([counts, [0]]).astype(np.uint32))
    assert np.array_equal(out, ref)
    assert int(to_host(d_tot, np.uint32)[0]) == total


def test_scan_in_place(torch_cuda, ctx, orc):
    counts = np.random.default_rng(1).integers(0, 9, size=50001).astype(np.uint32)
    buf = to_dev(torch_cuda, np.concatenate([counts, [0]]).astype(np.uint32))
    torch_cuda.cuda.synchronize()
    ctx.scan_offsets(buf, counts.size, buf)
    ref, _ = orc.scan(np.concatenate([counts, [0]]).astype(np.uint32))
    assert np.array_equal(to_host(buf, np.uint32), ref)


def _oracle_chain(orc, sc, u, tile=16):
    n = sc.count
    W = int(u[40])
    pre = orc.preprocess(sc.packed(), n, u, tile=tile)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, tile, 64 * n)
    return pre, offs, total, keys, vals


@pytest.mark.parametrize("key_bytes", [4, 8])
def test_emit_exact(torch_cuda, ctx, orc, key_bytes):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    pre, offs, total, keys, vals = _oracle_chain(orc, sc, u)
    _say(f"emit key_bytes={key_bytes} instances={total}")
    kdt = np.uint32 if key_bytes == 4 else np.uint64
    d_off = to_dev(torch_cuda, offs)
    d_meta = to_dev(torch_cuda, pre["meta"].ravel())
    d_k = empty_dev(torch_cuda, total, kdt)
    d_v = empty_dev(torch_cuda, total, np.uint32)
    torch_cuda.cuda.synchronize()
    ctx.emit(d_off, d_meta, sc.count, 512, 16, 64 * sc.count, key_bytes, d_k, d_v)
    assert np.array_equal(to_host(d_k, kdt)[:total].astype(np.uint64), keys)
    assert np.array_equal(to_host(d_v, np.uint32)[:total], vals)
    # capacity truncation keeps the first slots only (emit.wgsl:57-59)
    cap = total // 3
    d_k2 = empty_dev(torch_cuda, total, kdt)
    torch_cuda.cuda.synchronize()
    ctx.emit(d_off, d_meta, sc.count, 512, 16, cap, key_bytes, d_k2, d_v)
    got = to_host(d_k2, kdt).astype(np.uint64)
    assert np.array_equal(got[:cap], keys[:cap]) and not got[cap:].any()


def _gpu_sort(torch, ctx, keys, vals, key_bytes, begin, end, blocks=None, digits=None):
    n = keys.size
    kdt = np.uint32 if key_bytes == 4 else np.uint64
    k0 = to_dev(torch, keys.astype(kdt) if n else np.zeros(1, kdt))
    v0 = to_dev(torch, vals if n else np.zeros(1, np.uint32))
    k1 = empty_dev(torch, n, kdt)
    v1 = empty_dev(torch, n, np.uint32)
    torch.cuda.synchronize()
    idx = ctx.sort_pairs(key_bytes, k0, v0, k1, v1, n, begin, end, blocks, digits)
    ks = (k1 if idx else k0)
    vs = (v1 if idx else v0)
    return to_host(ks, kdt)[:n].astype(np.uint64), to_host(vs, np.uint32)[:n]


@pytest.mark.parametrize("n,key_bytes,bits", [
    (0, 4, 32), (1, 4, 32), (100, 4, 32), (4095, 4, 29), (4096, 4, 29), (4097, 8, 45),
    (100_000, 4, 32), (1_000_003, 4, 29), (3_000_000, 8, 45), (2_000_000, 8, 64)])
def test_sort_stable_exact(torch_cuda, ctx, n, key_bytes, bits):
    _say(f"sort n={n} key_bytes={key_bytes} bits={bits}")
    rng = np.random.default_rng(n + bits)
    hi = (1 << bits) - 1 if bits < 64 else np.iinfo(np.uint64).max
    keys = rng.integers(0, hi, size=n, dtype=np.uint64, endpoint=True)
    # heavy ties: half the keys drawn from a few values
    if n > 10:
        keys[: n // 2] = keys[rng.integers(0, 8, n // 2)]
    vals = rng.permutation(n).astype(np.uint32)
    k, v = _gpu_sort(torch_cuda, ctx, keys, vals, key_bytes, 0, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


@pytest.mark.parametrize("blocks,digits", [(b, d) for d in (8, 9, 10, 11) for b in (1, 2, 4)])
@pytest.mark.parametrize("n,key_bytes,bits,begin", [
    (0, 4, 32, 0), (1, 4, 29, 0), (8193, 8, 33, 0), (300_000, 4, 32, 8),
    (1_000_003, 4, 29, 0), (2_500_000, 8, 34, 0), (600_000, 8, 64, 0)])
def test_sort_geometries_stable_exact(torch_cuda, ctx, blocks, digits, n, key_bytes, bits, begin):
    """Every onesweep geometry (1024/512/256-thread blocks, 8- to 11-bit
    digits up to 11 bits) gives the stable sort on bits [begin, bits)."""
    _say(f"sort geometry blocks={blocks} digits={digits} n={n} key_bytes={key_bytes} bits=[{begin},{bits})")
    rng = np.random.default_rng(n + bits + 7 * digits)
    hi = (1 << bits) - 1 if bits < 64 else np.iinfo(np.uint64).max
    keys = rng.integers(0, hi, size=n, dtype=np.uint64, endpoint=True)
    if n > 10:
        keys[: n // 2] = keys[rng.integers(0, 8, n // 2)]   # heavy ties
    vals = rng.permutation(n).astype(np.uint32)
    k, v = _gpu_sort(torch_cuda, ctx, keys, vals, key_bytes, begin, bits, blocks, digits)
    width = bits - begin
    win = (keys >> np.uint64(begin)) & (np.uint64((1 << width) - 1) if width < 64 else np.iinfo(np.uint64).max)
    order = np.argsort(win, kind="stable")
    assert np.array_equal(v, vals[order])
    assert np.array_equal(k, keys[order])


@pytest.mark.parametrize("n,key_bytes,bits,digits", [
    (1_500_000, 4, 29, 8), (2_200_000, 4, 29, 8), (3_000_000, 4, 29, 8), (3_400_000, 4, 29, 8),
    (4_500_000, 4, 29, 8), (1_300_000, 8, 33, 9), (1_700_000, 8, 33, 9), (2_200_000, 8, 33, 9)])
def test_sort_tile_lengths(torch_cuda, ctx, n, key_bytes, bits, digits):
    """Key counts between one and a few full tiles per resident block, where
    the onesweep tile length follows the device-side count (each of the
    shortened lengths of the default 32- and 64-bit geometries on a 256-CU
    part, and the full one): still the stable sort."""
    _say(f"sort tile lengths n={n} key_bytes={key_bytes} bits={bits} digits={digits}")
    rng = np.random.default_rng(n + digits)
    keys = rng.integers(0, (1 << bits) - 1, size=n, dtype=np.uint64, endpoint=True)
    keys[: n // 3] = keys[rng.integers(0, 64, n // 3)]   # ties
    vals = rng.permutation(n).astype(np.uint32)
    k, v = _gpu_sort(torch_cuda, ctx, keys, vals, key_bytes, 0, bits, 1, digits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v, vals[order])
    assert np.array_equal(k, keys[order])


def test_sort_geometry_rejects_unknown_widths(torch_cuda, ctx):
    keys = np.arange(10, dtype=np.uint64)
    for blocks, digits in ((3, 8), (1, 7), (1, 12)):
        with pytest.raises(lib.GsrError) as e:
            _gpu_sort(torch_cuda, ctx, keys, keys.astype(np.uint32), 4, 0, 32, blocks, digits)
        assert e.value.code == lib.GSR_ERR_ARGUMENT


def test_sort_bit_window_and_all_equal(torch_cuda, ctx):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 32, size=300_000, dtype=np.uint64)
    vals = np.arange(keys.size, dtype=np.uint32)
    # bits [8, 24) only: stable on that window
    k, v = _gpu_sort(torch_cuda, ctx, keys, vals, 4, 8, 24)
    win = (keys >> np.uint64(8)) & np.uint64(0xFFFF)
    order = np.argsort(win, kind="stable")
    assert np.array_equal(v, vals[order])
    same = np.full(70_000, 12345, np.uint64)
    k, v = _gpu_sort(torch_cuda, ctx, same, np.arange(same.size, dtype=np.uint32), 4, 0, 32)
    assert np.array_equal(v, np.arange(same.size, dtype=np.uint32))


def test_sort_emitted_instances_exact(torch_cuda, ctx, orc):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    pre, offs, total, keys, vals = _oracle_chain(orc, sc, u)
    ks, vs = orc.sort_stable(keys, vals)
    for kb in (4, 8):
        k, v = _gpu_sort(torch_cuda, ctx, keys, vals, kb, 0, 26)
        assert np.array_equal(k, ks) and np.array_equal(v, vs)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 100_003])
def test_tile_ranges_chunk_edges(torch_cuda, ctx, orc, n):
    """Key counts at the range kernel's row (64) and chunk (2048) edges, with
    tile runs crossing them and empty tiles between runs."""
    rng = np.random.default_rng(n)
    num_tiles = 300
    tiles = np.sort(rng.integers(0, num_tiles, size=n)).astype(np.uint64)
    ks = (tiles << np.uint64(16)) | rng.integers(0, 1 << 16, size=n).astype(np.uint64)
    ks.sort()
    ref = orc.tile_ranges(ks, num_tiles)
    for kb, kdt in ((4, np.uint32), (8, np.uint64)):
        d_k = to_dev(torch_cuda, ks.astype(kdt))
        d_r = to_dev(torch_cuda, np.full(2 * num_tiles, 7, np.uint32))
        torch_cuda.cuda.synchronize()
        ctx.tile_ranges(d_k, kb, n, num_tiles, d_r)
        assert np.array_equal(to_host(d_r, np.uint32), ref), (n, kb)


@pytest.mark.parametrize("key_bytes", [4, 8])
def test_tile_ranges_exact(torch_cuda, ctx, orc, key_bytes):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    pre, offs, total, keys, vals = _oracle_chain(orc, sc, u)
    ks, vs = orc.sort_stable(keys, vals)
    kdt = np.uint32 if key_bytes == 4 else np.uint64
    ref = orc.tile_ranges(ks, 32 * 32)
    d_k = to_dev(torch_cuda, ks.astype(kdt))
    d_r = to_dev(torch_cuda, np.full(2 * 1024, 7, np.uint32))  # stale contents must be cleared
    torch_cuda.cuda.synchronize()
    ctx.tile_ranges(d_k, key_bytes, ks.size, 1024, d_r)
    assert np.array_equal(to_host(d_r, np.uint32), ref)
    ctx.tile_ranges(d_k, key_bytes, 0, 1024, d_r)
    assert not to_host(d_r, np.uint32).any()


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_raster_bit_exact(torch_cuda, ctx, orc, tile):
    sc, cam = scenes.workload_small()
    cam.width, cam.height = 500, 380      # partial edge tiles
    u = cam.uniform(sc.count)
    W, H = 500, 380
    pre, offs, total, keys, vals = _oracle_chain(orc, sc, u, tile)
    ks, vs = orc.sort_stable(keys, vals)
    TX, TY = (W + tile - 1) // tile, (H + tile - 1) // tile
    ranges = orc.tile_ranges(ks, TX * TY)
    _say(f"raster tile={tile} instances={total}")
    img_o, T_o, last_o = orc.raster(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    d = {k: to_dev(torch_cuda, np.ascontiguousarray(pre[k]).ravel()) for k in ("conic", "xy", "rgb")}
    d_r = to_dev(torch_cuda, ranges)
    d_v = to_dev(torch_cuda, vs)
    out = empty_dev(torch_cuda, 3 * W * H, np.float32)
    T = empty_dev(torch_cuda, W * H, np.float32)
    last = empty_dev(torch_cuda, W * H, np.uint32)
    torch_cuda.cuda.synchronize()
    ctx.rasterize(W, H, tile, sc.count, vs.size, d_r, d_v, d["conic"], d["xy"], d["rgb"], out, T, last)
    img = to_host(out, np.float32).reshape(3, H, W)
    assert np.abs(img - img_o).max() <= 1e-4
    assert np.array_equal(img, img_o)
    assert np.array_equal(to_host(T, np.float32).reshape(H, W), T_o)
    assert np.array_equal(to_host(last, np.uint32).reshape(H, W), last_o)


def test_raster_without_splats_is_black(torch_cuda, ctx):
    """gsr_rasterize with no splats but ranges that name instances (values
    that cannot name a splat): every pixel keeps T = 1 and no contributor,
    and no record is read."""
    W, H, tile = 40, 24, 16
    TX, TY = (W + tile - 1) // tile, (H + tile - 1) // tile
    vs = np.array([3, 9, 1, 7], np.uint32)
    ranges = np.zeros((TX * TY, 2), np.uint32)
    ranges[0] = (0, 4)
    ranges[TX * TY - 1] = (2, 4)
    empty = to_dev(torch_cuda, np.zeros(4, np.float32))
    out = empty_dev(torch_cuda, 3 * W * H, np.float32)
    T = empty_dev(torch_cuda, W * H, np.float32)
    last = empty_dev(torch_cuda, W * H, np.uint32)
    d_r, d_v = to_dev(torch_cuda, ranges.ravel()), to_dev(torch_cuda, vs)
    torch_cuda.cuda.synchronize()
    _say("raster without splats")
    ctx.rasterize(W, H, tile, 0, vs.size, d_r, d_v, empty, empty, empty, out, T, last)
    assert not to_host(out, np.float32).any()
    assert (to_host(T, np.float32) == 1.0).all()
    assert not to_host(last, np.uint32).any()


@pytest.mark.parametrize("n,key_bytes,bits", [((1 << 30) - 1, 4, 29), (300_000_001, 8, 34)])
def test_sort_at_the_key_limit(torch_cuda, ctx, n, key_bytes, bits):
    """The sort at its documented limit (2^30 - 1 keys: the look-back words
    carry 30-bit counts) and 300M 64-bit keys, generated and checked on the
    GPU: keys and values equal torch's stable sort (values are the input
    positions, so ties must keep input order), and one key more is refused."""
    torch = torch_cuda
    _say(f"sort at the limit n={n} key_bytes={key_bytes} bits={bits}")
    g = torch.Generator(device="cuda").manual_seed(n % 1000)
    kt = torch.int32 if key_bytes == 4 else torch.int64
    keys = torch.randint(0, 1 << bits, (n,), dtype=kt, device="cuda", generator=g)
    keys[: n // 4] = keys[: n // 4] & 0xFF  # many ties
    k1 = torch.empty_like(keys)
    v0 = torch.arange(n, dtype=torch.int32, device="cuda")
    v1 = torch.empty_like(v0)
    k0 = keys.clone()
    torch.cuda.synchronize()
    idx = ctx.sort_pairs(key_bytes, k0, v0, k1, v1, n, 0, bits)
    ks, vs = (k1, v1) if idx else (k0, v0)
    ref_k, ref_i = torch.sort(keys, stable=True)
    assert torch.equal(ks, ref_k)
    assert torch.equal(vs, ref_i.to(torch.int32))
    del ref_k, ref_i, ks, vs, k0, k1, v0, v1, keys
    torch.cuda.empty_cache()
    if key_bytes == 4:  # refused before any buffer is touched
        d = torch.empty(1, dtype=torch.int32, device="cuda")
        with pytest.raises(lib.GsrError, match="exceed the sort limit"):
            ctx.sort_pairs(4, d, d, d, d, 1 << 30, 0, 32)
