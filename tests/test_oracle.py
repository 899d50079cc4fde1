"""Pins of the CPU oracle (no GPU): every oracle function is checked against a
closed form, brute force, a special case or its FP64 twin, chosen so that a
dropped term, a wrong sign or a transposed operand fails one of them."""
import math

import numpy as np
import pytest

from gsr import scenes


def _identity_splat_scene(n=1, pos=(0.0, 0.0, 0.0), lscale=-2.0, opacity=0.0, dc=(0.0, 0.0, 0.0), rest=None):
    s = scenes.Scene(pos=np.array([pos] * n, np.float32), lscale=np.full((n, 3), lscale, np.float32),
                     rot=np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1)),
                     dc=np.array([dc] * n, np.float32), opacity=np.full(n, opacity, np.float32),
                     rest=np.zeros((n, 45), np.float32) if rest is None else rest.astype(np.float32))
    return s


def test_exp_matches_libm(orc):
    x = np.linspace(-87.0, 88.7, 400001, dtype=np.float32)
    y = orc.exp_f32(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert (np.abs(y - ref) / ulp).max() < 2.0
    assert orc.exp_f32(np.array([0.0], np.float32))[0] == 1.0
    assert orc.exp_f32(np.array([-100.0], np.float32))[0] == 0.0
    assert np.isinf(orc.exp_f32(np.array([89.0], np.float32))[0])


def test_exp_blend_matches_libm_on_the_blend_range(orc):
    # every 7th float of [-5.55, 0] (about 160M values) plus the ends; the
    # bound was measured over every float of the range
    lo = np.float32(-5.55).view(np.uint32)
    bits = np.arange(np.uint32(0x80000000), lo + 1, 7, dtype=np.uint32)
    for chunk in np.array_split(bits, 16):
        x = chunk.view(np.float32)
        y = orc.exp_blend_f32(x).astype(np.float64)
        ref = np.exp(x.astype(np.float64))
        ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
        assert (np.abs(y - ref) / ulp).max() < 2.7
    ends = orc.exp_blend_f32(np.array([0.0, -0.0, -5.55, -np.log(255.0)], np.float32))
    assert ends[0] == 1.0 and ends[1] == 1.0
    assert ends[2] < np.float32(1.0 / 255.0) <= ends[3] * np.float32(1.0 + 3e-7)
    # monotone on a dense grid (alpha thresholds depend on it)
    x = np.linspace(-5.55, 0.0, 2_000_001, dtype=np.float32)
    y = orc.exp_blend_f32(x)
    assert (np.diff(y) >= 0).all()


def test_blend_power_matches_calculate_power(orc):
    """power_f32 (dx (a dx + b dy) + c dy^2, a = -A/2, b = -B, c = -C/2) is
    calculatePower's -0.5 (A dx^2 + C dy^2) - B dx dy (raseter.wgsl:113-117):
    exact on small integers (a wrong sign of B, a swapped dx/dy or a lost
    half changes these), and within a few FP32 roundings of the FP64 formula
    on random positive-definite conics, near and far from the centre."""
    A = np.array([1, 2, 1, 1, 1, 3], np.float32)
    B = np.array([0, 0.5, 1, 0, -1, 0.25], np.float32)
    C = np.array([1, 1, 1, 4, 2, 0.5], np.float32)
    dx = np.array([3, 1, 1, 1, 2, -4], np.float32)
    dy = np.array([4, 2, 1, 0, -1, 2], np.float32)
    want = -0.5 * (A.astype(np.float64) * dx * dx + C.astype(np.float64) * dy * dy) - B.astype(np.float64) * dx * dy
    assert np.array_equal(orc.power_f32(A, B, C, dx, dy), want.astype(np.float32))
    assert orc.power_f32(1, 0, 4, 1, 0)[0] == -0.5 and orc.power_f32(1, 0, 4, 0, 1)[0] == -2.0
    rng = np.random.default_rng(11)
    n = 1_000_000
    A = rng.uniform(1e-3, 3.4, n).astype(np.float32)
    C = rng.uniform(1e-3, 3.4, n).astype(np.float32)
    B = (rng.uniform(-0.99, 0.99, n) * np.sqrt(A.astype(np.float64) * C)).astype(np.float32)
    scale = np.where(rng.random(n) < 0.5, 4.0, 300.0)
    dx = (rng.uniform(-1, 1, n) * scale).astype(np.float32)
    dy = (rng.uniform(-1, 1, n) * scale).astype(np.float32)
    p32 = orc.power_f32(A, B, C, dx, dy).astype(np.float64)
    A64, B64, C64, dx64, dy64 = (v.astype(np.float64) for v in (A, B, C, dx, dy))
    p64 = -0.5 * (A64 * dx64 * dx64 + C64 * dy64 * dy64) - B64 * dx64 * dy64
    mag = 0.5 * A64 * dx64 ** 2 + np.abs(B64 * dx64 * dy64) + 0.5 * C64 * dy64 ** 2
    assert (np.abs(p32 - p64) <= 6 * 2.0 ** -24 * mag + 1e-30).all()
    assert (p64 <= 0).all()


def test_blend_transmit_is_the_rounded_product(orc):
    """transmit_f32 (T - alpha T in one FMA) is T (1 - alpha) (raseter.wgsl:93)
    rounded once: bit-equal to the FP32 rounding of the exact product, which
    FP64 holds exactly for FP32 operands."""
    rng = np.random.default_rng(12)
    T = rng.uniform(1e-4, 1.0, 1_000_000).astype(np.float32)
    a = rng.uniform(1 / 255, 0.99, 1_000_000).astype(np.float32)
    exact = (T.astype(np.float64) * (1.0 - a.astype(np.float64))).astype(np.float32)
    assert np.array_equal(orc.transmit_f32(T, a), exact)
    assert orc.transmit_f32(0.5, 0.25)[0] == np.float32(0.375)


def _blend_alpha_f64(power, op):
    """raseter.wgsl:84-92 in FP64 with libm exp: skip (0) when power > 0 or alpha < 1/255."""
    p = np.asarray(power, np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        a = np.minimum(np.float64(np.float32(0.99)), np.asarray(op, np.float64) * np.exp(p))
    return np.where((p > 0) | (a < np.float64(np.float32(1.0 / 255.0))), 0.0, a)


def test_exp_blend_exponent_by_shift_equals_ldexp(orc):
    """The GPU kernels scale the blend exponential's 2^f by 2^k with one
    shift-and-add on the bits; the oracle uses ldexp.  Every float power in
    [-5.55, 0] gives the same bits both ways (exhaustive: 1.08e9 floats).
    Below about -87 the result underflows, where the bit form is wrong: the
    check sees that."""
    assert orc.exp_blend_shift_mismatches(0.0, 5.55) == 0
    assert orc.exp_blend_shift_mismatches(100.0, 104.0) > 0


def test_blend_decision_matches_fp64_at_every_power(orc):
    """The FP32 blend (power range cut, exp_blend_f32, 0.99 clamp, 1/255 test) takes
    the FP64 decision at every power from 0 down to -inf, including the powers
    below about -8.7e6 where exp_blend_f32's 2^k scale wraps: those must be
    skipped, never blended at the clamp."""
    far = np.array([-1e4, -8.8e6, -1e7, -1e30, -np.inf, -3.4e38], np.float32)
    sweep = -np.geomspace(1e-7, 3e38, 20001).astype(np.float32)
    near_cut = np.linspace(-6.5, -4.5, 20001, dtype=np.float32)
    powers = np.concatenate([far, sweep, near_cut, [0.0, -0.0, 1e-6, 2.0]]).astype(np.float32)
    ops = np.array([1e-3, 0.00392, 0.004, 0.05, 0.5, 0.9, 0.99, 0.999999, 1.0], np.float32)
    P, O = np.meshgrid(powers, ops, indexing="ij")
    a32 = orc.blend_alpha_f32(P, O).astype(np.float64)
    a64 = _blend_alpha_f64(P.astype(np.float64), O.astype(np.float64))
    assert np.isfinite(a32).all()
    # the far powers are skipped for every opacity
    assert not a32[: far.size].any()
    # decisions agree except where FP64 alpha is within 4 ulp of the 1/255 threshold
    thr = np.float64(np.float32(1.0 / 255.0))
    edge = np.abs(np.minimum(0.99, O.astype(np.float64) * np.exp(P.astype(np.float64))) - thr) <= 4 * np.spacing(np.float32(thr))
    differ = (a32 > 0) != (a64 > 0)
    assert not (differ & ~edge).any()
    assert edge.sum() < 50
    both = (a32 > 0) & (a64 > 0)
    ulp = np.spacing(a64[both].astype(np.float32)).astype(np.float64)
    assert (np.abs(a32[both] - a64[both]) / ulp).max() < 3.5
    # a NaN power (non-finite conic) passes both range tests and takes the clamp, as in the kernels
    assert orc.blend_alpha_f32(np.array([np.nan], np.float32), np.float32(0.5))[0] == np.float32(0.99)


@pytest.mark.parametrize("dist", [20.0, 300.0, 2500.0, 1e4, 1e6])
def test_raster_far_thin_splat_matches_fp64(orc, dist):
    """A thin splat (conic (3, 0, 1e-9), opacity 0.9) whose centre lies `dist`
    pixels to the side of a 16x16 tile, as a needle splat near the camera with
    an off-screen centre puts it: the FP32 raster blends exactly the pixels the
    FP64 raster blends (none once the power is below -5.55 everywhere)."""
    conic = [[3.0, 0.0, 1e-9, 0.9]]
    xy = [[-dist, 8.0]]
    rgb = [[0.3, 0.6, 0.9, 0.0]]
    img, T, last = _raster_one(orc, conic, xy, rgb)
    img64, T64, last64 = _raster_one(orc, conic, xy, rgb, fp64=True)
    assert np.array_equal(last, last64)
    assert np.abs(T - T64).max() < 1e-6 and np.abs(img - img64).max() < 1e-6
    if dist >= 20.0:
        assert not last.any() and (T == 1.0).all() and not img.any()


def test_raster_far_splat_mixed_powers(orc):
    """One 64x64 frame in which one splat (wide in x, very steep in y, centre just
    right of the frame) meets powers from about -0.2 at its own row down to
    about -9e6 at the frame's top and bottom rows, across the wrap of the
    exponential's scale: FP32 and FP64 rasters blend the same pixels."""
    W = H = 64
    conic = [[0.01, 0.0, 2e4, 0.95]]
    rgb = [[1.0, 1.0, 1.0, 0.0]]
    img, T, last = _raster_one(orc, conic, [[70.0, 30.0]], rgb, W=W, H=H)
    img64, T64, last64 = _raster_one(orc, conic, [[70.0, 30.0]], rgb, W=W, H=H, fp64=True)
    assert np.array_equal(last, last64) and last[30].any() and not last[0].any() and not last[63].any()
    assert np.abs(T - T64).max() < 1e-6
    for dist in (3e3, 4e4):
        img, T, last = _raster_one(orc, conic, [[-dist, 30.0]], rgb, W=W, H=H)
        assert not last.any() and (T == 1.0).all()


def _subset(sc, keep):
    return scenes.Scene(pos=sc.pos[keep], lscale=sc.lscale[keep], rot=sc.rot[keep], dc=sc.dc[keep],
                        opacity=sc.opacity[keep], rest=sc.rest[keep])


def test_edge_frame_far_needles_blend_nowhere(orc):
    """The edge frame's far needle splats cover every tile of the frame and meet
    every pixel at powers below -8.7e6 (the class the blend's exponential cannot
    evaluate): removing them leaves the FP32 and the FP64 frames bit for bit as
    they were, with the instance count lower by exactly their tiles."""
    from edge_scenes import edge_frame, powers_in_rect
    sc, cam = edge_frame()
    u = cam.uniform(sc.count)
    pre = orc.preprocess(sc.packed(), sc.count, u)
    far = [g for g in range(4096, sc.count)
           if pre["tiles"][g] > 0 and np.nanmax(powers_in_rect(pre, g, cam.width, cam.height)) < -8.7e6]
    assert len(far) == 15 and all(pre["tiles"][g] == 19 * 13 for g in far)
    keep = np.setdiff1d(np.arange(sc.count), far)
    sub = _subset(sc, keep)
    for fp64 in (False, True):
        a, Ta, _, Ia, _ = orc.render(sc.packed(), sc.count, u, fp64=fp64)
        b, Tb, _, Ib, _ = orc.render(sub.packed(), sub.count, cam.uniform(sub.count), fp64=fp64)
        assert Ia == Ib + 15 * 19 * 13
        assert np.array_equal(a, b) and np.array_equal(Ta, Tb), fp64


def test_indefinite_pancake_conic(orc):
    """The planted pancake of tests/edge_scenes.py: in FP32 its 2-D
    determinant rounds below zero, giving a visible splat with an indefinite
    conic (A < 0, C < 0) that blends some pixels; in FP64 the same splat's
    conic is positive definite."""
    from edge_scenes import edge_frame, indefinite_frame
    sc, cam = indefinite_frame()
    u = cam.uniform(sc.count)
    g = sc.count - 1
    c32 = orc.preprocess(sc.packed(), sc.count, u)["conic"][g]
    c64 = orc.preprocess(sc.packed(), sc.count, u, fp64=True)["conic"][g]
    assert c32[0] < 0 and c32[2] < 0 and c32[0] * c32[2] - c32[1] ** 2 < 0
    assert c64[0] > 0 and c64[2] > 0 and c64[0] * c64[2] - c64[1] ** 2 > 0
    a, _, _, _, _ = orc.render(sc.packed(), sc.count, u)
    base, bcam = edge_frame()
    b, _, _, _, _ = orc.render(base.packed(), base.count, bcam.uniform(base.count))
    assert (np.abs(a - b).max(0) > 0).sum() > 0


def test_edge_frame_fp32_vs_fp64(orc):
    """The whole edge frame (near-plane, behind-camera, off-screen, degenerate
    and huge splats, near needles): FP32 and FP64 frames take the same tile
    decisions (instance count) and the FP32 image stays within the error the
    FP32 arithmetic explains (a few cliff pixels where alpha or T lands on a
    threshold)."""
    from edge_scenes import edge_frame
    sc, cam = edge_frame()
    u = cam.uniform(sc.count)
    a, Ta, la, Ia, _ = orc.render(sc.packed(), sc.count, u)
    b, Tb, lb, Ib, _ = orc.render(sc.packed(), sc.count, u, fp64=True)
    assert Ia == Ib
    assert np.isfinite(a).all() and np.isfinite(Ta).all()
    d = np.abs(a - b)
    assert d.mean() < 3e-5 and d.max() < 0.05 and (la != lb).mean() < 0.02


@pytest.mark.parametrize("n", [0, 1, 7, 4096, 100003])
def test_scan_is_exclusive_prefix_sum(orc, n):
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 300, size=n + 1).astype(np.uint32)
    counts[n] = 0
    out, total = orc.scan(counts)
    ref = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]])
    assert np.array_equal(out.astype(np.uint64), ref)
    assert total == int(counts.astype(np.uint64).sum())


def test_emit_matches_brute_force(orc):
    rng = np.random.default_rng(5)
    n, W, tile = 300, 200, 16
    TX = (W + tile - 1) // tile
    x0 = rng.integers(0, TX, n); x1 = np.minimum(x0 + rng.integers(0, 4, n), TX - 1)
    y0 = rng.integers(0, 9, n); y1 = np.minimum(y0 + rng.integers(0, 4, n), 8)
    d16 = rng.integers(0, 65536, n)
    vis = rng.random(n) < 0.8
    meta = np.stack([(y0 << 16) | x0, (y1 << 16) | x1, d16, np.zeros(n, int)], 1).astype(np.uint32)
    counts = np.where(vis, (x1 - x0 + 1) * (y1 - y0 + 1), 0).astype(np.uint32)
    offs, total = orc.scan(np.concatenate([counts, [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, meta, n, W, tile, 1 << 30)
    ek, ev = [], []
    for g in range(n):
        if not vis[g]:
            continue
        for ty in range(y0[g], y1[g] + 1):
            for tx in range(x0[g], x1[g] + 1):
                ek.append(((ty * TX + tx) << 16) | d16[g]); ev.append(g)
    assert np.array_equal(keys, np.array(ek, np.uint64)) and np.array_equal(vals, np.array(ev, np.uint32))
    # capacity drops the tail slots (emit.wgsl:57-59)
    k2, v2 = orc.emit(offs, meta, n, W, tile, 100)
    assert k2.size == 100 and np.array_equal(k2, keys[:100])


@pytest.mark.parametrize("n,bits", [(0, 29), (1, 29), (1000, 8), (50000, 29), (200000, 45)])
def test_sort_is_stable_lexicographic(orc, n, bits):
    rng = np.random.default_rng(bits + n)
    keys = rng.integers(0, 1 << min(bits, 20), size=n).astype(np.uint64) << np.uint64(max(0, bits - 20))
    vals = np.arange(n, dtype=np.uint32)
    rng.shuffle(vals)
    k, v = orc.sort_stable(keys, vals)
    order = np.lexsort((np.arange(n), keys))      # stable by key, ties by input position
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    k2, v2 = orc.sort_radix(keys, vals, bits)
    assert np.array_equal(k2, k) and np.array_equal(v2, v)


def test_tile_ranges_closed_form(orc):
    rng = np.random.default_rng(9)
    tiles = np.sort(rng.integers(0, 50, 2000)).astype(np.uint64)
    keys = (tiles << np.uint64(16)) | rng.integers(0, 65536, 2000).astype(np.uint64)
    keys = np.sort(keys)
    r = orc.tile_ranges(keys, 64).reshape(64, 2)
    t = (keys >> np.uint64(16)).astype(np.int64)
    for tt in range(64):
        idx = np.nonzero(t == tt)[0]
        if idx.size:
            assert r[tt, 0] == idx[0] and r[tt, 1] == idx[-1] + 1
        else:
            assert r[tt, 0] == 0 and r[tt, 1] == 0
    assert orc.tile_ranges(keys[:0], 64).sum() == 0


def _cam(W=64, H=64, eye=(0.0, 0.0, -4.0), fov=60.0):
    return scenes.Camera(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov, W, H)


def test_preprocess_isotropic_on_axis_closed_form(orc):
    """Identity rotation, isotropic scale s, splat at the origin on the optical
    axis: Sigma = s^2 I, J diagonal (fx/z), conic = 1 / (s^2 f^2 / z^2 + 0.3)."""
    s = 0.05
    sc = _identity_splat_scene(lscale=math.log(s), opacity=0.0)
    cam = _cam(256, 256)
    u = cam.uniform(1)
    for fp64 in (False, True):
        o = orc.preprocess(sc.packed(), 1, u, fp64=fp64)
        f = float(u[36]); z = 4.0
        var = s * s * f * f / (z * z) + 0.3
        tol = 1e-4 if not fp64 else 1e-6
        assert abs(o["conic"][0, 0] - 1.0 / var) < tol * (1.0 / var)
        assert abs(o["conic"][0, 1]) < 1e-6
        assert abs(o["conic"][0, 2] - 1.0 / var) < tol * (1.0 / var)
        assert o["conic"][0, 3] == 0.5                       # sigmoid(0)
        assert abs(o["xy"][0, 0] - 127.5) < 1e-3 and abs(o["xy"][0, 1] - 127.5) < 1e-3
        r = math.ceil(3.0 * math.sqrt(var))
        x0 = int(math.floor((127.5 - r) / 16)); x1 = int(math.floor((127.5 + r) / 16))
        assert o["meta"][0, 0] == (x0 << 16) | x0 and o["meta"][0, 1] == (x1 << 16) | x1
        assert o["tiles"][0] == (x1 - x0 + 1) ** 2


def test_preprocess_sh_degree0_and_clamp(orc):
    """Zero higher bands: colour = C0 * dc + 0.5, clamped below at 0 (no upper clamp)."""
    dc = (1.0, -3.0, 5.0)
    sc = _identity_splat_scene(dc=dc)
    o = orc.preprocess(sc.packed(), 1, _cam().uniform(1))
    c0 = 0.28209479177387814
    exp = [max(c0 * d + 0.5, 0.0) for d in dc]
    assert np.allclose(o["rgb"][0, :3], exp, rtol=1e-6, atol=1e-7)
    assert o["rgb"][0, 1] == 0.0 and o["rgb"][0, 2] > 1.0


def test_preprocess_sh_band1_sign(orc):
    """Band-1 term -C1*y*f0 + C1*z*f1 - C1*x*f2 with d = normalize(p - eye): a splat on
    the +z axis seen from -z has d = (0,0,1), so only f1 contributes (+C1*f1)."""
    rest = np.zeros((1, 45), np.float32)
    rest[0, 0] = 1.0; rest[0, 1] = 2.0; rest[0, 2] = 4.0
    sc = _identity_splat_scene(pos=(0.0, 0.0, 0.5), rest=rest)
    o = orc.preprocess(sc.packed(), 1, _cam().uniform(1), fp64=True)
    assert abs(o["rgb"][0, 0] - (0.4886025119029199 * 2.0 + 0.5)) < 1e-6
    assert abs(o["rgb"][0, 1] - 0.5) < 1e-7   # other channels have zero coefficients


def test_preprocess_near_and_frustum_cull(orc):
    sc = _identity_splat_scene(n=3)
    sc.pos[0] = (0.0, 0.0, -3.9)     # 0.1 in front of the eye: near-culled
    sc.pos[1] = (50.0, 0.0, 0.0)     # far off to the side: frustum-culled
    sc.pos[2] = (0.0, 0.0, 0.0)
    o = orc.preprocess(sc.packed(), 3, _cam().uniform(3))
    assert o["tiles"][0] == 0 and not o["conic"][0].any() and not o["xy"][0].any()
    assert o["tiles"][1] == 0 and o["conic"][1].any() and not o["rgb"][1].any()
    assert o["tiles"][2] > 0


def test_preprocess_fp32_vs_fp64(orc):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    a = orc.preprocess(sc.packed(), sc.count, u)
    b = orc.preprocess(sc.packed(), sc.count, u, fp64=True)
    vis = a["tiles"] > 0
    assert np.abs(a["xy"] - b["xy"]).max() < 2e-3
    rel = np.abs(a["conic"] - b["conic"]) / (np.abs(b["conic"]) + 1e-3)
    assert rel.max() < 1e-3
    assert np.abs(a["rgb"][vis] - b["rgb"][vis]).max() < 1e-5
    # integer outputs differ only for splats straddling a float->int cliff
    assert (a["tiles"] != b["tiles"]).mean() < 1e-3
    assert (a["meta"] != b["meta"]).any(1).mean() < 1e-2


def _raster_one(orc, conic, xy, rgb, W=16, H=16, fp64=False):
    n = len(xy)
    vals = np.arange(n, dtype=np.uint32)
    tiles = ((W + 15) // 16) * ((H + 15) // 16)
    ranges = np.tile(np.array([0, n], np.uint32), tiles)   # every tile holds every splat
    return orc.raster(W, H, 16, ranges, vals, np.array(conic), np.array(xy), np.array(rgb), fp64=fp64)


def test_raster_single_splat_closed_form(orc):
    """One isotropic splat: C(x,y) = rgb * min(0.99, op * exp(-r^2 / (2 s^2)))
    where alpha >= 1/255, else 0."""
    s2 = 4.0
    op = 0.8
    img, T, last = _raster_one(orc, [[1 / s2, 0, 1 / s2, op]], [[7.3, 8.6]], [[0.2, 0.5, 1.0, 0]])
    yy, xx = np.mgrid[0:16, 0:16]
    a = op * np.exp(-((xx - 7.3) ** 2 + (yy - 8.6) ** 2) / (2 * s2))
    a = np.minimum(a, 0.99)
    a = np.where(a >= 1 / 255, a, 0.0)
    assert np.allclose(img[2], a * 1.0, atol=2e-6)
    assert np.allclose(img[0], a * 0.2, atol=2e-6)
    assert np.allclose(T, 1 - a, atol=2e-6)
    assert np.array_equal(last, (a > 0).astype(np.uint32))


def test_raster_opaque_front_stops(orc):
    """Opaque splats at one pixel: the first blends (alpha clamped to 0.99,
    T = 0.01); the second would take T to 0.01 * 0.01 < 1e-4, so the pixel stops
    there without blending it, and the third is never reached."""
    big = 1e-4
    img, T, last = _raster_one(orc, [[big, 0, big, 1.0], [big, 0, big, 1.0], [big, 0, big, 1.0]],
                               [[8, 8], [8, 8], [8, 8]], [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]])
    p = (8, 8)
    assert abs(img[0][p] - 0.99) < 1e-6 and img[1][p] == 0.0 and img[2][p] == 0.0
    assert last[p] == 1 and abs(T[p] - 0.01) < 1e-6


def test_render_fp32_vs_fp64_small(orc):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    a, Ta, la, Ia, _ = orc.render(sc.packed(), sc.count, u)
    b, Tb, lb, Ib, _ = orc.render(sc.packed(), sc.count, u, fp64=True)
    assert Ia == Ib == 84916
    d = np.abs(a - b)
    assert d.mean() < 1e-5 and (d > 1e-4).mean() < 1e-3
