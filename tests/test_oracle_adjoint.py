"""Pins of the FP64 adjoint (oracle.backward) — CPU only.

1. Central finite differences (eps 1e-4, FP64 forward) on 256 randomly chosen
   splats, every one of the 59 parameters, on a 384-splat scene and on the
   1M-splat 1080p scene of the benchmark.  A component whose +eps and -eps
   frames differ in any discrete decision (tile membership, alpha cutoff,
   saturation stop, colour clamp) sits on a cliff of the renderer and is
   excluded; the excluded share is bounded.
2. The adjoint's forward image equals the FP64 forward.
3. Decisions taken from FP32 give the same gradient as FP64 decisions where
   the two frames agree (the branch choice is the only difference).
"""
import numpy as np

from gsr import scenes


def _scene():
    sc = scenes.uniform_scene(384, seed=7, extent=1.0, scale_mu=-2.5, scale_sigma=0.4)
    cam = scenes.Camera((0.3, 0.2, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, 96, 80)
    return sc, cam


def test_adjoint_image_is_fp64_forward(orc):
    sc, cam = _scene()
    u = cam.uniform(sc.count)
    dLdC = np.random.default_rng(0).normal(size=(3, 80, 96))
    _, _, img = orc.backward(sc.packed(), sc.count, u, dLdC, decisions_f32=False, want_image=True)
    ref, _, _, _, _ = orc.render(sc.packed(), sc.count, u, fp64=True)
    assert np.abs(img - ref).max() < 1e-12


def test_adjoint_matches_central_differences(orc):
    sc, cam = _scene()
    u = cam.uniform(sc.count)
    n = sc.count
    rng = np.random.default_rng(1)
    dLdC = rng.normal(size=(3, cam.height, cam.width))
    planar = sc.planar()
    grad, _, _ = orc.backward(sc.packed(), n, u, dLdC, decisions_f32=False)
    live = np.nonzero(np.abs(grad).sum(0) > 0)[0]
    chosen = rng.choice(live, size=min(256, live.size), replace=False)
    fd, same = orc.finite_differences(sc.packed(), n, u, dLdC, chosen, eps=1e-4)
    an = grad[:, chosen]
    errs = (np.abs(fd - an) / np.maximum(np.maximum(np.abs(an), np.abs(fd)), 1e-4))[same]
    assert same.mean() > 0.9, same.mean()
    assert np.quantile(errs, 0.999) < 1e-3 and errs.max() < 1e-2, (errs.max(), np.quantile(errs, 0.999))
    # every parameter group is exercised on the same branch
    for sl in (slice(0, 3), slice(3, 6), slice(6, 10), slice(10, 13), slice(13, 14), slice(14, 59)):
        assert same[sl].sum() > 0.8 * same[sl].size


def test_adjoint_matches_central_differences_on_the_1m_scene(orc):
    """The same check on the 1,000,000-splat 1920x1080 scene of the benchmark
    (dL/dRGB ~ N(0, 1)): central differences (eps 1e-4, FP64 frames) of 256
    randomly chosen visible splats, every one of their 59 parameters, against
    the FP64 adjoint of the whole frame.  Each perturbed frame is rendered only
    over the pixels where the splat can blend (oracle.finite_differences_window,
    checked against whole-frame differences below): every other pixel is the
    same in both frames.
    Position components cross discrete decisions most often (a shift of
    1.5e-2 px moves some pixel of the splat's footprint across the 1/255
    cutoff or a stop), so their branch share is bounded separately."""
    sc, cam = scenes.workload_1m_1080p()
    u = cam.uniform(sc.count)
    n = sc.count
    rng = np.random.default_rng(2)
    dLdC = rng.normal(size=(3, cam.height, cam.width))
    grad, _, _ = orc.backward(sc.packed(), n, u, dLdC, decisions_f32=False)
    pre = orc.preprocess(sc.packed(), n, u, fp64=True)
    chosen = np.sort(rng.choice(np.nonzero(pre["tiles"] > 0)[0], size=256, replace=False))
    fd, same = orc.finite_differences_window(sc.packed(), n, u, dLdC, chosen, eps=1e-4)
    an = grad[:, chosen]
    errs = (np.abs(fd - an) / np.maximum(np.maximum(np.abs(an), np.abs(fd)), 1e-4))[same]
    assert same.mean() > 0.95, same.mean()
    assert np.quantile(errs, 0.999) < 1e-3 and errs.max() < 1e-2, (errs.max(), np.quantile(errs, 0.999))
    assert same[0:3].mean() > 0.5
    for sl in (slice(3, 6), slice(6, 10), slice(10, 13), slice(13, 14), slice(14, 59)):
        assert same[sl].mean() > 0.95


def test_windowed_differences_equal_whole_frame_differences(orc):
    sc, cam = _scene()
    u = cam.uniform(sc.count)
    dLdC = np.random.default_rng(3).normal(size=(3, cam.height, cam.width))
    chosen = np.arange(0, sc.count, 12)
    a, sa = orc.finite_differences(sc.packed(), sc.count, u, dLdC, chosen)
    b, sb = orc.finite_differences_window(sc.packed(), sc.count, u, dLdC, chosen)
    assert np.array_equal(sa, sb)
    assert np.abs(a - b).max() <= 1e-8 * max(1.0, np.abs(a).max())


def test_fp32_decisions_agree_with_fp64_decisions(orc):
    sc, cam = _scene()
    u = cam.uniform(sc.count)
    dLdC = np.random.default_rng(4).normal(size=(3, 80, 96))
    a, a2, _ = orc.backward(sc.packed(), sc.count, u, dLdC, decisions_f32=True)
    b, b2, _ = orc.backward(sc.packed(), sc.count, u, dLdC, decisions_f32=False)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel < 1e-2
    same = np.isclose(a2, b2, rtol=1e-9, atol=1e-12).all(1)
    assert same.mean() > 0.9


def test_gradient_mass_bounds_the_gradient(orc):
    sc, cam = _scene()
    u = cam.uniform(sc.count)
    dLdC = np.random.default_rng(5).normal(size=(3, 80, 96))
    g, g2, _, m2, pm = orc.backward(sc.packed(), sc.count, u, dLdC, want_mass=True)
    assert (m2[:, :9] >= np.abs(g2[:, :9]) * (1 - 1e-12)).all()
    assert (pm >= np.abs(g) * (1 - 1e-9) - 1e-300).all()
    assert (m2[:, :9] > 0).any(1).sum() > sc.count // 2


def test_fp32_twin_matches_fp64_adjoint(orc):
    """The FP32 twin of the raster adjoint (the kernel's arithmetic on the CPU:
    T recovered through 1 / (1 - alpha), the dL/dRGB-weighted light from
    behind as one running sum) against the FP64 adjoint on the same FP32
    forward decisions: every 2-D gradient element within 1e-3 |g64| + 1e-4 m
    (m the summed magnitude of the element's terms)."""
    sc, cam = _scene()
    n = sc.count
    u = cam.uniform(n)
    W, H = cam.width, cam.height
    dLdC = np.random.default_rng(6).normal(size=(3, H, W))
    pre = orc.preprocess(sc.packed(), n, u, tile=16)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, 16, 64 * n)
    ks, vs = orc.sort_stable(keys, vals)
    ranges = orc.tile_ranges(ks, ((W + 15) // 16) * ((H + 15) // 16))
    _, T, last = orc.raster(W, H, 16, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    twin = orc.raster_backward_f32(W, H, 16, ranges, vs, pre["conic"], pre["xy"], pre["rgb"], T, last,
                                   np.asarray(dLdC, np.float32), n)
    _, ref2, _, m2, _ = orc.backward(sc.packed(), n, u, dLdC, decisions_f32=True, want_mass=True)
    a, b, m = twin[:, :9], ref2[:, :9], m2[:, :9]
    assert (np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-4 * m + 1e-30).all(), np.max(np.abs(a - b) / (1e-3 * np.abs(b) + 1e-4 * m + 1e-30))
    assert np.linalg.norm(a - b) < 1e-4 * np.linalg.norm(b)
