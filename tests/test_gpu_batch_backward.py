"""Batched backward: gradients of a multi-view batch frame summed over views.

The batch forward with saved state keeps every view's T_final and last
contributor; gsr_backward_batch replays every (view, tile) and sums the
parameter gradient over views.  Checks:
* the batch forward with save_state is bit-identical to the batch forward
  without it;
* against the FP64 CPU adjoint summed over the views (each view's adjoint is
  pinned by FP64 finite differences in tests/test_oracle_adjoint.py), with the
  tolerances of tests/test_gpu_backward.py and the views' masses summed;
* against the sum of per-view plain backwards on the GPU (the same per-view
  arithmetic; only the float summation order of the atomics and of the view
  sum differs): per parameter group, the norm of the difference relative to
  the norm of the summed per-view magnitudes sum_v |g_v| (the scale the view
  sum's rounding is relative to) within 1e-6, and every view's 2-D gradients
  norm-wise within 1e-5.  Log-scale and rotation get 5e-4: they come from the
  conic moments through the covariance chain, which amplifies the moments'
  float-atomic summation noise, and the batch sums dL/dSigma over the views
  before that chain where the per-view backwards each run it (the same sum,
  rounded differently).  Measured over runs of the 1M-splat 4-view case:
  6e-5 to 1.1e-4, against 1.4e-5 between two runs of one plain backward
  (printed).  The FP64 adjoint check above is the accuracy bar;
* the state checks (no saved state, other cameras, other count)."""
import sys

import numpy as np
import pytest

from conftest import to_dev, to_host
from gsr import lib, scenes

pytestmark = pytest.mark.gpu

GROUPS = {"position": slice(0, 3), "log-scale": slice(3, 6), "rotation": slice(6, 10), "dc": slice(10, 13),
          "opacity": slice(13, 14), "sh-rest": slice(14, 59)}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _layout(us):
    offs = [0]
    for u in us:
        offs.append(offs[-1] + 3 * int(u[40]) * int(u[41]))
    return offs


def _batch_grads(torch, ctx, params, n, us, dl_views, tile=16, want_2d=False):
    offs = ctx.batch_layout(n, us, lib.options(tile_size=tile))
    assert offs == _layout(us)
    out = torch.empty(offs[-1], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.render_batch(params, n, us, lib.options(tile_size=tile, save_state=True), out)
    d = to_dev(torch, np.concatenate([np.ascontiguousarray(x, np.float32).ravel() for x in dl_views]))
    grad = torch.empty(59 * n, dtype=torch.float32, device="cuda")
    g2 = torch.empty(len(us) * n * 12, dtype=torch.float32, device="cuda") if want_2d else None
    ctx.backward_batch(params, n, us, d, grad, g2)
    ctx.synchronize()
    g2h = to_host(g2, np.float32).reshape(len(us), n, 12).astype(np.float64) if want_2d else None
    return to_host(grad, np.float32).reshape(59, n).astype(np.float64), g2h, to_host(out, np.float32)


def _plain_grads(torch, ctx, params, n, u, dl, tile=16):
    W, H = int(u[40]), int(u[41])
    out = torch.empty(3 * W * H, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.forward(params, n, u, lib.options(tile_size=tile, save_state=True), out)
    grad = torch.empty(59 * n, dtype=torch.float32, device="cuda")
    g2 = torch.empty(n * 12, dtype=torch.float32, device="cuda")
    ctx.backward(params, n, u, to_dev(torch, np.ascontiguousarray(dl, np.float32).ravel()), grad, g2)
    ctx.synchronize()
    return (to_host(grad, np.float32).reshape(59, n).astype(np.float64),
            to_host(g2, np.float32).reshape(n, 12).astype(np.float64))


def _norm_rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def _check_view_sum(grad, want, mag):
    for name, sl in GROUPS.items():
        rel = np.linalg.norm(grad[sl] - want[sl]) / max(np.linalg.norm(mag[sl]), 1e-30)
        _say(f"{name}: |diff| / |sum_v |g_v|| {rel:.2e}; norm rel {_norm_rel(grad[sl], want[sl]):.2e}")
        assert rel < (5e-4 if name in ("log-scale", "rotation") else 1e-6), (name, rel)


def _check_fp64(grad, ref, mass):
    for name, sl in GROUPS.items():
        a, b, m = grad[sl], ref[sl], mass[sl]
        rel = _norm_rel(a, b)
        ratio = np.abs(a - b) / np.maximum(1e-3 * np.abs(b) + 1e-4 * m, 1e-30)
        _say(f"params/{name}: norm rel {rel:.2e}; ratio p99.9 {np.quantile(ratio, 0.999):.3f}")
        assert rel < 1e-3, (name, rel)
        assert np.quantile(ratio, 0.999) <= 1.0, name


def _views(n, views, seed, res=((256, 256), (512, 384), (640, 480))):
    cams = scenes.orbit_cameras(views, seed=seed, resolutions=res)
    return [c.uniform(n) for c in cams]


def test_batch_forward_with_state_is_bit_identical(torch_cuda, ctx):
    sc = scenes.clustered_scene(30_000, seed=8)
    us = _views(sc.count, 5, seed=8)
    params = to_dev(torch_cuda, sc.planar())
    offs = ctx.batch_layout(sc.count, us)
    a = torch_cuda.empty(offs[-1], dtype=torch_cuda.float32, device="cuda")
    b = torch_cuda.empty_like(a)
    torch_cuda.cuda.synchronize()
    ctx.render_batch(params, sc.count, us, lib.options(), a)
    ctx.render_batch(params, sc.count, us, lib.options(save_state=True), b)
    ctx.synchronize()
    assert torch_cuda.equal(a, b)


def test_batch_backward_matches_fp64_adjoint_sum(torch_cuda, ctx, orc):
    sc = scenes.clustered_scene(20_000, seed=9)
    us = _views(sc.count, 3, seed=9)
    rng = np.random.default_rng(4)
    dls = [rng.normal(size=(3, int(u[41]), int(u[40]))) for u in us]
    params = to_dev(torch_cuda, sc.planar())
    _say(f"batch backward vs FP64: n={sc.count} views={len(us)}")
    grad, g2, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us, dls, want_2d=True)
    ref = np.zeros((59, sc.count))
    mass = np.zeros((59, sc.count))
    for v, (u, dl) in enumerate(zip(us, dls)):
        r, r2, _, m2, pm = orc.backward(sc.packed(), sc.count, u, dl, decisions_f32=True, want_mass=True)
        ref += r
        mass += pm
        # each view's 2-D gradient against that view's FP64 adjoint
        a, b = g2[v][:, :9], r2[:, :9]
        ratio = np.abs(a - b) / np.maximum(1e-3 * np.abs(b) + 1e-4 * m2[:, :9], 1e-30)
        _say(f"view {v} grad2d: norm rel {_norm_rel(a, b):.2e}; ratio p99.9 {np.quantile(ratio, 0.999):.3f}")
        assert _norm_rel(a, b) < 1e-3
        assert np.quantile(ratio, 0.999) <= 1.0
    _check_fp64(grad, ref, mass)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batch_backward_equals_sum_of_view_backwards(torch_cuda, ctx, tile):
    sc = scenes.clustered_scene(40_000, seed=10)
    us = _views(sc.count, 4, seed=10)
    rng = np.random.default_rng(5)
    dls = [rng.normal(size=(3, int(u[41]), int(u[40]))) for u in us]
    params = to_dev(torch_cuda, sc.planar())
    _say(f"batch backward vs per-view: tile {tile}")
    grad, g2, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us, dls, tile=tile, want_2d=True)
    want = np.zeros_like(grad)
    mag = np.zeros_like(grad)
    for v, (u, dl) in enumerate(zip(us, dls)):
        gv, g2v = _plain_grads(torch_cuda, ctx, params, sc.count, u, dl, tile=tile)
        want += gv
        mag += np.abs(gv)
        assert _norm_rel(g2[v][:, :9], g2v[:, :9]) < 1e-5, v
        assert not g2[v][:, 9:].any()
    _check_view_sum(grad, want, mag)


def test_repeated_batch_backward_gives_the_same_gradient(torch_cuda, ctx):
    """The batched adjoint clears the 2-D gradient sums it reads, and the
    next backward of the same shape skips its fill: two backward calls on one
    saved forward, a third after a forward with other cameras and a fourth
    back on the first, must each equal a fresh run's gradients (up to the
    float atomics' order)."""
    sc = scenes.clustered_scene(30_000, seed=13)
    us = _views(sc.count, 3, seed=13)
    us2 = _views(sc.count, 3, seed=14)
    rng = np.random.default_rng(6)
    dls = [rng.normal(size=(3, int(u[41]), int(u[40]))) for u in us]
    params = to_dev(torch_cuda, sc.planar())
    _say(f"repeated batch backward: n={sc.count} views=3")
    first, first2d, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us, dls, want_2d=True)
    d = to_dev(torch_cuda, np.concatenate([np.ascontiguousarray(x, np.float32).ravel() for x in dls]))
    grad = torch_cuda.empty(59 * sc.count, dtype=torch_cuda.float32, device="cuda")
    ctx.backward_batch(params, sc.count, us, d, grad)   # same saved forward again
    ctx.synchronize()
    again = to_host(grad, np.float32).reshape(59, sc.count).astype(np.float64)
    assert _norm_rel(again, first) < 1e-5
    other, _, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us2, dls)
    assert _norm_rel(other, first) > 1e-2   # other cameras: another gradient
    back, back2d, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us, dls, want_2d=True)
    assert _norm_rel(back, first) < 1e-5
    assert _norm_rel(back2d, first2d) < 1e-5


def test_batch_backward_state_checks(torch_cuda, ctx):
    sc = scenes.clustered_scene(10_000, seed=11)
    us = _views(sc.count, 2, seed=11)
    params = to_dev(torch_cuda, sc.planar())
    offs = ctx.batch_layout(sc.count, us)
    out = torch_cuda.empty(offs[-1], dtype=torch_cuda.float32, device="cuda")
    d = torch_cuda.zeros(offs[-1], dtype=torch_cuda.float32, device="cuda")
    grad = torch_cuda.empty(59 * sc.count, dtype=torch_cuda.float32, device="cuda")
    torch_cuda.cuda.synchronize()
    ctx.render_batch(params, sc.count, us, lib.options(), out)   # no saved state
    with pytest.raises(lib.GsrError) as e:
        ctx.backward_batch(params, sc.count, us, d, grad)
    assert e.value.code == lib.GSR_ERR_STATE
    ctx.render_batch(params, sc.count, us, lib.options(save_state=True), out)
    other = _views(sc.count, 2, seed=12)
    with pytest.raises(lib.GsrError) as e:
        ctx.backward_batch(params, sc.count, other, d, grad)      # other cameras
    assert e.value.code == lib.GSR_ERR_STATE
    with pytest.raises(lib.GsrError) as e:
        ctx.backward_batch(params, sc.count - 1, us, d, grad)     # other count
    assert e.value.code == lib.GSR_ERR_STATE
    ctx.backward_batch(params, sc.count, us, d, grad)           # zero dL: zero gradient
    ctx.synchronize()
    assert not to_host(grad, np.float32).any()


def test_batch_backward_1m_4_views_1080p(torch_cuda, ctx, orc):
    """1M splats, 4 views at 1920x1080: against the FP64 adjoint summed over
    the views and against the sum of the views' plain backwards."""
    sc = scenes.clustered_scene(1_000_000, seed=1)
    cams = scenes.orbit_cameras(4, seed=14, radius=8.0, resolutions=((1920, 1080),), fov_range=(50.0, 50.0))
    us = [c.uniform(sc.count) for c in cams]
    rng = np.random.default_rng(6)
    dls = [rng.normal(size=(3, 1080, 1920)).astype(np.float32) for _ in us]
    params = to_dev(torch_cuda, sc.planar())
    _say("batch backward 1M x 4 views 1080p")
    grad, _, _ = _batch_grads(torch_cuda, ctx, params, sc.count, us, dls)
    want = np.zeros_like(grad)
    mag = np.zeros_like(grad)
    for u, dl in zip(us, dls):
        gv = _plain_grads(torch_cuda, ctx, params, sc.count, u, dl)[0]
        want += gv
        mag += np.abs(gv)
    # run-to-run spread of one plain backward (float-atomic order), for scale
    again = _plain_grads(torch_cuda, ctx, params, sc.count, us[-1], dls[-1])[0]
    for name, sl in GROUPS.items():
        rr = np.linalg.norm(again[sl] - gv[sl]) / max(np.linalg.norm(np.abs(gv[sl])), 1e-30)
        _say(f"plain rerun {name}: |diff| / |g| {rr:.2e}")
    assert np.isfinite(grad).all()
    _say("FP64 adjoint of each view")
    ref = np.zeros_like(grad)
    mass = np.zeros_like(grad)
    for u, dl in zip(us, dls):
        r, _, _, _, pm = orc.backward(sc.packed(), sc.count, u, dl, decisions_f32=True, want_mass=True)
        ref += r
        mass += pm
    _check_fp64(grad, ref, mass)
    _check_view_sum(grad, want, mag)
