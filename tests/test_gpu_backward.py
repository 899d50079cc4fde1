"""Hand-written backward kernels against the FP64 CPU adjoint.

The FP64 adjoint (oracle, decisions taken from the FP32 forward) is itself
pinned by FP64 central finite differences in tests/test_oracle_adjoint.py.
Tolerances.
* Against the FP64 adjoint: within 1e-3 relative norm-wise per parameter
  group, and element-wise |g - g64| <= 1e-3 |g64| + 1e-4 m for 99.9% of the
  elements, where m is the summed magnitude of every product the gradient's
  terms are built from (the FP64 adjoint reports it; for parameters it is
  |J|^T of the 2-D masses).  The last 0.1% are terms FP32 cannot resolve,
  e.g. a splat centred within 1e-3 px of a pixel centre, where x_splat -
  x_pixel loses most of its bits.
* Against the FP32 twin of the raster adjoint (oracle.raster_backward_f32,
  the kernel's arithmetic replayed on the CPU with double sums): every
  element within 1e-4 m.  Only the summation order (wave reductions, atomics)
  differs, so this pins the kernel itself.
The forward with saved state must be bit-identical to the plain forward."""
import sys

import numpy as np
import pytest

from conftest import to_dev, to_host
from gsr import lib, scenes

pytestmark = pytest.mark.gpu

GROUPS = {"position": slice(0, 3), "log-scale": slice(3, 6), "rotation": slice(6, 10), "dc": slice(10, 13),
          "opacity": slice(13, 14), "sh-rest": slice(14, 59)}
G2 = {"xy": slice(0, 2), "conic": slice(2, 5), "opacity": slice(5, 6), "rgb": slice(6, 9)}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gpu_grads(torch, ctx, sc, u, dLdC, tile=16):
    params = to_dev(torch, sc.planar())
    out = torch.empty(3 * int(u[40]) * int(u[41]), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(tile_size=tile, save_state=True), out)
    d = to_dev(torch, np.ascontiguousarray(dLdC, np.float32).ravel())
    grad = torch.empty_like(params)
    g2 = torch.empty(sc.count * 12, dtype=torch.float32, device="cuda")
    ctx.backward(params, sc.count, u, d, grad, g2)
    ctx.synchronize()
    return (to_host(grad, np.float32).reshape(59, sc.count).astype(np.float64),
            to_host(g2, np.float32).reshape(sc.count, 12).astype(np.float64),
            to_host(out, np.float32))


def _compare(name, got, ref, mass, axis_groups, axis):
    for gname, sl in axis_groups.items():
        a = got[sl] if axis == 0 else got[:, sl]
        b = ref[sl] if axis == 0 else ref[:, sl]
        m = mass[sl] if axis == 0 else mass[:, sl]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) == 0
            continue
        rel = np.linalg.norm(a - b) / nb
        bound = 1e-3 * np.abs(b) + 1e-4 * m
        ratio = np.abs(a - b) / np.maximum(bound, 1e-30)
        _say(f"{name}/{gname}: norm rel err {rel:.2e}; error/bound p99.9 {np.quantile(ratio, 0.999):.3f} "
             f"max {ratio.max():.3f}; share above the bound {(ratio > 1.0).mean():.2e}")
        assert rel < 1e-3, (gname, rel)
        assert np.quantile(ratio, 0.999) <= 1.0, (gname, np.quantile(ratio, 0.999))


def _compare_twin(torch, ctx, orc, sc, u, dLdC, g2, m2, tile=16):
    """GPU 2-D gradients vs the FP32 CPU twin on the same (bit-exact) forward."""
    n = sc.count
    W, H = int(u[40]), int(u[41])
    pre = orc.preprocess(sc.packed(), n, u, tile=tile)
    offs, total = orc.scan(np.concatenate([pre["tiles"], [0]]).astype(np.uint32))
    keys, vals = orc.emit(offs, pre["meta"], n, W, tile, 64 * n)
    ks, vs = orc.sort_radix(keys, vals, 40)
    TX, TY = (W + tile - 1) // tile, (H + tile - 1) // tile
    ranges = orc.tile_ranges(ks, TX * TY)
    _, T, last = orc.raster(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"])
    twin = orc.raster_backward_f32(W, H, tile, ranges, vs, pre["conic"], pre["xy"], pre["rgb"], T, last,
                                   np.asarray(dLdC, np.float32), n)
    err = np.abs(g2[:, :9] - twin[:, :9]) / np.maximum(1e-4 * m2[:, :9], 1e-30)
    _say(f"twin: error / (1e-4 mass) max {err.max():.3f}")
    assert err.max() <= 1.0, err.max()


@pytest.mark.parametrize("workload", ["small", "clustered"])
def test_backward_matches_fp64_adjoint(torch_cuda, ctx, orc, workload):
    if workload == "small":
        sc, cam = scenes.workload_small()
    else:
        sc = scenes.clustered_scene(60_000, seed=21)
        cam = scenes.Camera((0.0, 0.0, -8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, 640, 360)
    u = cam.uniform(sc.count)
    rng = np.random.default_rng(2)
    dLdC = rng.normal(size=(3, cam.height, cam.width))
    _say(f"backward {workload}: n={sc.count} {cam.width}x{cam.height}")
    grad, g2, _ = _gpu_grads(torch_cuda, ctx, sc, u, dLdC)
    ref, ref2, _, m2, pm = orc.backward(sc.packed(), sc.count, u, dLdC, decisions_f32=True, want_mass=True)
    _compare("grad2d", g2, ref2, m2, G2, axis=1)
    _compare("params", grad, ref, pm, GROUPS, axis=0)
    _compare_twin(torch_cuda, ctx, orc, sc, u, dLdC, g2, m2)


def test_forward_with_state_is_bit_identical(torch_cuda, ctx):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar())
    a = torch_cuda.empty(3 * 512 * 512, dtype=torch_cuda.float32, device="cuda")
    b = torch_cuda.empty_like(a)
    torch_cuda.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(save_state=False), a)
    ctx.forward(params, sc.count, u, lib.options(save_state=True), b)
    ctx.synchronize()
    assert torch_cuda.equal(a, b)


def test_autograd_function(torch_cuda, ctx):
    from gsr_autograd import render
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar()).requires_grad_(True)
    w = torch_cuda.randn(3, 512, 512, device="cuda", generator=torch_cuda.Generator("cuda").manual_seed(0))
    img = render(params, ctx, u)
    loss = (img * w).sum()
    loss.backward()
    g_direct, _, _ = _gpu_grads(torch_cuda, ctx, sc, u, w.cpu().numpy())
    got = params.grad.detach().cpu().numpy().astype(np.float64)
    assert np.allclose(got, g_direct, rtol=1e-4, atol=1e-6 * np.abs(g_direct).max())


def test_backward_requires_saved_state(torch_cuda, ctx):
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar())
    torch_cuda.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(save_state=False), None)
    grad = torch_cuda.empty_like(params)
    d = torch_cuda.zeros(3 * 512 * 512, device="cuda")
    with pytest.raises(lib.GsrError) as e:
        ctx.backward(params, sc.count, u, d, grad)
    assert e.value.code == lib.GSR_ERR_STATE


def test_backward_of_overflowed_forward_reports_capacity(torch_cuda, ctx):
    """gsr_forward is asynchronous: an instance overflow of the forward is
    reported by the next gsr_synchronize, a backward enqueued behind the failed
    forward writes NaN gradients instead of replaying a truncated frame, and
    the failed frame's state is dropped."""
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar())
    out = torch_cuda.empty(3 * 512 * 512, dtype=torch_cuda.float32, device="cuda")
    grad = torch_cuda.zeros_like(params)
    g2 = torch_cuda.zeros(sc.count * 12, dtype=torch_cuda.float32, device="cuda")
    d = torch_cuda.ones(3 * 512 * 512, dtype=torch_cuda.float32, device="cuda")
    torch_cuda.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(save_state=True, capacity=1000), out)
    ctx.backward(params, sc.count, u, d, grad, g2)
    with pytest.raises(lib.GsrError) as e:
        ctx.synchronize()
    assert e.value.code == lib.GSR_ERR_CAPACITY and "84916" in str(e.value)
    assert torch_cuda.isnan(grad).all().item() and torch_cuda.isnan(g2).all().item()
    with pytest.raises(lib.GsrError) as e:
        ctx.backward(params, sc.count, u, d, grad)
    assert e.value.code == lib.GSR_ERR_STATE
    ctx.synchronize()   # nothing left to report
    ctx.forward(params, sc.count, u, lib.options(save_state=True), out)
    ctx.backward(params, sc.count, u, d, grad)
    ctx.synchronize()
    assert torch_cuda.isfinite(grad).all().item()


def test_backward_rejects_another_camera_or_parameters(torch_cuda, ctx):
    """The raster adjoint replays the saved forward while the preprocess adjoint
    recomputes the projection from its arguments: a backward with a moved camera
    of the same size, or with other parameters, is refused."""
    sc, cam = scenes.workload_small()
    u = cam.uniform(sc.count)
    moved = scenes.Camera((0.1, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 512, 512).uniform(sc.count)
    params = to_dev(torch_cuda, sc.planar())
    other = params.clone()
    grad = torch_cuda.empty_like(params)
    d = torch_cuda.zeros(3 * 512 * 512, dtype=torch_cuda.float32, device="cuda")
    torch_cuda.cuda.synchronize()
    ctx.forward(params, sc.count, u, lib.options(save_state=True), None)
    for cam_u, p in ((moved, params), (u, other)):
        with pytest.raises(lib.GsrError) as e:
            ctx.backward(p, sc.count, cam_u, d, grad)
        assert e.value.code == lib.GSR_ERR_STATE
    ctx.backward(params, sc.count, u, d, grad)
    ctx.synchronize()


def test_backward_1m_1080p(torch_cuda, ctx, orc):
    sc, cam = scenes.workload_1m_1080p()
    u = cam.uniform(sc.count)
    rng = np.random.default_rng(3)
    dLdC = rng.normal(size=(3, 1080, 1920))
    _say("backward 1M 1080p")
    grad, g2, _ = _gpu_grads(torch_cuda, ctx, sc, u, dLdC)
    ref, ref2, _, m2, pm = orc.backward(sc.packed(), sc.count, u, dLdC, decisions_f32=True, want_mass=True)
    _compare("grad2d", g2, ref2, m2, G2, axis=1)
    _compare("params", grad, ref, pm, GROUPS, axis=0)
    _compare_twin(torch_cuda, ctx, orc, sc, u, dLdC, g2, m2)
