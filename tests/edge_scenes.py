"""Edge-case scenes shared by the CPU oracle pins and the GPU parity tests.

edge_scene: uniform splats plus planted cases for preprocess: the near plane,
behind the camera, off screen, huge and tiny scales, unnormalised and
degenerate quaternions, saturated opacities.

edge_frame: edge_scene plus what a real PLY scene puts in front of the raster
and its adjoint (raseter.wgsl:80-103, preprocess.wgsl:103-132):
* needle splats close to the camera whose centres lie 2,500 to 2,900 px off
  screen but whose 3-sigma squares cover the whole frame: every pixel meets
  them at powers of about -1e7, where the blend's exponential is out of its
  range and the alpha test must skip them;
* diagonal needles of the same kind that some pixels do blend;
* huge near-degenerate covariances: an edge-on pancake (scales e^3, e^3,
  e^-12) whose 2-D covariance is singular but for the 0.3 low-pass, and a
  huge isotropic splat far behind everything that covers the frame.
"""
from __future__ import annotations

import numpy as np

from gsr import scenes

FRAME_CAMERA = scenes.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 300, 200)


def edge_scene(n=4096, seed=11):
    sc = scenes.uniform_scene(n, seed=seed, extent=2.5, scale_mu=-3.0, scale_sigma=1.5)
    sc.pos[0] = (0.0, 0.0, -3.85)
    sc.pos[1] = (0.0, 0.0, -5.0)
    sc.pos[2] = (40.0, 0.0, 0.0)
    sc.lscale[3] = (1.0, 1.0, 1.0)
    sc.lscale[4] = (-12.0, -12.0, -12.0)
    sc.rot[5] = (3.0, 0.0, 0.0, 0.0)
    sc.rot[6] = (0.0, 0.0, 0.0, 1e-3)
    sc.opacity[7] = 30.0
    sc.opacity[8] = -30.0
    return sc


def _needles(cam):
    """Needles at depths 0.4/0.7/1.2 in front of the eye, centres far off screen
    (pixel offsets below), long axis across the offset (so the square reaches
    the frame) or diagonal; thin axes e^-9."""
    fx, fy = cam.uniform(1)[36:38]
    pos, ls, rot = [], [], []
    for d in (0.4, 0.7, 1.2):
        for opx, opy, axis in ((150, -2500, 0), (150, 2700, 0), (-2600, 100, 1), (2900, 100, 1),
                               (-1800, -1800, 2), (2100, 2300, 3)):
            x = (opx - 149.5) / fx * d
            y = (opy - 99.5) / fy * d
            sig_px = max(abs(opx - 150), abs(opy - 100)) / 2.0
            s = float(np.log(sig_px * d / fx))
            l = [-9.0, -9.0, -9.0]
            q = (1.0, 0.0, 0.0, 0.0)
            if axis < 2:
                l[axis] = s
            else:
                l[0] = s
                ang = np.pi / 4 if axis == 2 else -np.pi / 4
                q = (float(np.cos(ang / 2)), 0.0, 0.0, float(np.sin(ang / 2)))
            pos.append((x, y, -4.0 + d))
            ls.append(l)
            rot.append(q)
    # edge-on pancake: scales (e^3, e^3, e^-12) turned 90 degrees about x
    c = float(np.cos(np.pi / 4))
    pos.append((0.3, -0.2, -1.0)); ls.append([3.0, 3.0, -12.0]); rot.append((c, c, 0.0, 0.0))
    # huge isotropic splat behind the scene
    pos.append((0.0, 0.0, 6.0)); ls.append([4.0, 4.0, 4.0]); rot.append((1.0, 0.0, 0.0, 0.0))
    return np.array(pos, np.float32), np.array(ls, np.float32), np.array(rot, np.float32)


def edge_frame(n=4096, seed=11):
    """(scene, camera) of the edge frame: edge_scene(n, seed) then the planted
    needles and huge covariances (their opacity logit 2, colour from a DC of 0.5
    and seeded higher bands)."""
    base = edge_scene(n, seed)
    cam = scenes.Camera(*[getattr(FRAME_CAMERA, k) for k in ("eye", "target", "up", "fovy_deg", "width", "height")])
    pos, ls, rot = _needles(cam)
    k = len(pos)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    sc = scenes.Scene(pos=np.concatenate([base.pos, pos]), lscale=np.concatenate([base.lscale, ls]),
                      rot=np.concatenate([base.rot, rot]),
                      dc=np.concatenate([base.dc, np.full((k, 3), 0.5, np.float32)]),
                      opacity=np.concatenate([base.opacity, np.full(k, 2.0, np.float32)]),
                      rest=np.concatenate([base.rest, rng.normal(0.0, 0.1, (k, 45)).astype(np.float32)]))
    return sc, cam


# A flat pancake (scales e^4.36, e^4.51, e^-12.8) seen almost edge-on: its
# 2-D covariance is singular but for the 0.3 low-pass, and the FP32
# determinant a c - b^2 (entries near 1e7) rounds below zero, so the FP32
# conic has A < 0 and C < 0: an indefinite q that some pixels off its axes
# still meet in (0, q_max] (preprocess.wgsl:103-110 has no guard).
INDEFINITE_PANCAKE = dict(pos=(0.17569424, 0.18081267, -3.2771184), lscale=(4.3554664, 4.508484, -12.785446),
                          rot=(0.20866793, 0.21627764, 0.7852127, 0.54140806), opacity=0.7369, dc=0.5)


def indefinite_frame(n=4096, seed=11):
    """edge_frame(n, seed) plus INDEFINITE_PANCAKE as its last splat."""
    sc, cam = edge_frame(n, seed)
    p = INDEFINITE_PANCAKE
    sc = scenes.Scene(pos=np.concatenate([sc.pos, np.array([p["pos"]], np.float32)]),
                      lscale=np.concatenate([sc.lscale, np.array([p["lscale"]], np.float32)]),
                      rot=np.concatenate([sc.rot, np.array([p["rot"]], np.float32)]),
                      dc=np.concatenate([sc.dc, np.full((1, 3), p["dc"], np.float32)]),
                      opacity=np.concatenate([sc.opacity, np.array([p["opacity"]], np.float32)]),
                      rest=np.concatenate([sc.rest, np.zeros((1, 45), np.float32)]))
    return sc, cam


def powers_in_rect(pre, g, W, H, tile=16):
    """FP64 powers of splat g at every pixel of the tiles its rectangle covers
    (NaN elsewhere), from the oracle's FP32 preprocess outputs."""
    m = pre["meta"][g]
    x0, y0, x1, y1 = int(m[0] & 0xFFFF), int(m[0] >> 16), int(m[1] & 0xFFFF), int(m[1] >> 16)
    A, B, C, _ = pre["conic"][g].astype(np.float64)
    px, py = pre["xy"][g].astype(np.float64)
    yy, xx = np.mgrid[0:H, 0:W]
    inrect = (xx // tile >= x0) & (xx // tile <= x1) & (yy // tile >= y0) & (yy // tile <= y1)
    dx, dy = px - xx, py - yy
    p = -0.5 * (A * dx * dx + C * dy * dy) - B * dx * dy
    return np.where(inrect & (pre["tiles"][g] > 0), p, np.nan)


def power_conditioning(pre, W, H, tile=16):
    """Per splat: over the pixels of its rectangle where it can blend (power in
    [-5.55, 0] and alpha >= 1/255), the largest ratio of the summed magnitudes
    of the power's terms (A dx^2 / 2, C dy^2 / 2, B dx dy) to |power|.  FP32
    evaluates the power with an absolute error of about 6e-8 times that sum,
    so a ratio of 1e4 leaves alpha good to about 6e-4 only: the FP32 frame
    then differentiates a visibly different function than the FP64 one."""
    n = pre["tiles"].size
    cond = np.zeros(n)
    yy, xx = np.mgrid[0:H, 0:W]
    for g in np.nonzero(pre["tiles"])[0]:
        m = pre["meta"][g]
        x0, y0, x1, y1 = int(m[0] & 0xFFFF), int(m[0] >> 16), int(m[1] & 0xFFFF), int(m[1] >> 16)
        sl = (slice(y0 * tile, min(H, (y1 + 1) * tile)), slice(x0 * tile, min(W, (x1 + 1) * tile)))
        A, B, C, op = pre["conic"][g].astype(np.float64)
        px, py = pre["xy"][g].astype(np.float64)
        dx, dy = px - xx[sl], py - yy[sl]
        terms = 0.5 * abs(A) * dx * dx + 0.5 * abs(C) * dy * dy + np.abs(B * dx * dy)
        p = -0.5 * (A * dx * dx + C * dy * dy) - B * dx * dy
        with np.errstate(over="ignore"):
            cand = (p <= 0) & (p >= -5.55) & (op * np.exp(p) >= 1 / 255)
        if cand.any():
            cond[g] = (terms[cand] / np.maximum(np.abs(p[cand]), 1e-2)).max()
    return cond
