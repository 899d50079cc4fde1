"""Synthetic splat scenes and cameras for the renderer's workloads.

Scenes are produced in the loader's packed block layout
(src/loader/loader.js:75-85): pos[3N] | log-scale[3N] | rot[4N] (w,x,y,z) |
f_dc[3N] | opacity-logit[N] | f_rest[45N], and can be converted to the planar
[59][N] layout the device keeps (`packed_to_planar`).

Cameras follow the reference camera (src/renderer/camera.js:87-98, 135-142):
view = wgpu-matrix `mat4.lookAt(eye, target, up)`, projection =
`mat4.perspective(fovY, W/H, 0.2, 1000)` with clip z in [0, 1], focal =
(W / (2 tanFovX), H / (2 tanFovY)), tanFovX = tanFovY * W / H.  The 176-byte
uniform block mirrors GlobalUniforms (src/shaders/preprocess.wgsl:2-12).

Seeds: every workload uses numpy PCG64 seeded with its index.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

Z_NEAR = 0.2     # camera.js:13
Z_FAR = 1000.0   # camera.js:14
PARAMS_PER_SPLAT = 59

# Planar channel offsets ([59][N] layout).
CH_POS, CH_SCALE, CH_ROT, CH_DC, CH_OPACITY, CH_REST = 0, 3, 6, 10, 13, 14


@dataclass
class Scene:
    pos: np.ndarray      # (N,3) f32
    lscale: np.ndarray   # (N,3) f32
    rot: np.ndarray      # (N,4) f32, (w,x,y,z) = PLY rot_0..3
    dc: np.ndarray       # (N,3) f32
    opacity: np.ndarray  # (N,)  f32 logit
    rest: np.ndarray     # (N,45) f32, column c + 15*channel

    @property
    def count(self) -> int:
        return int(self.pos.shape[0])

    def packed(self) -> np.ndarray:
        return np.concatenate([self.pos.ravel(), self.lscale.ravel(), self.rot.ravel(),
                               self.dc.ravel(), self.opacity.ravel(), self.rest.ravel()]
                              ).astype(np.float32)

    def planar(self) -> np.ndarray:
        return np.ascontiguousarray(np.concatenate(
            [self.pos.T, self.lscale.T, self.rot.T, self.dc.T, self.opacity[None, :], self.rest.T],
            axis=0).astype(np.float32))

    @staticmethod
    def from_planar(p: np.ndarray) -> "Scene":
        p = np.asarray(p, dtype=np.float32)
        return Scene(pos=np.ascontiguousarray(p[0:3].T), lscale=np.ascontiguousarray(p[3:6].T),
                     rot=np.ascontiguousarray(p[6:10].T), dc=np.ascontiguousarray(p[10:13].T),
                     opacity=np.ascontiguousarray(p[13]), rest=np.ascontiguousarray(p[14:59].T))

    @staticmethod
    def from_packed(packed: np.ndarray, n: int) -> "Scene":
        a = np.asarray(packed, dtype=np.float32)
        o = 0

        def take(k):
            nonlocal o
            v = a[o:o + n * k].reshape(n, k) if k > 1 else a[o:o + n]
            o += n * k
            return np.ascontiguousarray(v)
        return Scene(pos=take(3), lscale=take(3), rot=take(4), dc=take(3), opacity=take(1), rest=take(45))


def packed_to_planar(packed: np.ndarray, n: int) -> np.ndarray:
    return Scene.from_packed(packed, n).planar()


def _common_attributes(rng: np.random.Generator, n: int, scale_mu: float, scale_sigma: float):
    lscale = rng.normal(scale_mu, scale_sigma, size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opacity = rng.normal(0.0, 1.0, size=n)
    dc = rng.normal(0.0, 0.5, size=(n, 3))
    rest = rng.normal(0.0, 0.1, size=(n, 45))
    return lscale, q, opacity, dc, rest


def uniform_scene(n: int, seed: int, extent: float = 1.0, scale_mu: float = -4.0,
                  scale_sigma: float = 0.5) -> Scene:
    """Means uniform in [-extent, extent]^3 (the 16,384-splat oracle scene uses extent 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-extent, extent, size=(n, 3))
    lscale, q, op, dc, rest = _common_attributes(rng, n, scale_mu, scale_sigma)
    return Scene(pos.astype(np.float32), lscale.astype(np.float32), q.astype(np.float32),
                 dc.astype(np.float32), op.astype(np.float32), rest.astype(np.float32))


def clustered_scene(n: int, seed: int, clusters: int = 64, half_extent: float = 3.0,
                    sigma: float = 0.3, scale_mu: float = -5.0, scale_sigma: float = 0.7) -> Scene:
    """Means from a mixture of isotropic Gaussian clusters (uneven tile loads)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-half_extent, half_extent, size=(clusters, 3))
    which = rng.integers(0, clusters, size=n)
    pos = centres[which] + rng.normal(0.0, sigma, size=(n, 3))
    lscale, q, op, dc, rest = _common_attributes(rng, n, scale_mu, scale_sigma)
    return Scene(pos.astype(np.float32), lscale.astype(np.float32), q.astype(np.float32),
                 dc.astype(np.float32), op.astype(np.float32), rest.astype(np.float32))


# ---- wgpu-matrix conventions -------------------------------------------------

def _normalize(v):
    return v / math.sqrt(float(np.dot(v, v)))


def look_at(eye, target, up) -> np.ndarray:
    """wgpu-matrix mat4.lookAt: column-major world->view, camera looks down -z."""
    eye = np.asarray(eye, dtype=np.float64)
    z = _normalize(eye - np.asarray(target, dtype=np.float64))
    x = _normalize(np.cross(np.asarray(up, dtype=np.float64), z))
    y = _normalize(np.cross(z, x))
    m = np.zeros(16)
    m[0], m[1], m[2] = x[0], y[0], z[0]
    m[4], m[5], m[6] = x[1], y[1], z[1]
    m[8], m[9], m[10] = x[2], y[2], z[2]
    m[12] = -float(np.dot(x, eye))
    m[13] = -float(np.dot(y, eye))
    m[14] = -float(np.dot(z, eye))
    m[15] = 1.0
    return m.astype(np.float32)


def perspective(fovy: float, aspect: float, near: float = Z_NEAR, far: float = Z_FAR) -> np.ndarray:
    """wgpu-matrix mat4.perspective: column-major, clip z in [0, 1]."""
    f = math.tan(math.pi * 0.5 - 0.5 * fovy)
    m = np.zeros(16)
    m[0] = f / aspect
    m[5] = f
    m[11] = -1.0
    rinv = 1.0 / (near - far)
    m[10] = far * rinv
    m[14] = far * near * rinv
    return m.astype(np.float32)


def mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major a*b with float32 storage (wgpu-matrix mat4.multiply)."""
    A = a.astype(np.float64).reshape(4, 4).T
    B = b.astype(np.float64).reshape(4, 4).T
    return (A @ B).T.ravel().astype(np.float32)


@dataclass
class Camera:
    eye: tuple
    target: tuple
    up: tuple
    fovy_deg: float
    width: int
    height: int

    def uniform(self, count: int = 0) -> np.ndarray:
        """The 176-byte GlobalUniforms block as 44 float32 slots."""
        aspect = self.width / self.height
        view = look_at(self.eye, self.target, self.up)
        proj = perspective(math.radians(self.fovy_deg), aspect)
        vp = mat4_mul(proj, view)
        tan_y = math.tan(math.radians(self.fovy_deg) / 2.0)
        tan_x = tan_y * aspect
        u = np.zeros(44, dtype=np.float32)
        u[0:16] = view
        u[16:32] = vp
        u[32:35] = np.asarray(self.eye, dtype=np.float32)
        u.view(np.uint32)[35] = count
        u[36:38] = (self.width / (2 * tan_x), self.height / (2 * tan_y))
        u[38:40] = (tan_x, tan_y)
        u[40:42] = (self.width, self.height)
        return u


def orbit_cameras(views: int, seed: int, radius: float = 6.0,
                  resolutions=((256, 256), (512, 384), (640, 480), (800, 600), (1024, 768)),
                  fov_range=(40.0, 70.0)) -> list:
    """Cameras on a sphere looking at the origin; up (0,1,0), or (0,0,1) when the
    view direction is within 0.99 of the y axis."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cams = []
    for _ in range(views):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = tuple(float(v) for v in d * radius)
        up = (0.0, 0.0, 1.0) if abs(d[1]) > 0.99 else (0.0, 1.0, 0.0)
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        fov = float(rng.uniform(*fov_range))
        cams.append(Camera(eye, (0.0, 0.0, 0.0), up, fov, int(w), int(h)))
    return cams


# ---- the workloads ----------------------------------------------------------

def workload_small():
    """16,384 splats, means U[-1,1]^3, 512x512, eye (0,0,-4), 60 deg (seed 0)."""
    return uniform_scene(16384, seed=0), Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 512, 512)


def workload_1m_1080p():
    """1,000,000 clustered splats, 1920x1080, eye (0,0,-8), 50 deg (seed 1)."""
    return clustered_scene(1_000_000, seed=1), Camera((0.0, 0.0, -8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, 1920, 1080)


def workload_batch64():
    """500,000 clustered splats, 64 orbit cameras of mixed resolution (seed 3)."""
    return clustered_scene(500_000, seed=3), orbit_cameras(64, seed=3)


def workload_6m_4k():
    """6,000,000 splats in 256 clusters over [-10,10]^3, 8 cameras at 3840x2160, 45 deg (seed 4)."""
    scene = clustered_scene(6_000_000, seed=4, clusters=256, half_extent=10.0, sigma=0.8,
                            scale_mu=-5.5, scale_sigma=0.8)
    cams = orbit_cameras(8, seed=4, radius=18.0, resolutions=((3840, 2160),), fov_range=(45.0, 45.0))
    return scene, cams
