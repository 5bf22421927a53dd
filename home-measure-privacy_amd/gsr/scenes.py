"""Seeded synthetic scenes and cameras for the rasterizer workloads.

Layouts follow trainer/gpu.js:
  * Gaussian record, 16 float32 (gpu.js:5-6, packGaussians gpu.js:688-694):
      0-2 mean, 3 pad, 4-6 logScale, 7 opacityLogit, 8-11 rotation wxyz,
      12-14 colour, 15 pad
  * camera uniform, 80 bytes (gpu.js:16, cameraUniform gpu.js:696-701):
      f32[0..11] world->camera rows [R|t], f32[12..15] fx fy cx cy,
      f32[16..17] width height, u32[18] Gaussian count, u32[19] pixels
  * SH coefficients (extension): float32 [n][16][3], coefficient-major; the DC
    coefficient is in colour units (colour = dc + sum_{l>=1} Y_lm f_lm), so a
    degree-0 scene is the reference's colour copy.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) run as a counter
stream: element i of stream s is mix(seed_s + (i+1) * golden_gamma).  Uniforms
take the top 53 bits; normals use Box-Muller on pairs of uniforms.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based SplitMix64 streams; every draw advances a 64-bit counter."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def bits(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * GAMMA
            self.state = self.state + np.uint64(n) * GAMMA
            return _mix(z)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self, n: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))  # 1 - u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return mean + std * z


@dataclass
class Camera:
    """Pinhole camera, COLMAP axes (x right, y down, z forward), colmap.js:45-54."""

    view: np.ndarray  # 12 floats, rows [R | t]
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int

    def uniform(self, count: int) -> bytes:
        """The 80-byte camera uniform of cameraUniform (gpu.js:696-701)."""
        f = np.zeros(20, dtype=np.float32)
        f[:12] = np.asarray(self.view, dtype=np.float32)
        f[12:18] = [self.fx, self.fy, self.cx, self.cy, self.width, self.height]
        u = f.view(np.uint32)
        u[18] = count
        u[19] = self.width * self.height
        return f.tobytes()

    @property
    def pixels(self) -> int:
        return self.width * self.height

    @property
    def tiles(self) -> int:
        return ((self.width + 15) // 16) * ((self.height + 15) // 16)

    def centre(self) -> np.ndarray:
        v = np.asarray(self.view, dtype=np.float64).reshape(3, 4)
        return -(v[:, :3].T @ v[:, 3])

    def crop(self, x0: int, y0: int, w: int, h: int) -> "Camera":
        """The sub-image [x0, x0+w) x [y0, y0+h) as a camera of its own."""
        return Camera(self.view.copy(), self.fx, self.fy, self.cx - x0, self.cy - y0, w, h)


def look_at(eye, at, width: int, height: int, fov_deg: float) -> Camera:
    """scene-test.html:16-24 lookAt with up = (0, 1, 0); horizontal field of view.

    When the view direction is within 0.999 of the up vector, up = (0, 0, 1)
    is used instead (the reference's rig never looks straight up or down).
    """
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(at, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    if abs(f @ up) > 0.999:
        up = np.array([0.0, 0.0, 1.0])
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    rows = np.stack([r, d, f])
    t = -(rows @ eye)
    view = np.concatenate([np.concatenate([rows[i], [t[i]]]) for i in range(3)])
    fx = (width / 2.0) / math.tan(math.radians(fov_deg) / 2.0)
    return Camera(view.astype(np.float64), fx, fx, width / 2.0, height / 2.0, width, height)


@dataclass
class Scene:
    gs: np.ndarray  # float32 [n, 16]
    sh: np.ndarray | None  # float32 [n, 16, 3] or None
    sh_degree: int
    cams: list = field(default_factory=list)
    targets: list = field(default_factory=list)  # float32 [H, W, 3] per camera

    @property
    def n(self) -> int:
        return self.gs.shape[0]


def gaussians(rng: Rng, n: int, sh_degree: int, means: np.ndarray | None = None):
    """means ~ U[-1,1]^3, logScale ~ N(-4, 0.5), q = normalise(N(0,1)^4),
    opacityLogit ~ N(0,1), colour ~ U[0,1]; SH: DC ~ U[0,1], bands N(0, 0.1)."""
    gs = np.zeros((n, 16), dtype=np.float32)
    gs[:, 0:3] = rng.uniform(3 * n, -1.0, 1.0).reshape(n, 3) if means is None else means
    gs[:, 4:7] = rng.normal(3 * n, -4.0, 0.5).reshape(n, 3)
    gs[:, 7] = rng.normal(n)
    q = rng.normal(4 * n).reshape(n, 4)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    gs[:, 8:12] = q
    col = rng.uniform(3 * n).reshape(n, 3)
    gs[:, 12:15] = col
    sh = None
    if sh_degree > 0:
        sh = np.zeros((n, 16, 3), dtype=np.float32)
        sh[:, 0, :] = col
        k = (sh_degree + 1) ** 2
        sh[:, 1:k, :] = rng.normal(n * (k - 1) * 3, 0.0, 0.1).reshape(n, k - 1, 3)
    return gs, sh


def unit_vector(rng: Rng) -> np.ndarray:
    v = rng.normal(3)
    return v / np.linalg.norm(v)


def fibonacci_sphere(k: int) -> np.ndarray:
    i = np.arange(k) + 0.5
    phi = np.arccos(1 - 2 * i / k)
    theta = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(theta) * np.sin(phi), np.cos(phi), np.sin(theta) * np.sin(phi)], axis=1)


RESOLUTIONS = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def config(idx: int, seed: int = 0, n: int | None = None, views: int | None = None, with_targets: bool = True) -> Scene:
    """The five rasterizer workloads.  ``n`` / ``views`` override the sizes (tests).

    0: 10k Gaussians, SH0, one 256x256 camera at radius 3, FOV 60, L1 target.
    1: 300k, SH3, 1280x720, FOV 60 (forward only).
    2: 1M, SH3, 1920x1080, FOV 60.
    3: 1M, SH3, 32 views, resolution uniform over RESOLUTIONS, Fibonacci sphere,
       radius U[2.5, 3.5], FOV U[45, 75].
    4: 3M, SH3, means from 64 Gaussian clusters (sigma 0.3) in a 4x4x2 box,
       64 views at 1600x1066.
    """
    rng = Rng(0x5EED0000 + 7919 * idx + seed)
    sizes = {0: 10_000, 1: 300_000, 2: 1_000_000, 3: 1_000_000, 4: 3_000_000}
    n = sizes[idx] if n is None else n
    deg = 0 if idx == 0 else 3
    means = None
    if idx == 4:
        centres = np.stack([rng.uniform(64, -2, 2), rng.uniform(64, -2, 2), rng.uniform(64, -1, 1)], axis=1)
        pick = (rng.uniform(n) * 64).astype(np.int64).clip(0, 63)
        means = centres[pick] + rng.normal(3 * n, 0.0, 0.3).reshape(n, 3)
        means = np.clip(means, [-2, -2, -1], [2, 2, 1]).astype(np.float32)
    gs, sh = gaussians(rng, n, deg, means)
    cams = []
    if idx in (0, 1, 2):
        W, H = {0: (256, 256), 1: (1280, 720), 2: (1920, 1080)}[idx]
        cams.append(look_at(3.0 * unit_vector(rng), [0, 0, 0], W, H, 60.0))
    elif idx == 3:
        k = 32 if views is None else views
        dirs = fibonacci_sphere(k)
        radius = rng.uniform(k, 2.5, 3.5)
        fov = rng.uniform(k, 45.0, 75.0)
        res = (rng.uniform(k) * 4).astype(np.int64).clip(0, 3)
        for v in range(k):
            W, H = RESOLUTIONS[res[v]]
            cams.append(look_at(radius[v] * dirs[v], [0, 0, 0], W, H, fov[v]))
    else:
        k = 64 if views is None else views
        dirs = fibonacci_sphere(k)
        for v in range(k):
            cams.append(look_at(4.0 * dirs[v], [0, 0, 0], 1600, 1066, 60.0))
    targets = []
    if with_targets:
        for c in cams:
            targets.append(rng.uniform(c.pixels * 3).astype(np.float32).reshape(c.height, c.width, 3))
    return Scene(gs, sh, deg, cams, targets)
