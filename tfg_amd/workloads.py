"""Seeded ray batches for the traversal microbenchmark and parity tests.

Primary rays follow Camara::generarRayo (camara.h:42-59) with a uniform
sub-pixel jitter; bounce rays are cosine-sampled about the geometric normal
at primary hit points (render.h:379-413); shadow rays run from the bounce
origins to a point light with tmax = light distance (escenario.h:152-171).
All arrays use the tfg_ray layout: (n, 8) float32 = ox, oy, oz, tmin, dx,
dy, dz, tmax.
"""
from __future__ import annotations

import numpy as np

from .scenes import Scene

EPS = 1e-3   # ray offset, render.h:27 / escenario.h:40
MISS = 0xFFFFFFFF


def primary_rays(scene: Scene, width: int, height: int, seed: int, tmin: float = 0.0) -> np.ndarray:
    f = scene.camera.frame(width, height)
    pos, front, up, left = f[0:3], f[4:7], f[8:11], f[12:15]
    plane_w, plane_h = f[7], f[11]
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:height, 0:width]
    jx = rng.random(xs.size, dtype=np.float32)
    jy = rng.random(xs.size, dtype=np.float32)
    u = (xs.ravel().astype(np.float32) + jx) / np.float32(width)
    v = (ys.ravel().astype(np.float32) + jy) / np.float32(height)
    px = (np.float32(2) * u - np.float32(1)) * (plane_w * np.float32(0.5))
    py = (np.float32(1) - np.float32(2) * v) * (plane_h * np.float32(0.5))
    d = front[None, :] * f[3] + left[None, :] * px[:, None] + up[None, :] * py[:, None]
    d = d.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((xs.size, 8), np.float32)
    r[:, 0:3] = pos
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def _hit_frames(scene: Scene, rays: np.ndarray, hits: np.ndarray, rng):
    """Hit points and geometric normals facing the incoming ray; rays that
    missed borrow a random hit of the batch."""
    prim = hits["prim"]
    ok = np.nonzero(prim != MISS)[0]
    if ok.size == 0:
        raise ValueError("no primary ray hit the scene")
    idx = np.arange(rays.shape[0])
    miss = prim == MISS
    idx[miss] = ok[rng.integers(0, ok.size, int(miss.sum()))]
    r = rays[idx].astype(np.float64)
    t = hits["t"][idx].astype(np.float64)
    p = r[:, 0:3] + t[:, None] * r[:, 4:7]
    tv = scene.verts[prim[idx]].astype(np.float64).reshape(-1, 3, 3)
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    flip = np.einsum("ij,ij->i", n, r[:, 4:7]) > 0
    n[flip] = -n[flip]
    return p, n


def bounce_rays(scene: Scene, rays: np.ndarray, hits: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p, n = _hit_frames(scene, rays, hits, rng)
    r1 = rng.random(p.shape[0])
    r2 = rng.random(p.shape[0])
    phi = 2.0 * np.pi * r1
    st = np.sqrt(r2)
    loc = np.stack([st * np.cos(phi), st * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - r2))], 1)
    # tangent frame as render.h:396-413
    t = np.where(np.abs(n[:, 0:1]) > np.abs(n[:, 1:2]),
                 np.stack([n[:, 2], np.zeros_like(n[:, 0]), -n[:, 0]], 1),
                 np.stack([np.zeros_like(n[:, 0]), -n[:, 2], n[:, 1]], 1))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    d = loc[:, 0:1] * t + loc[:, 1:2] * b + loc[:, 2:3] * n
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = np.zeros((p.shape[0], 8), np.float32)
    out[:, 0:3] = p + n * EPS
    out[:, 3] = 0.0
    out[:, 4:7] = d
    out[:, 7] = np.inf
    return out


def shadow_rays(scene: Scene, rays: np.ndarray, hits: np.ndarray, seed: int, light: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p, n = _hit_frames(scene, rays, hits, rng)
    o = p + n * EPS
    lp = scene.lights[light, 0:3].astype(np.float64)
    d = lp[None, :] - o
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    out = np.zeros((p.shape[0], 8), np.float32)
    out[:, 0:3] = o
    out[:, 3] = 0.0
    out[:, 4:7] = d
    out[:, 7] = dist
    return out


def random_rays(n: int, seed: int, extent: float = 1.5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r
