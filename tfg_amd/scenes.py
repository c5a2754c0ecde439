"""Seeded procedural scenes for the benchmark and parity workloads.

The reference renders meshes loaded through assimp (cargarModelo.h:83-224)
next to analytic planes and spheres (main.cpp:164-345).  The workloads here
are generated instead, so every run is reproducible from its seed:

* ``cornell``        - room of 12 axis-aligned quads (24 triangles; five walls,
                       a closing wall behind the camera and a 6-quad box), a
                       subdivided icosphere and one point light.
* ``displaced_sphere`` - UV sphere displaced along its normals by seeded
                       Perlin noise.
* ``instances``      - seeded displaced spheres and boxes flattened into one
                       triangle list inside a closed room (or on a ground
                       quad under a black sky), several point lights.

Every scene is returned as a :class:`Scene` of plain numpy arrays: triangle
vertices (N, 9), per-corner shading normals (N, 9; all-zero rows mean "use
the geometric normal"), a material index per triangle, the material table,
point lights and a camera.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Material table columns (tfg_material, include/tfg_types.h):
#   kd[3], roughness, ks[3], ior, kt[3], pad, le[3], pad  -> 16 floats
MAT_FLOATS = 16


def material(kd=(0.0, 0.0, 0.0), ks=(0.0, 0.0, 0.0), roughness=0.0, kt=(0.0, 0.0, 0.0), ior=1.0, le=(0.0, 0.0, 0.0)):
    m = np.zeros(MAT_FLOATS, np.float32)
    m[0:3] = kd
    m[3] = roughness
    m[4:7] = ks
    m[7] = ior
    m[8:11] = kt
    m[12:15] = le
    return m


@dataclasses.dataclass
class Camera:
    """Pinhole camera parameterised like Camara (camara.h:15-59)."""

    pos: tuple
    front: tuple
    up: tuple
    fov_deg: float = 45.0

    def frame(self, width: int, height: int) -> np.ndarray:
        """16 floats: pos, focal, front, plane_w, up, plane_h, left, pad (tfg_camera).

        Same construction as camara.h:29-38 (left = front x up, up = left x
        front) and RenderKernel.cu:70-74 (focal 1, plane height 2 tan(fov/2)).
        """
        f = np.asarray(self.front, np.float64)
        f /= np.linalg.norm(f)
        u = np.asarray(self.up, np.float64)
        u /= np.linalg.norm(u)
        left = np.cross(f, u)
        left /= np.linalg.norm(left)
        up = np.cross(left, f)
        up /= np.linalg.norm(up)
        focal = np.float32(1.0)
        fov = np.float32(self.fov_deg * math.pi / 180.0)
        plane_h = np.float32(2.0 * focal * np.tan(np.float32(fov / np.float32(2.0))))
        plane_w = np.float32(plane_h * (np.float32(width) / np.float32(height)))
        out = np.zeros(16, np.float32)
        out[0:3] = self.pos
        out[3] = focal
        out[4:7] = f
        out[7] = plane_w
        out[8:11] = up
        out[11] = plane_h
        out[12:15] = left
        return out


@dataclasses.dataclass
class Scene:
    name: str
    verts: np.ndarray        # (N, 9) float32
    normals: np.ndarray      # (N, 9) float32
    mat_id: np.ndarray       # (N,) uint32
    materials: np.ndarray    # (M, 16) float32
    lights: np.ndarray       # (L, 8) float32: pos, pad, intensity, pad
    camera: Camera
    seed: int = 0
    analytic: np.ndarray | None = None   # (K, 12) float32 tfg_analytic records (type/material as uint32 bits)

    @property
    def n_tris(self) -> int:
        return int(self.verts.shape[0])

    def shading_records(self) -> np.ndarray:
        """(N, 16) uint32 view of tfg_tri_shading records (64 B each)."""
        rec = np.zeros((self.n_tris, 16), np.float32)
        rec[:, 0:3] = self.normals[:, 0:3]
        rec[:, 4:7] = self.normals[:, 3:6]
        rec[:, 8:11] = self.normals[:, 6:9]
        v = self.verts.reshape(-1, 3, 3)
        rec[:, 12:15] = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        u = rec.view(np.uint32)
        u[:, 3] = self.mat_id
        smooth = np.any(self.normals[:, 0:3] != self.normals[:, 3:6], axis=1) | np.any(
            self.normals[:, 0:3] != self.normals[:, 6:9], axis=1)
        u[:, 7] = smooth.astype(np.uint32)
        return u


# --------------------------------------------------------------- geometry

def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    p0, p1, p2, p3 = (np.asarray(p, np.float32) for p in (p0, p1, p2, p3))
    return np.stack([np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])]).astype(np.float32)


def box(lo, hi):
    """12 triangles of an axis-aligned box, outward winding."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = [
        (x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
        (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1),
    ]
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2)]
    return np.concatenate([quad(*(c[i] for i in f)) for f in faces])


def icosphere(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron: 20 * 4**level triangles; returns (verts, normals)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    f = np.array(
        [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mids])
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    n = v[f]                                   # (F, 3, 3) unit normals
    p = n * radius + np.asarray(center)
    return p.reshape(-1, 9).astype(np.float32), n.reshape(-1, 9).astype(np.float32)


def _perlin_tables(seed: int):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256).astype(np.int64)
    return np.concatenate([perm, perm])


_GRAD = np.array([(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
                  (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (0, -1, 1), (-1, 1, 0), (0, -1, -1)],
                 np.float64)


def perlin3(p: np.ndarray, seed: int) -> np.ndarray:
    """Improved Perlin noise (Perlin, SIGGRAPH 2002) at points p (K, 3), seeded permutation."""
    perm = _perlin_tables(seed)
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    pi &= 255
    fade = pf * pf * pf * (pf * (pf * 6 - 15) + 10)
    out = np.zeros(p.shape[0])

    def grad(h, x, y, z):
        g = _GRAD[h & 15]
        return g[:, 0] * x + g[:, 1] * y + g[:, 2] * z

    X, Y, Z = pi[:, 0], pi[:, 1], pi[:, 2]
    x, y, z = pf[:, 0], pf[:, 1], pf[:, 2]
    u, v, w = fade[:, 0], fade[:, 1], fade[:, 2]
    A = perm[X] + Y
    AA = perm[A] + Z
    AB = perm[A + 1] + Z
    B = perm[X + 1] + Y
    BA = perm[B] + Z
    BB = perm[B + 1] + Z
    lerp = lambda t, a, b: a + t * (b - a)  # noqa: E731
    out = lerp(w,
               lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                    lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
               lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                    lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))))
    return out


def displaced_uv_sphere(nu: int, nv: int, radius: float, amplitude: float, freq: float, seed: int,
                        center=(0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0)):
    """UV sphere with nu segments and nv rings (2*nu*nv triangles), displaced
    along the normal by amplitude * perlin(freq * n).  Returns (verts, normals)."""
    th = np.linspace(0.0, math.pi, nv + 1)
    ph = np.linspace(0.0, 2.0 * math.pi, nu + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")             # (nv+1, nu+1)
    n = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    n[0, :] = (0.0, 1.0, 0.0)
    n[-1, :] = (0.0, -1.0, 0.0)
    n[:, -1] = n[:, 0]
    flat = n.reshape(-1, 3)
    d = radius + amplitude * perlin3(flat * freq + 17.0, seed)
    pos = (flat * d[:, None]) * np.asarray(scale) + np.asarray(center)
    pos = pos.reshape(nv + 1, nu + 1, 3)
    # vertex normals from the displaced surface: accumulate face normals
    idx = np.arange((nv + 1) * (nu + 1)).reshape(nv + 1, nu + 1)
    a, b = idx[:-1, :-1], idx[:-1, 1:]
    c, e = idx[1:, :-1], idx[1:, 1:]
    tris = []
    tris.append(np.stack([a[1:], c[1:], e[1:]], -1).reshape(-1, 3))     # skip degenerate top row
    tris.append(np.stack([a[:-1], e[:-1], b[:-1]], -1).reshape(-1, 3))  # skip degenerate bottom row
    tris.append(np.stack([a[0], c[0], e[0]], -1).reshape(-1, 3))        # top cap
    tris.append(np.stack([a[-1], e[-1], b[-1]], -1).reshape(-1, 3))     # bottom cap
    F = np.concatenate(tris)
    P3 = pos.reshape(-1, 3)
    # outward orientation: flip winding where the face normal points inward,
    # then accumulate the (now outward) face normals into vertex normals
    ctr = np.asarray(center)
    fn = np.cross(P3[F[:, 1]] - P3[F[:, 0]], P3[F[:, 2]] - P3[F[:, 0]])
    inward = np.einsum("ij,ij->i", fn, P3[F].mean(1) - ctr) < 0
    F[inward] = F[inward][:, [0, 2, 1]]
    fn[inward] = -fn[inward]
    vn = np.zeros_like(P3)
    for k in range(3):
        np.add.at(vn, F[:, k], fn)
    # seam: merge normals of duplicated vertices on the first/last column
    vn = vn.reshape(nv + 1, nu + 1, 3)
    seam = vn[:, 0] + vn[:, -1]
    vn[:, 0] = seam
    vn[:, -1] = seam
    vn[0, :] = vn[0, :].sum(0)
    vn[-1, :] = vn[-1, :].sum(0)
    vn = vn.reshape(-1, 3)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    verts = P3[F].reshape(-1, 9).astype(np.float32)
    norms = vn[F].reshape(-1, 9).astype(np.float32)
    return verts, norms


def subdivided_box(lo, hi, n: int):
    """Axis-aligned box with each face split into n x n quads (12 n^2 triangles)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    out = []
    g = np.linspace(0.0, 1.0, n + 1)
    U, V = np.meshgrid(g, g, indexing="ij")
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        for side in (0, 1):
            P = np.zeros((n + 1, n + 1, 3))
            P[..., axis] = hi[axis] if side else lo[axis]
            P[..., a1] = lo[a1] + U * (hi[a1] - lo[a1])
            P[..., a2] = lo[a2] + V * (hi[a2] - lo[a2])
            p00, p10, p01, p11 = P[:-1, :-1], P[1:, :-1], P[:-1, 1:], P[1:, 1:]
            t1 = np.stack([p00, p10, p11], 2).reshape(-1, 9)
            t2 = np.stack([p00, p11, p01], 2).reshape(-1, 9)
            tri = np.concatenate([t1, t2])
            if side == 0:   # keep outward winding
                tri = tri.reshape(-1, 3, 3)[:, [0, 2, 1]].reshape(-1, 9)
            out.append(tri)
    return np.concatenate(out).astype(np.float32)


# ----------------------------------------------------------------- scenes

def cornell(level: int = 4, seed: int = 0) -> Scene:
    """12-quad room (walls x,y in [-3,3], z in [-8, 4.5]) + L-`level` icosphere (r = 0.3).

    Light and camera follow the reference's base scene (main.cpp:174-177,
    RenderKernel.cu:81-83): light (0, 2.9, -4) with intensity 100,
    camera (0, -0.75, 4.25) looking down -z.
    """
    X0, X1, Y0, Y1, Z0, Z1 = -3.0, 3.0, -3.0, 3.0, -8.0, 4.5
    white = material(kd=(0.75, 0.75, 0.75))
    red = material(kd=(0.75, 0.15, 0.15))
    green = material(kd=(0.15, 0.75, 0.15))
    quads, mats = [], []
    quads.append(quad((X0, Y0, Z1), (X1, Y0, Z1), (X1, Y0, Z0), (X0, Y0, Z0))); mats += [0, 0]   # floor
    quads.append(quad((X0, Y1, Z0), (X1, Y1, Z0), (X1, Y1, Z1), (X0, Y1, Z1))); mats += [0, 0]   # ceiling
    quads.append(quad((X0, Y0, Z0), (X1, Y0, Z0), (X1, Y1, Z0), (X0, Y1, Z0))); mats += [0, 0]   # back
    quads.append(quad((X1, Y0, Z1), (X0, Y0, Z1), (X0, Y1, Z1), (X1, Y1, Z1))); mats += [0, 0]   # front (behind camera)
    quads.append(quad((X0, Y0, Z1), (X0, Y0, Z0), (X0, Y1, Z0), (X0, Y1, Z1))); mats += [1, 1]   # left red
    quads.append(quad((X1, Y0, Z0), (X1, Y0, Z1), (X1, Y1, Z1), (X1, Y1, Z0))); mats += [2, 2]   # right green
    b = box((-1.6, -3.0, -5.2), (-0.4, -1.2, -4.0))
    quads.append(b)
    mats += [0] * 12
    room = np.concatenate(quads)
    sv, sn = icosphere(level, 0.3, center=(0.9, -2.7, -3.6))
    verts = np.concatenate([room, sv]).astype(np.float32)
    normals = np.concatenate([np.zeros_like(room), sn]).astype(np.float32)
    mat_id = np.array(mats + [0] * sv.shape[0], np.uint32)
    lights = np.zeros((1, 8), np.float32)
    lights[0, 0:3] = (0.0, 2.9, -4.0)
    lights[0, 4:7] = (100.0, 100.0, 100.0)
    cam = Camera(pos=(0.0, -0.75, 4.25), front=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0))
    return Scene("cornell", verts, normals, mat_id, np.stack([white, red, green]), lights, cam, seed)


def displaced_sphere(nu: int = 1024, nv: int = 512, seed: int = 1) -> Scene:
    """Radius-1 UV sphere (2*nu*nv triangles; 1,048,576 at the defaults)
    displaced by 0.05 * Perlin noise, one point light, camera at distance 2.4."""
    v, n = displaced_uv_sphere(nu, nv, 1.0, 0.05, 4.0, seed)
    mats = np.stack([material(kd=(0.7, 0.7, 0.7))])
    lights = np.zeros((1, 8), np.float32)
    lights[0, 0:3] = (2.0, 3.0, 2.5)
    lights[0, 4:7] = (20.0, 20.0, 20.0)
    cam = Camera(pos=(0.0, 0.0, 2.4), front=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_deg=45.0)
    return Scene("displaced_sphere", v, n, np.zeros(v.shape[0], np.uint32), mats, lights, cam, seed)


# Room around the instance grid (enclosed=True): x, y, z extents.  The
# instances span x in [-10.4, 10.4], z in [-23.5, -2.7], y in [-1, 1]; the
# lights sit at y 6..10 inside it and the camera at z = 8.
ROOM = ((-14.0, 14.0), (-1.0, 12.0), (-28.0, 12.0))


def instances(n_inst: int = 64, sphere_res=(256, 128), box_res: int = 35, seed: int = 3,
              enclosed: bool = True) -> Scene:
    """`n_inst` seeded instances (half displaced spheres, half boxes), 4 point
    lights, Lambertian albedo U[0.2,0.8] and GGX roughness U[0.1,0.8].

    enclosed=True (the default) puts them in a closed room, as the
    reference's scenes are closed rooms (main.cpp:164-345: floor, ceiling,
    back and side walls): floor, ceiling, back wall, a wall behind the
    camera and two coloured side walls, 12 triangles, all Lambertian, so
    every path ends on a surface and indirect light reaches every pixel.
    enclosed=False gives the open variant: a large ground quad under a
    black sky (most second-bounce rays escape).

    At the defaults: 32 x 65,536 + 32 x 14,700 + 12 = 2,567,564 triangles
    (open: + 2)."""
    rng = np.random.default_rng(seed)
    verts, norms, mids, mats = [], [], [], []
    side = int(math.ceil(math.sqrt(n_inst)))
    spacing = 2.6
    for i in range(n_inst):
        gx, gz = i % side, i // side
        cx = (gx - (side - 1) / 2) * spacing + rng.uniform(-0.3, 0.3)
        cz = -(gz * spacing) - 4.0 + rng.uniform(-0.3, 0.3)
        kd = rng.uniform(0.2, 0.8, 3)
        rough = rng.uniform(0.1, 0.8)
        ks = np.full(3, rng.uniform(0.04, 0.3))
        mats.append(material(kd=kd, ks=ks, roughness=rough))
        if i % 2 == 0:
            r = rng.uniform(0.6, 1.0)
            v, n = displaced_uv_sphere(sphere_res[0], sphere_res[1], r, 0.05, rng.uniform(2.0, 6.0),
                                       int(rng.integers(1 << 30)), center=(cx, r - 1.0, cz))
        else:
            h = rng.uniform(0.8, 1.8)
            w = rng.uniform(0.8, 1.6)
            v = subdivided_box((cx - w / 2, -1.0, cz - w / 2), (cx + w / 2, -1.0 + h, cz + w / 2), box_res)
            n = np.zeros_like(v)
        verts.append(v)
        norms.append(n)
        mids.append(np.full(v.shape[0], i, np.uint32))
    if enclosed:
        (X0, X1), (Y0, Y1), (Z0, Z1) = ROOM
        walls = [
            (quad((X0, Y0, Z1), (X1, Y0, Z1), (X1, Y0, Z0), (X0, Y0, Z0)), 0),   # floor
            (quad((X0, Y1, Z0), (X1, Y1, Z0), (X1, Y1, Z1), (X0, Y1, Z1)), 1),   # ceiling
            (quad((X0, Y0, Z0), (X1, Y0, Z0), (X1, Y1, Z0), (X0, Y1, Z0)), 1),   # back
            (quad((X1, Y0, Z1), (X0, Y0, Z1), (X0, Y1, Z1), (X1, Y1, Z1)), 1),   # behind the camera
            (quad((X0, Y0, Z1), (X0, Y0, Z0), (X0, Y1, Z0), (X0, Y1, Z1)), 2),   # left
            (quad((X1, Y0, Z0), (X1, Y0, Z1), (X1, Y1, Z1), (X1, Y1, Z0)), 3),   # right
        ]
        for q, k in walls:
            verts.append(q)
            norms.append(np.zeros_like(q))
            mids.append(np.full(2, n_inst + k, np.uint32))
        mats.append(material(kd=(0.6, 0.6, 0.6), ks=(0.04, 0.04, 0.04), roughness=0.6))   # floor
        mats.append(material(kd=(0.7, 0.7, 0.7)))     # ceiling, back and front walls
        mats.append(material(kd=(0.7, 0.25, 0.2)))    # left wall
        mats.append(material(kd=(0.2, 0.45, 0.7)))    # right wall
    else:
        ground = quad((-40.0, -1.0, 10.0), (40.0, -1.0, 10.0), (40.0, -1.0, -60.0), (-40.0, -1.0, -60.0))
        verts.append(ground)
        norms.append(np.zeros_like(ground))
        mids.append(np.full(2, n_inst, np.uint32))
        mats.append(material(kd=(0.6, 0.6, 0.6), ks=(0.04, 0.04, 0.04), roughness=0.6))
    lights = np.zeros((4, 8), np.float32)
    for k in range(4):
        lights[k, 0:3] = (rng.uniform(-10, 10), rng.uniform(6, 10), rng.uniform(-22, 2))
        lights[k, 4:7] = rng.uniform(150.0, 300.0, 3)
    cam = Camera(pos=(0.0, 4.0, 8.0), front=(0.0, -0.35, -1.0), up=(0.0, 1.0, 0.0), fov_deg=60.0)
    return Scene("instances" if enclosed else "instances_open", np.concatenate(verts).astype(np.float32),
                 np.concatenate(norms).astype(np.float32),
                 np.concatenate(mids), np.stack(mats), lights, cam, seed)


VIEW_RESOLUTIONS = ((256, 256), (384, 216), (512, 288), (640, 360), (768, 432), (854, 480), (960, 540),
                    (1024, 576), (1152, 648), (1280, 720), (1366, 768), (1440, 810), (1600, 900), (1680, 945),
                    (1792, 1008), (1920, 1080))


def views(n: int = 16, seed: int = 11):
    """`n` seeded camera poses over the instances scene, with resolutions from
    256x256 to 1920x1080 (VIEW_RESOLUTIONS, cycled): list of (Camera, W, H).

    Poses orbit the instance grid (centre (0, 0, -13)) at radius U[9, 16] and
    height U[1.5, 7], looking at a seeded point near the grid centre."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        phi = rng.uniform(0.0, 2.0 * math.pi)
        rad = rng.uniform(9.0, 16.0)
        pos = np.array([rad * math.cos(phi), rng.uniform(1.5, 7.0), -13.0 + rad * math.sin(phi)])
        target = np.array([rng.uniform(-3, 3), rng.uniform(-1, 1), -13.0 + rng.uniform(-3, 3)])
        front = target - pos
        W, H = VIEW_RESOLUTIONS[i % len(VIEW_RESOLUTIONS)]
        out.append((Camera(pos=tuple(pos), front=tuple(front / np.linalg.norm(front)), up=(0.0, 1.0, 0.0),
                           fov_deg=float(rng.uniform(40.0, 70.0))), W, H))
    return out


def analytic_records(items) -> np.ndarray:
    """items: (type, a, b, s, c, material) tuples -> (K, 12) float32 tfg_analytic."""
    out = np.zeros((len(items), 12), np.float32)
    u = out.view(np.uint32)
    for i, (typ, a, b, sval, c, mat) in enumerate(items):
        out[i, 0:3] = a
        u[i, 3] = typ
        out[i, 4:7] = b
        out[i, 7] = sval
        out[i, 8:11] = c
        u[i, 11] = mat
    return out


def reference_scene(mesh_level: int = 3, seed: int = 0) -> Scene:
    """The reference's base scene (main.cpp:164-345): five coloured planes, seven
    spheres (two glass, two plastic, three diffuse) and one point light, as
    analytic primitives, plus a small icosphere mesh so the CWBVH path runs."""
    PL, SP = 1, 0
    mats, items = [], []

    def add(kd=(0, 0, 0), ks=(0, 0, 0), kt=(0, 0, 0), ior=1.0):
        mats.append(material(kd=kd, ks=ks, kt=kt, ior=ior))
        return len(mats) - 1

    items.append((PL, (1, 0, 0), (0, 0, 0), -3.0, (0, 0, 0), add(kd=(0.1, 0.8, 0.9))))     # left
    items.append((PL, (-1, 0, 0), (0, 0, 0), -3.0, (0, 0, 0), add(kd=(0.9, 0.1, 0.6))))    # right
    items.append((PL, (0, 1, 0), (0, 0, 0), -3.0, (0, 0, 0), add(kd=(0.8, 0.8, 0.8))))     # floor
    items.append((PL, (0, -1, 0), (0, 0, 0), -3.0, (0, 0, 0), add(kd=(0.8, 0.8, 0.8))))    # ceiling
    items.append((PL, (0, 0, 1), (0, 0, 0), -8.0, (0, 0, 0), add(kd=(0.8, 0.8, 0.8))))     # back
    spheres = [((-1.5, -1.0, -3.8), 1.1, dict(ks=(0.04,) * 3, kt=(0.95,) * 3, ior=1.5)),
               ((1.8, 1.5, -4.5), 0.8, dict(kd=(0.9, 0.1, 0.1), ks=(0.3,) * 3)),
               ((0.2, -2.2, -4.0), 0.7, dict(kd=(0.3, 0.5, 0.2))),
               ((-1.2, 1.8, -5.0), 0.6, dict(kd=(0.2, 0.8, 0.2), ks=(0.2,) * 3)),
               ((2.0, -2.0, -4.2), 0.5, dict(ks=(0.04,) * 3, kt=(0.95,) * 3, ior=1.5)),
               ((-0.5, 0.5, -4.8), 0.6, dict(kd=(0.1, 0.1, 0.9), ks=(0.4,) * 3)),
               ((1.0, -0.5, -5.5), 0.9, dict(kd=(0.9, 0.8, 0.2)))]
    for c, r, kw in spheres:
        items.append((SP, c, (0, 0, 0), r, (0, 0, 0), add(**kw)))
    mesh_mat = add(kd=(0.75, 0.75, 0.75))
    v, n = icosphere(mesh_level, 0.5, center=(1.2, 0.6, -3.2))
    lights = np.zeros((1, 8), np.float32)
    lights[0, 0:3] = (0.0, 2.9, -4.0)
    lights[0, 4:7] = (100.0, 100.0, 100.0)
    cam = Camera(pos=(0.0, -0.75, 4.25), front=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0))
    return Scene("reference", v, n, np.full(v.shape[0], mesh_mat, np.uint32), np.stack(mats), lights, cam, seed,
                 analytic=analytic_records(items))


# main.cpp:769-799: the mesh models of the reference's "concurso" scene, as
# (file, scale, rotation X/Y/Z in degrees, translation, diffuse, specular).
# columna_suelo.glb and cabeza.glb are named there but absent from the
# reference's tree; their loads fail and add nothing, as they do there.
REFERENCE_MODELS = (
    ("concurso/columna2.glb", 0.5, (0.0, -90.0, 0.0), (-1.7, -3.3, -6.5), (0.75, 0.65, 0.55), (0.01,) * 3),
    ("concurso/columna1.glb", 0.5, (0.0, -90.0, 0.0), (1.7, -3.1, -2.5), (0.75, 0.65, 0.55), (0.01,) * 3),
    ("concurso/columna3.glb", 0.5, (0.0, 180.0, 0.0), (1.7, -3.4, -6.5), (0.75, 0.65, 0.55), (0.01,) * 3),
    ("concurso/columna_suelo.glb", 0.5, (0.0, -90.0, 0.0), (-1.7, -3.1, -2.5), (0.75, 0.65, 0.55), (0.01,) * 3),
    ("concurso/suelo2.glb", 0.5, (0.0, 0.0, 0.0), (-4.0, -4.8, -5.2), (0.1, 0.2, 0.4), (0.0,) * 3),
    ("concurso/suelo_grieta.glb", 1.0, (0.0, 0.0, 0.0), (0.0, -3.3, -5.2), (0.73, 0.73, 0.73), (0.0,) * 3),
    ("concurso/esfera_rota.glb", 0.5, (30.0, 110.0, 0.0), (1.0, -2.8, -2.0), (0.85, 0.4, 0.3), (0.0,) * 3),
    ("concurso/cabeza.glb", 0.15, (0.0, 80.0, 30.0), (0.0, -1.4, -2.3), (0.73, 0.73, 0.73), (0.0,) * 3),
)


def reference_models(model_dir: str, seed: int = 0) -> Scene:
    """The reference's mesh scene (main.cpp:745-830): the base scene's
    analytic planes, spheres and light (main.cpp:164-345) plus the glTF
    models of REFERENCE_MODELS loaded from `model_dir` (the reference's
    modelos/ folder) with tfg_amd.gltf and placed as cargarModelo places
    them.  Missing model files are skipped."""
    import os

    from . import gltf

    base = reference_scene(seed=seed)
    mats = [m for m in base.materials[:-1]]   # the base scene's icosphere material is not used
    verts, norms, mids = [], [], []
    for rel, scale, rot, trans, kd, ks in REFERENCE_MODELS:
        path = os.path.join(model_dir, rel)
        if not os.path.exists(path):
            continue
        v, n = gltf.load(path)
        v, n = gltf.place(v, n, scale, rot[0], rot[1], rot[2], trans)
        mats.append(material(kd=kd, ks=ks))
        verts.append(v)
        norms.append(n)
        mids.append(np.full(v.shape[0], len(mats) - 1, np.uint32))
    if not verts:
        raise FileNotFoundError(f"no reference model under {model_dir}")
    return Scene("reference_models", np.concatenate(verts), np.concatenate(norms), np.concatenate(mids),
                 np.stack(mats), base.lights, base.camera, seed, analytic=base.analytic)


def random_soup(n: int, seed: int, extent: float = 1.0) -> np.ndarray:
    """n random small triangles in [-extent, extent]^3 (tests)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-extent, extent, (n, 1, 3))
    e = rng.normal(0.0, extent * 0.15, (n, 3, 3))
    return (c + e).reshape(n, 9).astype(np.float32)
