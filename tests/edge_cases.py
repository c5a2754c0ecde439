"""Edge-case geometry and rays for the traversal parity tests.

Scene: an 8 x 8 grid of quad pairs on z = 0 over [-1, 1]^2 (vertices at
multiples of 1/4, exact in FP32), a closed cube [-0.5, 0.5]^2 x [1.5, 2.5]
and four zero-area triangles at z = 0.5.

Ray families (each a (n, 8) float32 array in tfg_ray layout, plus what the
watertight test guarantees for it):
  axis     - straight down from z = 1 through every grid vertex, every edge
             midpoint (shared edges, including the diagonals) and the cell
             centres, with +0 and -0 in the zero direction components: every
             ray hits the plane at exactly t = 1;
  tmax     - the same rays with tmax = 1 (accepted: t <= tmax) and with tmax
             one ulp below 1 (rejected);
  corners  - from the cube's centre toward its 8 corners, 12 edge midpoints,
             6 face centres and 2000 random directions: a closed mesh, so
             every ray hits it;
  surface  - starting on the plane (tmin = 0) pointing down and up: the plane
             hit at t = 0 is never accepted (t > tmin);
  oblique  - from random points above the plane aimed exactly at interior
             grid vertices (each shared by up to six triangles): every ray
             hits the plane;
  grazing  - inside the plane (direction z = 0), edge-on to its triangles;
  degenerate - aimed at the zero-area triangles, which are never reported.
"""
import numpy as np

GRID = 8
N_PLANE = 2 * GRID * GRID
N_CUBE = 12


def scene() -> np.ndarray:
    """(n, 9) float32 triangles: plane grid, cube, degenerate triangles."""
    tris = []
    s = 2.0 / GRID
    for j in range(GRID):
        for i in range(GRID):
            x0, y0 = -1.0 + i * s, -1.0 + j * s
            a, b, c, d = (x0, y0, 0.0), (x0 + s, y0, 0.0), (x0 + s, y0 + s, 0.0), (x0, y0 + s, 0.0)
            tris.append(a + b + c)
            tris.append(a + c + d)
    lo, hi = (-0.5, -0.5, 1.5), (0.5, 0.5, 2.5)
    corner = [(x, y, z) for z in (lo[2], hi[2]) for y in (lo[1], hi[1]) for x in (lo[0], hi[0])]
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    for q in faces:
        p = [corner[k] for k in q]
        tris.append(p[0] + p[1] + p[2])
        tris.append(p[0] + p[2] + p[3])
    tris.append((-0.3, -0.3, 0.5, 0.0, 0.0, 0.5, 0.3, 0.3, 0.5))     # collinear
    tris.append((0.2, -0.6, 0.5, 0.2, -0.6, 0.5, 0.2, -0.6, 0.5))     # a point
    tris.append((-0.6, 0.4, 0.5, -0.2, 0.4, 0.5, -0.6, 0.4, 0.5))     # repeated vertex
    tris.append((0.5, 0.5, 0.5, 0.5, 0.7, 0.5, 0.5, 0.6, 0.5))        # collinear along y
    return np.asarray(tris, np.float32)


def _rays(o, d, tmin=0.0, tmax=np.inf):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    r = np.zeros((o.shape[0], 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def plane_points():
    """Grid vertices, edge midpoints (horizontal, vertical, diagonal) and cell centres."""
    s = 2.0 / GRID
    k = -1.0 + s * np.arange(GRID + 1)
    m = -1.0 + s * (np.arange(GRID) + 0.5)
    pts = [(x, y) for y in k for x in k]
    pts += [(x, y) for y in k for x in m]
    pts += [(x, y) for y in m for x in k]
    pts += [(x, y) for y in m for x in m]          # cell centres = diagonal midpoints
    pts += [(x + s / 4, y + s / 4) for y in m for x in m]   # on the diagonals, off-centre
    return np.asarray(pts, np.float32)


def families(seed: int = 3) -> dict:
    rng = np.random.default_rng(seed)
    p = plane_points()
    n = p.shape[0]
    o = np.column_stack([p, np.ones(n, np.float32)])
    down = np.tile(np.float32([0.0, 0.0, -1.0]), (n, 1))
    down_neg0 = np.tile(np.float32([-0.0, -0.0, -1.0]), (n, 1))
    axis = np.concatenate([_rays(o, down), _rays(o, down_neg0)])
    below = np.nextafter(np.float32(1.0), np.float32(0.0))
    tmax = np.concatenate([_rays(o, down, tmax=1.0), _rays(o, down_neg0, tmax=1.0),
                           _rays(o, down, tmax=below), _rays(o, down_neg0, tmax=below)])
    c = np.float32([0.0, 0.0, 2.0])
    targets = [(x, y, z) for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (1.5, 2.5)]
    targets += [(0.0, y, z) for y in (-0.5, 0.5) for z in (1.5, 2.5)]
    targets += [(x, 0.0, z) for x in (-0.5, 0.5) for z in (1.5, 2.5)]
    targets += [(x, y, 2.0) for x in (-0.5, 0.5) for y in (-0.5, 0.5)]
    targets += [(0.0, 0.0, 1.5), (0.0, 0.0, 2.5), (0.0, -0.5, 2.0), (0.0, 0.5, 2.0), (-0.5, 0.0, 2.0),
                (0.5, 0.0, 2.0)]
    dirs = np.asarray(targets, np.float32) - c
    rnd = rng.normal(size=(2000, 3)).astype(np.float32)
    corners = _rays(np.tile(c, (dirs.shape[0] + 2000, 1)), np.concatenate([dirs, rnd]))
    sp = p[rng.choice(n, 100, replace=False)]
    so = np.column_stack([sp, np.zeros(100, np.float32)])
    surface = np.concatenate([_rays(so, np.tile(np.float32([0, 0, -1]), (100, 1))),
                              _rays(so, np.tile(np.float32([0, 0, 1]), (100, 1)))])
    # interior vertices only: a ray aimed at the mesh's outer boundary may
    # pass just outside it after the FP32 rounding of its direction
    inner = [j * (GRID + 1) + i for j in range(1, GRID) for i in range(1, GRID)]
    vi = rng.choice(inner, 2000)
    tv = np.column_stack([p[vi], np.zeros(2000, np.float32)])
    oo = np.column_stack([rng.uniform(-1, 1, (2000, 2)), rng.uniform(0.5, 1.5, 2000)]).astype(np.float32)
    oblique = _rays(oo, tv - oo)
    go = np.column_stack([rng.uniform(-1.5, 1.5, (100, 2)), np.zeros(100)]).astype(np.float32)
    gd = np.column_stack([rng.normal(size=(100, 2)), np.zeros(100)]).astype(np.float32)
    grazing = _rays(go, gd)
    degen_t = np.float32([[0.0, 0.0, 0.5], [0.2, -0.6, 0.5], [-0.4, 0.4, 0.5], [0.5, 0.6, 0.5],
                          [-0.15, -0.15, 0.5], [0.15, 0.15, 0.5]])
    do = np.column_stack([degen_t[:, :2], np.full(degen_t.shape[0], 0.9, np.float32)])
    degenerate = np.concatenate([_rays(do, degen_t - do), _rays(do, np.tile(np.float32([0, 0, -1]),
                                                                           (degen_t.shape[0], 1)))])
    return {"axis": axis, "tmax": tmax, "corners": corners, "surface": surface, "oblique": oblique,
            "grazing": grazing, "degenerate": degenerate}


def smallest_containing(points: np.ndarray) -> np.ndarray:
    """For points on the plane, the smallest index of the grid triangles
    that contain each one (edges and vertices included), by exact 2D edge
    functions in FP64 (the coordinates are multiples of 1/16): the tie rule
    (equal t: the smaller primIdx wins) restated independently of the
    traversal."""
    v = scene()[:N_PLANE].astype(np.float64).reshape(-1, 3, 3)[:, :, :2]
    out = np.full(points.shape[0], -1, np.int64)
    for k in range(N_PLANE - 1, -1, -1):
        a, b, c = v[k]
        def edge(p, q):
            return (q[0] - p[0]) * (points[:, 1] - p[1]) - (q[1] - p[1]) * (points[:, 0] - p[0])
        e = np.stack([edge(a, b), edge(b, c), edge(c, a)], 1)
        inside = np.all(e >= 0, 1) | np.all(e <= 0, 1)
        out[inside] = k
    return out


def check_properties(fam: dict, hits: dict):
    """The watertight guarantees of the families (hits: name -> HIT_DTYPE array)."""
    miss = np.uint32(0xFFFFFFFF)
    h = hits["axis"]
    assert np.all(h["prim"] < N_PLANE) and np.all(h["t"] == np.float32(1.0))
    np.testing.assert_array_equal(h["prim"].astype(np.int64),
                                  smallest_containing(fam["axis"][:, 0:2].astype(np.float64)))
    h = hits["tmax"]
    k = h.shape[0] // 2
    assert np.all(h["prim"][:k] < N_PLANE) and np.all(h["t"][:k] == np.float32(1.0))
    assert np.all(h["prim"][k:] == miss)
    h = hits["corners"]
    assert np.all((h["prim"] >= N_PLANE) & (h["prim"] < N_PLANE + N_CUBE))
    h = hits["surface"]
    k = h.shape[0] // 2
    assert np.all(h["prim"][:k] == miss)                      # plane at t = 0 not accepted
    assert np.all((h["prim"][k:] == miss) | (h["prim"][k:] >= N_PLANE))
    assert np.all(hits["oblique"]["prim"] < N_PLANE)
    for name in fam:
        assert not np.any(hits[name]["prim"][hits[name]["prim"] != miss] >= N_PLANE + N_CUBE), name
