"""CPU pins of the oracle's RNG, camera and single watertight triangle test.

RNG (rng.h:12-35): the oracle's uint32 recurrences against a restatement in
Python's unbounded integers reduced mod 2**32 by hand, so an overflow, a
signed shift or a wrong constant shows; the float draw against exact
arithmetic (u32 -> nearest float32, times 2**-32).
Camera (camara.h:29-59, RenderKernel.cu:68-90): invariants (orthonormal
basis, left = front x up, unit directions, the centre pixel maps to front,
the corner rays lie on the 45 degree field of view).
Watertight test (Woop, Benthin and Wald, JCGT 2013): hits at known t, u, v,
parallel and edge-on rays miss, the tmin / tmax ends, no leak through a
shared edge, and agreement with an FP64 Moller-Trumbore written here.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from tfg_amd import scenes

M32 = (1 << 32) - 1


def _pcg32(state: int):
    """rng.h:12-17 on unbounded Python integers: (new state, output)."""
    state = (state * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return state, (word >> 22) ^ word


def _seed(px: int, py: int, frame: int) -> int:
    """rng.h:25-35 on unbounded Python integers."""
    s = ((py << 16) & M32) | (px & 0xFFFF)
    s ^= ((frame & M32) * 719393) & M32
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 0x27D4EB2D) & M32
    s ^= s >> 15
    return s if s else 1


def test_pcg32_matches_unbounded_integers():
    L = oracle.lib()
    rng = np.random.default_rng(1)
    starts = [0, 1, 2, 0x7FFFFFFF, 0x80000000, M32] + [int(x) for x in rng.integers(0, 1 << 32, 40)]
    for s0 in starts:
        st = C.c_uint32(s0)
        s = s0
        for _ in range(16):
            s, want = _pcg32(s)
            assert L.orc_pcg32(C.byref(st)) == want
            assert st.value == s
    # the NumPy restatement used by the cache-batch tests agrees too
    from oracle import mlp_oracle
    st = rng.integers(0, 1 << 32, 4096, dtype=np.uint64).astype(np.uint32)
    want = [_pcg32(int(x)) for x in st]
    got = mlp_oracle.pcg32(st)
    np.testing.assert_array_equal(got, [w for _, w in want])
    np.testing.assert_array_equal(st, [s for s, _ in want])


def test_stream_draws_and_float_conversion():
    L = oracle.lib()
    n = 64
    for px, py, frame in [(0, 0, 0), (1919, 1079, 5), (65535, 3, 1 << 20), (17, 400, 2**31 - 1)]:
        u = np.zeros(n, np.uint32)
        f = np.zeros(n, np.float32)
        L.orc_stream(px, py, frame, n, u.ctypes.data, f.ctypes.data)
        s = _seed(px, py, frame)
        assert L.orc_seed(px, py, frame) == s
        for k in range(n):
            s, w = _pcg32(s)
            assert u[k] == w
            # exact: u32 -> nearest float32, times a power of two
            assert f[k] == np.float32(w) * np.float32(2.0 ** -32)
        assert np.all((f >= 0) & (f <= 1))
    # nearest rounding makes 1.0 reachable: the largest u32 draws round up
    assert np.float32(np.uint32(M32)) * np.float32(2.0 ** -32) == 1.0


def test_seed_matches_unbounded_integers_and_separates_pixels():
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(3000):
        px, py = int(rng.integers(0, 65536)), int(rng.integers(0, 32768))
        frame = int(rng.integers(0, 2**31))
        assert L.orc_seed(px, py, frame) == _seed(px, py, frame)
    # every pixel of a 1080p frame gets its own stream (the mix is a bijection
    # on uint32 and (py << 16) | px does not overlap)
    for frame in (0, 7):
        seeds = {_seed(x, y, frame) for y in range(0, 1080, 3) for x in range(1920)}
        assert len(seeds) == 1920 * 360 and 0 not in seeds


CAM = np.dtype([("pos", np.float64, 3), ("front", np.float64, 3), ("up", np.float64, 3), ("left", np.float64, 3),
                ("w", np.float32), ("h", np.float32), ("f", np.float32), ("pad", np.float32)])
assert CAM.itemsize == 112


def _camera_default(W, H):
    c = np.zeros(1, CAM)
    oracle.lib().orc_camera_default(W, H, c.ctypes.data)
    return c[0]


def _camera_ray(cam, x, y, W, H, ox, oy):
    c = np.array(cam, CAM)
    o = np.zeros(3)
    d = np.zeros(3)
    oracle.lib().orc_camera_ray(c.ctypes.data, x, y, W, H, C.c_float(ox), C.c_float(oy), o.ctypes.data,
                                d.ctypes.data)
    return o, d


def test_default_camera_fields():
    W, H = 1920, 1080
    c = _camera_default(W, H)
    np.testing.assert_array_equal(c["pos"], [0.0, -0.75, 4.25])
    np.testing.assert_array_equal(c["front"], [0.0, 0.0, -1.0])
    np.testing.assert_array_equal(c["up"], [0.0, 1.0, 0.0])
    np.testing.assert_array_equal(c["left"], np.cross(c["front"], [0.0, 1.0, 0.0]))
    assert abs(c["h"] - 2.0 * math.tan(math.radians(22.5))) < 1e-6
    assert abs(c["w"] - c["h"] * W / H) < 1e-6 and c["f"] == 1.0


def test_camera_rays():
    W, H = 1920, 1080
    c = _camera_default(W, H)
    o, d = _camera_ray(c, W // 2, H // 2, W, H, 0.0, 0.0)
    np.testing.assert_array_equal(o, c["pos"])
    np.testing.assert_array_equal(d, c["front"])          # centre pixel -> front
    rng = np.random.default_rng(3)
    for _ in range(500):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        _, d = _camera_ray(c, x, y, W, H, float(rng.random()), float(rng.random()))
        assert abs(np.linalg.norm(d) - 1.0) < 1e-14
    # the top-left corner lies on the half field of view vertically and on
    # atan(tan(22.5 deg) * W / H) horizontally; +x in the image is +left
    _, d = _camera_ray(c, 0, 0, W, H, 0.0, 0.0)
    fz = d @ c["front"]
    assert abs(math.degrees(math.atan2(d @ c["up"], fz)) - 22.5) < 1e-5
    assert abs(math.atan2(d @ c["left"], fz) + math.atan(math.tan(math.radians(22.5)) * W / H)) < 1e-6
    _, d2 = _camera_ray(c, W - 1, H - 1, W, H, 1.0, 1.0)
    np.testing.assert_allclose(d2, d * [-1, -1, 1], atol=1e-7)


def test_camera_basis_for_any_orientation():
    rng = np.random.default_rng(4)
    L = oracle.lib()
    for _ in range(50):
        pos, front, up = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        fov = float(rng.uniform(20, 100))
        c = np.zeros(1, CAM)
        L.orc_camera_make(pos.ctypes.data, front.ctypes.data, up.ctypes.data, C.c_double(fov), 640, 360,
                          c.ctypes.data)
        c = c[0]
        B = np.stack([c["front"], c["up"], c["left"]])
        np.testing.assert_allclose(B @ B.T, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(c["front"], front / np.linalg.norm(front), atol=1e-12)
        np.testing.assert_allclose(c["left"], np.cross(c["front"], up) / np.linalg.norm(np.cross(c["front"], up)),
                                   atol=1e-12)
        assert c["up"] @ up > 0
        assert abs(c["h"] - 2 * math.tan(math.radians(fov) / 2)) < 2e-6 * max(1.0, c["h"])
        # the host camera of the renderer (scenes.Camera) builds the same frame
        fr = scenes.Camera(tuple(pos), tuple(front), tuple(up), fov).frame(640, 360)
        np.testing.assert_allclose(fr[4:7], c["front"], atol=1e-7)
        np.testing.assert_allclose(fr[8:11], c["up"], atol=1e-7)
        np.testing.assert_allclose(fr[12:15], c["left"], atol=1e-7)
        # float32 tan of NumPy and of the C library may differ in the last bit
        for a, b in ((fr[11], c["h"]), (fr[7], c["w"])):
            assert abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32))) <= 2


def _wt(tri, o, d, tmin=0.0, tmax=np.inf):
    v = np.ascontiguousarray(tri, np.float32).reshape(9)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    t, u, w = C.c_float(), C.c_float(), C.c_float()
    hit = oracle.lib().orc_tri_watertight(v.ctypes.data, o.ctypes.data, d.ctypes.data, C.c_float(tmin),
                                          C.c_float(tmax), C.byref(t), C.byref(u), C.byref(w))
    return (t.value, u.value, w.value) if hit else None


UNIT = [0, 0, 0, 1, 0, 0, 0, 1, 0]


def test_axis_aligned_hits_at_known_t_u_v():
    for x, y in [(0.25, 0.5), (0.125, 0.125), (0.5, 0.25), (0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (0.5, 0.5)]:
        for z, dz in [(5.0, -1.0), (-3.0, 1.0), (2.0, -0.5)]:
            h = _wt(UNIT, [x, y, z], [0, 0, dz])
            assert h is not None, (x, y, z)
            assert h[0] == abs(z / dz) and h[1] == x and h[2] == y       # u: weight of v1, v: of v2
    assert _wt(UNIT, [0.6, 0.6, 1.0], [0, 0, -1]) is None                 # outside the hypotenuse
    assert _wt(UNIT, [-0.01, 0.5, 1.0], [0, 0, -1]) is None


def test_parallel_and_edge_on_rays_miss():
    assert _wt(UNIT, [-1, 0.25, 1.0], [1, 0, 0]) is None                   # parallel, above
    assert _wt(UNIT, [-1, 0.25, 0.0], [1, 0, 0]) is None                   # edge-on, inside the plane
    assert _wt(UNIT, [0.25, -1, 0.0], [0, 1, 0]) is None
    assert _wt(UNIT, [0.25, 0.25, 1.0], [0, 0, 1]) is None                 # pointing away


def test_tmin_tmax_ends():
    o, d = [0.25, 0.25, 2.0], [0, 0, -1]
    assert _wt(UNIT, o, d, tmax=2.0)[0] == 2.0                             # t <= tmax accepted
    assert _wt(UNIT, o, d, tmax=float(np.nextafter(np.float32(2), np.float32(0)))) is None
    assert _wt(UNIT, o, d, tmin=2.0) is None                               # t > tmin required
    assert _wt(UNIT, o, d, tmin=float(np.nextafter(np.float32(2), np.float32(0))))[0] == 2.0


def test_no_leak_through_a_shared_edge():
    # the unit square split along its diagonal x = y; rays aimed at points of
    # the diagonal (exact in FP32) from random origins hit at least one half
    a = [0, 0, 0, 1, 0, 0, 1, 1, 0]
    b = [0, 0, 0, 1, 1, 0, 0, 1, 0]
    rng = np.random.default_rng(5)
    for k in range(3000):
        s = float(rng.integers(1, 64)) / 64.0
        o = np.float32([rng.uniform(-2, 3), rng.uniform(-2, 3), rng.uniform(0.2, 3) * (1 if k % 2 else -1)])
        d = np.float32([s, s, 0.0]) - o
        assert _wt(a, o, d) is not None or _wt(b, o, d) is not None, (o, d)


def _mt64(tri, o, d):
    """Moller-Trumbore in FP64: (t, u, v) of the plane hit, or None if parallel."""
    v0, v1, v2 = np.asarray(tri, np.float64).reshape(3, 3)
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-30:
        return None
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    return (e2 @ q) / det, u, v


def test_agrees_with_fp64_moller_trumbore():
    rng = np.random.default_rng(6)
    checked = 0
    for _ in range(4000):
        tri = rng.normal(size=9).astype(np.float32)
        o = rng.normal(size=3).astype(np.float32) * 3
        target = tri.reshape(3, 3).astype(np.float64).T @ rng.dirichlet([1, 1, 1]) + rng.normal(size=3) * 0.3
        d = (target - o).astype(np.float32)
        ref = _mt64(tri, o.astype(np.float64), d.astype(np.float64))
        got = _wt(tri, o, d)
        if ref is None:
            continue
        t, u, v = ref
        margin = min(u, v, 1 - u - v)
        if abs(margin) < 1e-4 or abs(t) < 1e-4:
            continue                                  # within rounding of an edge or the origin
        checked += 1
        if margin > 0 and t > 0:
            assert got is not None
            assert abs(got[0] - t) <= 1e-5 * abs(t)
            assert abs(got[1] - u) <= 1e-5 and abs(got[2] - v) <= 1e-5
        else:
            assert got is None
    assert checked > 3500
