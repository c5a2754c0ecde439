"""Scenes whose rendered pixels have closed forms, for the oracle integrator
(CPU) and the HIP path tracer (GPU) alike.

Each case gives a scene, the image size, the frame, max depth and RR depth,
and ``check(img, rtol)`` for the (H*W, 3) image of that one frame.  The
expected values are derived here from geometry and the radiometry of
render.h, not from either renderer:
  furnace      a diffuse plane of albedo rho inside an emissive sphere:
               every plane pixel is exactly rho * Le (the cosine draw's
               weight kd / pd cancels cos / pdf), every sky pixel Le;
  mixed lobes  a plane with kd and a perfect mirror ks in the same sphere:
               each path picks diffuse with pd = max kd / (max kd + max ks)
               and carries kd / pd * Le, or ks / ps * Le for the mirror
               (render.h:290-341); the frequencies follow pd;
  point light  a Lambertian plane under a point light with a spherical
               occluder, depth 1: kd / pi * I * cos / max(d, 1)^2 from the
               shading point p + 1e-3 n (escenario.h:144-184), 0 in the
               shadow and where the occluder (black) is seen;
  glass lens   a glass sphere in front of an emissive quad: a pixel is
               kt^2 * Le when the ray, refracted in and out by Snell's law
               (written here in angle form), reaches the quad, else 0.
Pixels within a small margin of a silhouette, shadow or quad edge are left
out of the per-pixel checks.
"""
import math
from dataclasses import dataclass
from typing import Callable

import numpy as np

import oracle
from tfg_amd import scenes

SP, PL = 0, 1
# keeps the mesh non-empty; below every plane and outside every view used here
FAR_TRI = [(-0.1, -10, 0), (0.1, -10, 0), (0, -10, 0.1)]


@dataclass
class Case:
    name: str
    scene: scenes.Scene
    W: int
    H: int
    frame: int
    max_depth: int
    rr_depth: int
    check: Callable = None


def _scene(name, tri_verts, tri_mats, materials, analytic, lights, cam):
    v = np.asarray(tri_verts, np.float32).reshape(-1, 9)
    tri = v.reshape(-1, 3, 3)
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    lt = np.zeros((len(lights), 8), np.float32)
    for k, (p, i) in enumerate(lights):
        lt[k, 0:3], lt[k, 4:7] = p, i
    return scenes.Scene(name, v, np.tile(ng, (1, 3)).astype(np.float32), np.asarray(tri_mats, np.uint32),
                        np.stack(materials), lt, cam, analytic=scenes.analytic_records(analytic))


def primary(case: Case):
    """Camera rays of every pixel as render.h builds them (camara.h:42-59):
    jitter = the first two draws of the pixel's stream, plane offsets in
    FP32, direction normalised in FP64."""
    W, H = case.W, case.H
    fr = case.scene.camera.frame(W, H)
    jit = np.zeros((H * W, 2), np.float32)
    u32 = np.zeros(2, np.uint32)
    for y in range(H):
        for x in range(W):
            oracle.lib().orc_stream(x, y, case.frame, 2, u32.ctypes.data, jit[y * W + x].ctypes.data)
    f32 = np.float32
    xs = np.tile(np.arange(W, dtype=f32), H)
    ys = np.repeat(np.arange(H, dtype=f32), W)
    u = (xs + jit[:, 0]) / f32(W)
    v = (ys + jit[:, 1]) / f32(H)
    px = (f32(2) * u - f32(1)) * (fr[7] * f32(0.5))
    py = (f32(1) - f32(2) * v) * (fr[11] * f32(0.5))
    F = fr.astype(np.float64)
    d = F[4:7] * F[3] + px[:, None].astype(np.float64) * F[12:15] + py[:, None].astype(np.float64) * F[8:11]
    return F[0:3], d / np.linalg.norm(d, axis=1, keepdims=True)


def _close(img, want, rtol):
    if rtol == 0:
        np.testing.assert_array_equal(img, want)
    else:
        np.testing.assert_allclose(img, want, rtol=rtol, atol=0)


def furnace() -> Case:
    rho, Le = np.float32([0.3, 0.5, 0.7]), np.float32([1.0, 2.0, 0.5])
    mats = [scenes.material(kd=rho), scenes.material(le=Le)]
    cam = scenes.Camera((0.0, 1.0, 0.0), (0.0, -0.3, -1.0), (0.0, 1.0, 0.0))
    sc = _scene("furnace", FAR_TRI, [0], mats, [(PL, (0, 1, 0), (0, 0, 0), -1.0, (0, 0, 0), 0),
                                                (SP, (0, 0, 0), (0, 0, 0), 50.0, (0, 0, 0), 1)], [], cam)
    case = Case("furnace", sc, 48, 32, 1, 4, 10)

    def check(img, rtol=0.0):
        o, d = primary(case)
        with np.errstate(divide="ignore"):
            t = np.where(d[:, 1] < 0, -2.0 / d[:, 1], np.inf)
        r = np.linalg.norm(o + d * np.minimum(t, 1e9)[:, None], axis=1)
        on_plane, on_sky = r < 49.0, r > 51.0
        zero = np.all(img == 0, axis=1)
        assert zero.sum() <= 2          # a cosine draw at cos <= 1e-3 ends the path (probability 1e-6)
        sel = on_plane & ~zero
        _close(img[sel], np.broadcast_to((rho * Le).astype(np.float64), img[sel].shape), rtol)
        _close(img[on_sky], np.broadcast_to(Le.astype(np.float64), img[on_sky].shape), rtol)
        assert on_plane.sum() > case.W * case.H // 2 and on_sky.sum() > 50

    case.check = check
    return case


def mixed_lobes() -> Case:
    kd, ks, Le = np.float32([0.5, 0.3, 0.2]), np.float32([0.2, 0.3, 0.1]), np.float32([1.0, 2.0, 0.5])
    mats = [scenes.material(kd=kd, ks=ks), scenes.material(le=Le)]
    cam = scenes.Camera((0.0, 1.0, 0.0), (0.0, -1.0, -0.3), (0.0, 1.0, 0.0))
    sc = _scene("mixed_lobes", FAR_TRI, [0], mats, [(PL, (0, 1, 0), (0, 0, 0), -1.0, (0, 0, 0), 0),
                                                    (SP, (0, 0, 0), (0, 0, 0), 50.0, (0, 0, 0), 1)], [], cam)
    case = Case("mixed_lobes", sc, 64, 48, 11, 4, 10)

    def check(img, rtol=0.0):
        mkd, mks = float(kd.max()), float(ks.max())
        pd, ps = mkd / (mkd + mks), mks / (mkd + mks)
        vd = ((kd / np.float32(pd)) * Le).astype(np.float64)
        vs = ((ks / np.float32(ps)) * Le).astype(np.float64)
        tol = max(rtol, 1e-6)
        is_d = np.all(np.isclose(img, vd, rtol=tol, atol=0), axis=1)
        is_s = np.all(np.isclose(img, vs, rtol=tol, atol=0), axis=1)
        n = case.W * case.H
        assert (is_d | is_s).sum() >= n - 2
        assert abs(is_d.sum() / n - pd) < 4 * math.sqrt(pd * ps / n), (is_d.sum() / n, pd)

    case.check = check
    return case


def point_light() -> Case:
    kd, I = np.float32([0.8, 0.6, 0.4]), np.float32([10.0, 20.0, 5.0])
    lp = np.array([0.3, 0.7, -0.2])
    oc, orad = np.array([0.6, 0.35, 0.2]), 0.15
    mats = [scenes.material(kd=kd), scenes.material()]
    cam = scenes.Camera((0.0, 2.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0), 60.0)
    sc = _scene("point_light", FAR_TRI, [0], mats, [(PL, (0, 1, 0), (0, 0, 0), 0.0, (0, 0, 0), 0),
                                                    (SP, tuple(oc), (0, 0, 0), orad, (0, 0, 0), 1)], [(lp, I)], cam)
    case = Case("point_light", sc, 48, 48, 3, 1, 5)

    def check(img, rtol=2e-5):
        o, d = primary(case)
        tc = (oc - o) @ d.T
        seen_d = np.linalg.norm(o + tc[:, None] * d - oc, axis=1)     # closest approach of the camera ray
        p = o + d * (-o[1] / d[:, 1])[:, None]
        ps = p + np.array([0, 1e-3, 0])
        dl = lp - ps
        dist = np.linalg.norm(dl, axis=1)
        dl /= dist[:, None]
        tl = np.clip(np.einsum("ij,ij->i", oc - ps, dl), 0, dist)
        shadow_d = np.linalg.norm(ps + tl[:, None] * dl - oc, axis=1)
        want = (kd / np.pi) * I / np.maximum(dist, 1.0)[:, None] ** 2 * dl[:, 1:2]
        lit = (seen_d > orad * 1.01) & (shadow_d > orad * 1.01)
        dark = (seen_d < orad * 0.99) | (shadow_d < orad * 0.99)
        np.testing.assert_allclose(img[lit], want[lit], rtol=max(rtol, 2e-5))
        assert not img[dark].any()
        n = case.W * case.H
        assert lit.sum() > 0.8 * n and dark.sum() > 20
        assert (dist[lit] < 1.0).sum() > 100                       # the max(d, 1) clamp is exercised

    case.check = check
    return case


def _refract_angle(d, n, eta):
    """Snell's law in angle form: n faces the incoming ray, eta = n1 / n2."""
    ci = -d @ n
    tang = d + ci * n
    si = np.linalg.norm(tang)
    st = eta * si
    if st >= 1.0:
        return None
    if si < 1e-12:
        return d.copy()
    return -n * math.sqrt(1 - st * st) + tang / si * st


def glass_lens() -> Case:
    kt, Le = np.float32([0.9, 0.8, 0.7]), np.float32([2.0, 3.0, 4.0])
    ior, half, zq = 1.5, 0.6, -3.0
    mats = [scenes.material(le=Le), scenes.material(kt=kt, ior=ior)]
    quad = [(-half, -half, zq), (half, -half, zq), (half, half, zq), (-half, -half, zq), (half, half, zq),
            (-half, half, zq)]
    cam = scenes.Camera((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 30.0)
    sc = _scene("glass_lens", quad, [0, 0], mats, [(SP, (0, 0, 0), (0, 0, 0), 1.0, (0, 0, 0), 1)], [], cam)
    case = Case("glass_lens", sc, 64, 64, 5, 3, 8)

    def check(img, rtol=0.0):
        o, d = primary(case)
        n = case.W * case.H
        want = np.zeros((n, 3))
        sure = np.zeros(n, bool)
        lit_val = ((kt * kt) * Le).astype(np.float64)
        n_lens = 0
        for i in range(n):
            b = o @ d[i]
            disc = b * b - (o @ o - 1.0)
            if disc <= 0.02:                       # misses or grazes the sphere
                sure[i] = disc < -0.02
                continue
            p1 = o + d[i] * (-b - math.sqrt(disc))
            t1 = _refract_angle(d[i], p1, 1.0 / ior)
            p2 = p1 - 2 * (p1 @ t1) * t1           # far end of the chord
            t2 = _refract_angle(t1, -p2, ior)
            assert t2 is not None                  # a sphere never traps the ray
            n_lens += 1
            if t2[2] >= 0:
                sure[i] = True
                continue
            q = p2 + t2 * ((zq - p2[2]) / t2[2])
            m = half - np.abs(q[:2])
            sure[i] = abs(m).min() > 2e-3
            if m.min() > 0:
                want[i] = lit_val
        assert n_lens > 300 and 50 < (want[:, 0] > 0).sum() < n_lens - 50
        _close(img[sure], want[sure], rtol)

    case.check = check
    return case


CASES = (furnace, mixed_lobes, point_light, glass_lens)


# ------------------------------------------------------ cache capture

R1, R2 = np.float32([0.6, 0.5, 0.4]), np.float32([0.7, 0.8, 0.9])


def two_planes() -> Case:
    """A diffuse floor (y = -1, albedo R1) and ceiling (y = 1, albedo R2), no
    light: cache-mode vertices alternate floor, ceiling, floor ...  The
    primary vertex never takes the head (its footprint spread is zero,
    render.h:86-102), so the head / query is the first ceiling vertex: on
    y = 1 with normal code (0.5, 0, 0.5), kd R2, a downward outgoing
    direction and normaliser throughput x reflectance = R1 * R2 exactly.  A
    tail at vertex k carries the suffix throughput r(0) r(1) ... r(k) of the
    alternating albedos and lies on the plane of k's parity.  Nothing is
    lit: pixels and radiance suffixes are zero."""
    mats = [scenes.material(kd=R1), scenes.material(kd=R2)]
    cam = scenes.Camera((0.0, 0.0, 0.0), (0.0, -1.0, -0.5), (0.0, 1.0, 0.0))
    sc = _scene("two_planes", FAR_TRI, [0], mats, [(PL, (0, 1, 0), (0, 0, 0), -1.0, (0, 0, 0), 0),
                                                   (PL, (0, -1, 0), (0, 0, 0), -1.0, (0, 0, 0), 1)], [], cam)
    return Case("two_planes", sc, 48, 32, 4, 8, 20)


def check_heads(head, norm, atol=1e-6):
    """head: (n, 16) packed vertices; norm: (n, 3) normalisers."""
    n = head.shape[0]
    np.testing.assert_allclose(head[:, 1], 1.0, atol=atol)                      # on the ceiling
    np.testing.assert_array_equal(head[:, 8:11], np.broadcast_to([0.5, 0.0, 0.5], (n, 3)))
    np.testing.assert_array_equal(head[:, [7, 11, 12]], np.broadcast_to(R2, (n, 3)))
    assert not head[:, [3, 13, 14, 15]].any()
    dm = 2.0 * head[:, 4:7].astype(np.float64) - 1.0
    np.testing.assert_allclose(np.linalg.norm(dm, axis=1), 1.0, atol=10 * atol)
    assert (dm[:, 1] < 0).all()
    np.testing.assert_array_equal(norm, np.broadcast_to(R1 * R2, (n, 3)))


def check_tails(tail, tsuf, max_depth, atol=1e-6):
    """tail: (n, 16) packed tail vertices; tsuf: (n, 3) suffix throughputs."""
    chain, c = [], np.float32([1, 1, 1])
    for k in range(max_depth):
        c = c * (R1 if k % 2 == 0 else R2)
        chain.append(c.copy())
    on_floor = np.isclose(tail[:, 1], -1.0, atol=atol)
    assert (on_floor | np.isclose(tail[:, 1], 1.0, atol=atol)).all()
    for row, ts, fl in zip(tail, tsuf, on_floor):
        np.testing.assert_array_equal(row[[7, 11, 12]], R1 if fl else R2)
        assert any((k % 2 == 0) == fl and np.array_equal(chain[k], ts) for k in range(2, max_depth)), (ts, fl)
