"""CPU pins of the oracle integrator's shading against closed forms.

Sampling and BRDF (render.h:383-390 and the GGX lobe of this build):
  - the cosine draw is cos(theta) = sqrt(1 - r2), phi = 2 pi r2 with
    E[cos] = 2/3, E[cos^2] = 1/2;
  - GGX D integrates to one over the projected hemisphere;
  - Smith G1 equals the tan form 2 / (1 + sqrt(1 + a^2 tan^2));
  - the visible normals integrate to one (Heitz 2014);
  - the BRDF is reciprocal;
  - the visible-normal draw reproduces the moments of its density;
  - E[G1(wi)] over draws equals the lobe's albedo integrated from the BRDF.
Analytic primitives (primitiva.h:137-187): distances of spheres and planes
at known points, tangent, parallel, behind and inside cases, triangles
against an FP64 Moller-Trumbore.
Integrator (render.h:33-466): the scenes of closed_form_scenes.py (furnace,
mixed diffuse and mirror lobes, point light over a Lambertian plane, glass
lens), whose pixels have closed forms; cache-mode records between two
diffuse planes.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import closed_form_scenes
import conditioning
import oracle
from tfg_amd import _lib, scenes

SP, PL = 0, 1


def _cos_local(r1, r2):
    out = np.zeros(3)
    oracle.lib().orc_cosine_local(r1, r2, out.ctypes.data)
    return out


def test_cosine_draw():
    g = (np.arange(128) + 0.5) / 128
    w = np.array([_cos_local(a, b) for a in g for b in g])
    r1 = np.repeat(g, 128)
    r2 = np.tile(g, 128)
    np.testing.assert_allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(w[:, 2], np.sqrt(1 - r2), atol=1e-14)
    np.testing.assert_allclose(np.mod(np.arctan2(w[:, 1], w[:, 0]), 2 * np.pi), 2 * np.pi * r1, atol=1e-12)
    assert abs(w[:, 2].mean() - 2 / 3) < 1e-4 and abs((w[:, 2] ** 2).mean() - 0.5) < 1e-4
    assert abs(w[:, 0].mean()) < 1e-12 and abs(w[:, 1].mean()) < 1e-12


def _D(a, nh):
    return oracle.lib().orc_ggx_D(a, nh)


def _G1(a, nv):
    return oracle.lib().orc_ggx_G1(a, nv)


ALPHAS = (0.1, 0.3, 0.6, 1.0)


def test_ggx_D_normalised():
    for a in ALPHAS:
        val, _ = integrate.quad(lambda mu: 2 * math.pi * _D(a, mu) * mu, 0, 1, points=[1 - a, 1 - a / 4], limit=200)
        assert abs(val - 1.0) < 1e-6, (a, val)


def test_ggx_G1_tan_form():
    for a in ALPHAS:
        for nv in np.linspace(0.02, 1.0, 50):
            tan2 = (1 - nv * nv) / (nv * nv)
            assert abs(_G1(a, nv) - 2 / (1 + math.sqrt(1 + a * a * tan2))) < 1e-12
        assert _G1(a, 0.0) == 0.0 and abs(_G1(a, 1.0) - 1.0) < 1e-15


def _sph(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def _hemi_grid(n_t=400, n_p=200):
    """Midpoint quadrature on the hemisphere: directions and solid-angle weights."""
    th = (np.arange(n_t) + 0.5) * (math.pi / 2) / n_t
    ph = (np.arange(n_p) + 0.5) * 2 * math.pi / n_p
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (math.pi / 2 / n_t) * (2 * math.pi / n_p)
    return _sph(T, P).reshape(-1, 3), w.reshape(-1)


def test_ggx_visible_normals_and_sampling():
    h, dw = _hemi_grid()
    Dv = np.vectorize(_D)
    g = (np.arange(160) + 0.5) / 160
    for a in (0.3, 0.6):
        D = Dv(a, h[:, 2])
        for th_o in (0.3, 1.0, 1.4):
            wo = np.array([math.sin(th_o), 0.0, math.cos(th_o)])
            dens = _G1(a, wo[2]) * np.maximum(h @ wo, 0) * D / wo[2]     # D_wo(h), Heitz 2014
            assert abs((dens * dw).sum() - 1.0) < 2e-3, (a, th_o)
            hs = np.zeros((g.size * g.size, 3))
            for i, u1 in enumerate(g):
                for j, u2 in enumerate(g):
                    oracle.lib().orc_ggx_vndf(wo.ctypes.data, a, u1, u2, hs[i * g.size + j].ctypes.data)
            np.testing.assert_allclose(np.linalg.norm(hs, axis=1), 1.0, atol=1e-12)
            for f in (lambda v: v[:, 2], lambda v: v[:, 0], lambda v: v[:, 0] ** 2):
                assert abs(f(hs).mean() - (f(h) * dens * dw).sum()) < 3e-3, (a, th_o)
            # reflected directions: E[G1(wi)] (the path weight with F = 1)
            # equals the albedo of the lobe integrated from the BRDF
            wi = 2 * (hs @ wo)[:, None] * hs - wo
            g1 = np.where(wi[:, 2] > 0, np.vectorize(_G1)(a, np.maximum(wi[:, 2], 0)), 0.0)
            one = np.ones(3, np.float32)
            f = np.zeros(3, np.float32)
            alb = 0.0
            for k in range(0, h.shape[0], 7):
                oracle.lib().orc_ggx_eval(one.ctypes.data, a, wo.ctypes.data, h[k].ctypes.data, f.ctypes.data)
                alb += f[0] * h[k, 2] * dw[k] * 7
            assert abs(g1.mean() - alb) < 1e-2, (a, th_o, g1.mean(), alb)


def test_ggx_eval_reciprocal_and_one_sided():
    rng = np.random.default_rng(7)
    ks = np.float32([0.04, 0.5, 1.0])
    f1, f2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for _ in range(300):
        a = rng.uniform(0.05, 1)
        wo, wi = rng.normal(size=3), rng.normal(size=3)
        wo[2], wi[2] = abs(wo[2]), abs(wi[2])
        wo /= np.linalg.norm(wo)
        wi /= np.linalg.norm(wi)
        oracle.lib().orc_ggx_eval(ks.ctypes.data, a, wo.ctypes.data, wi.ctypes.data, f1.ctypes.data)
        oracle.lib().orc_ggx_eval(ks.ctypes.data, a, wi.ctypes.data, wo.ctypes.data, f2.ctypes.data)
        np.testing.assert_allclose(f1, f2, rtol=1e-6)
        assert f1[2] >= f1[1] >= f1[0] > 0                                # Schlick grows with F0
        wi[2] = -wi[2]
        oracle.lib().orc_ggx_eval(ks.ctypes.data, a, wo.ctypes.data, wi.ctypes.data, f1.ctypes.data)
        assert not f1.any()


def _an(typ, a, s=0.0, b=(0, 0, 0), c=(0, 0, 0)):
    return scenes.analytic_records([(typ, a, b, s, c, 0)])


def _t(rec, o, d):
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    return oracle.lib().orc_analytic_t(rec.ctypes.data, o.ctypes.data, d.ctypes.data)


def test_analytic_sphere_and_plane():
    s = _an(SP, (0, 0, -5), 2.0)
    assert _t(s, (0, 0, 0), (0, 0, -1)) == 3.0                 # near root
    assert _t(s, (0, 0, -5), (0, 0, -1)) == 2.0                # from the centre: far root
    assert abs(_t(s, (0, 0, -6), (1, 0, 0)) - math.sqrt(3.0)) < 1e-6
    assert _t(s, (0, 0, 0), (0, 0, 1)) == -1.0                 # behind
    assert _t(s, (0, 2.5, 0), (0, 0, -1)) == -1.0              # passes above
    assert abs(_t(s, (0, 2.0, 0), (0, 0, -1)) - 5.0) < 1e-3    # tangent: one root
    p = _an(PL, (0, 1, 0), -3.0)                               # y = -3
    assert _t(p, (0, 0, 0), (0, -1, 0)) == 3.0
    assert abs(_t(p, (1, 1, 1), (0, -0.5, math.sqrt(0.75))) - 8.0) < 1e-6
    assert _t(p, (0, 0, 0), (0, 1, 0)) == -1.0                 # behind
    assert _t(p, (0, 0, 0), (1, 0, 0)) == -1.0                 # parallel
    assert abs(_t(p, (0, -4, 0), (0, 1, 0)) - 1.0) < 1e-7      # from the back side


def test_analytic_triangle_matches_fp64_moller_trumbore():
    rng = np.random.default_rng(8)
    n = 0
    for _ in range(2000):
        v = rng.normal(size=(3, 3))
        rec = _an(2, v[0], 0.0, v[1], v[2])
        vv = rec[0].reshape(3, 4)[:, :3].astype(np.float64)
        o = rng.normal(size=3) * 3
        d = vv.T @ rng.dirichlet([1, 1, 1]) + rng.normal(size=3) * 0.3 - o
        d /= np.linalg.norm(d)
        e1, e2 = vv[1] - vv[0], vv[2] - vv[0]
        pv = np.cross(d, e2)
        det = e1 @ pv
        sv = o - vv[0]
        u = sv @ pv / det
        qv = np.cross(sv, e1)
        w = d @ qv / det
        t = e2 @ qv / det
        if min(abs(u), abs(w), abs(1 - u - w), abs(det)) < 1e-3 or abs(t) < 1e-3:
            continue
        n += 1
        got = _t(rec, o, d)
        if u > 0 and w > 0 and u + w < 1 and t > 0:
            assert abs(got - t) < 1e-4 * max(1, t)
        else:
            assert got == -1.0
    assert n > 1500




@pytest.mark.parametrize("make", closed_form_scenes.CASES, ids=lambda f: f.__name__)
def test_integrator_closed_form_scenes(make):
    case = make()
    host = _lib.CwbvhHost(case.scene.verts, _lib.TFG_TRI_WATERTIGHT)
    ren = oracle.Renderer(case.scene, host.nodes, host.tris, case.W, case.H)
    img, _ = ren.render(case.frame, 1, case.max_depth, case.rr_depth)
    case.check(img)


def test_cache_capture_between_two_planes():
    """The closed forms of closed_form_scenes.two_planes on the oracle's
    cache-mode paths: training heads and tails (mode 2), queries (mode 1)."""
    case = closed_form_scenes.two_planes()
    host = _lib.CwbvhHost(case.scene.verts, _lib.TFG_TRI_WATERTIGHT)
    ren = oracle.Renderer(case.scene, host.nodes, host.tris, case.W, case.H)
    for mode in (2, 1):
        out = ren.render_nrc(case.frame, case.max_depth, case.rr_depth, mode)
        fl = out["flags"]
        assert not out["pixel"].any() and not out["lsuf"].any()
        sel = (fl & 2) != 0 if mode == 2 else (fl & 4) != 0
        assert sel.mean() > 0.95 and (fl & 1 == 1).all() == (mode == 2)
        closed_form_scenes.check_heads(out["head"][sel], out["norm"][sel])
        tl = (fl & 8) != 0
        if mode == 1:
            assert not tl.any()
            continue
        assert tl.sum() > 0.5 * case.W * case.H and not (tl & ~sel).any()
        closed_form_scenes.check_tails(out["tail"][tl], out["tsuf"][tl], case.max_depth)


def test_conditioning_probe():
    """The oracle's conditioning probe (Renderer.pixels_perturbed), which the
    per-path GPU tests use to show that their disagreeing paths are
    rounding-driven:
      * at delta 0 it reproduces the integrator bit for bit (it draws from
        its own stream, never from the path's PCG32 stream);
      * on the point-light scene every pixel whose camera ray and shadow ray
        pass farther than 1% of the occluder's radius from its silhouette is
        well-conditioned at the tests' delta;
      * a planted systematic error (1% too bright on 3% of the pixels of a
        Cornell box frame, as a wrong branch would give) is reported as
        ill-conditioned at about the base rate, far below the 90% bar."""
    case = closed_form_scenes.point_light()
    host = _lib.CwbvhHost(case.scene.verts, _lib.TFG_TRI_WATERTIGHT)
    W, H = case.W, case.H
    ren = oracle.Renderer(case.scene, host.nodes, host.tris, W, H)
    img, _ = ren.render(case.frame, 1, case.max_depth, case.rr_depth)
    ys, xs = np.mgrid[0:H, 0:W]
    xy = np.stack([xs.ravel(), ys.ravel()], 1)
    p0 = ren.pixels_perturbed(xy, case.frame, case.max_depth, case.rr_depth, 0.0, 5)
    assert np.array_equal(p0, img.astype(np.float32))
    ill = conditioning.ill_conditioned(ren, np.arange(W * H), img, case.frame, case.max_depth, case.rr_depth)
    o, d = closed_form_scenes.primary(case)
    oc, orad, lp = np.array([0.6, 0.35, 0.2]), 0.15, np.array([0.3, 0.7, -0.2])
    seen_d = np.linalg.norm(o + ((oc - o) @ d.T)[:, None] * d - oc, axis=1)
    ps = o + d * (-o[1] / d[:, 1])[:, None] + np.array([0, 1e-3, 0])
    dl = lp - ps
    dist = np.linalg.norm(dl, axis=1)
    dl /= dist[:, None]
    tl = np.clip(np.einsum("ij,ij->i", oc - ps, dl), 0, dist)
    shadow_d = np.linalg.norm(ps + tl[:, None] * dl - oc, axis=1)
    far = (np.abs(seen_d - orad) > 0.01 * orad) & (np.abs(shadow_d - orad) > 0.01 * orad)
    assert far.sum() > 0.9 * W * H
    assert not ill[far].any()

    sc = scenes.cornell(level=3)
    host = _lib.CwbvhHost(sc.verts, _lib.TFG_TRI_WATERTIGHT)
    W, H, md, rr = 64, 48, 5, 2
    ren = oracle.Renderer(sc, host.nodes, host.tris, W, H)
    ref, _ = ren.render(2, 1, md, rr)
    got = ref.copy()
    rng = np.random.default_rng(1)
    lit = np.nonzero(ref.max(1) > 0.15)[0]   # where 1% exceeds the tolerance
    planted = rng.choice(lit, size=int(0.03 * W * H), replace=False)
    got[planted] *= 1.01
    n_bad, n_expl, base = conditioning.explain(ren, got, ref, 2, md, rr)
    assert n_bad == planted.size
    assert n_expl < 0.3 * n_bad and base < 0.3
