"""GPU parity of the CWBVH traversal kernels against the CPU oracle.

Bar: bit-identical (t, u, v, primIdx) and bit-identical node/triangle visit
counts against oracle.cwbvh_trace on the same rays (both run the same FP32
operations in the same order), and any-hit flags equal to the oracle's.
"""
import numpy as np
import pytest

import oracle
from tfg_amd import _lib, scenes, workloads

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _trace_gpu(torch, dbvh, rays, counts=False):
    n = rays.shape[0]
    r = _dev(torch, rays)
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    cnt = torch.empty((n, 2), dtype=torch.int32, device="cuda") if counts else None
    dbvh.trace_closest(r, hits, n, counts=cnt)
    torch.cuda.synchronize()
    h = hits.cpu().numpy().view(np.uint32)
    return (h, cnt.cpu().numpy().view(np.uint32)) if counts else h


def _any_gpu(torch, dbvh, rays):
    n = rays.shape[0]
    r = _dev(torch, rays)
    occ = torch.empty(n, dtype=torch.int32, device="cuda")
    dbvh.trace_any(r, occ, n)
    torch.cuda.synchronize()
    return occ.cpu().numpy().view(np.uint32)


def _as_u32(h):
    out = np.zeros((h.shape[0], 4), np.uint32)
    out[:, 0] = h["t"].view(np.uint32)
    out[:, 1] = h["u"].view(np.uint32)
    out[:, 2] = h["v"].view(np.uint32)
    out[:, 3] = h["prim"]
    return out


CASES = [
    ("soup", lambda: scenes.random_soup(4000, 11)),
    ("cornell", lambda: scenes.cornell(level=4).verts),
    ("sphere256", lambda: scenes.displaced_sphere(256, 128).verts),
]


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("name,make", CASES)
def test_closest_bit_exact(torch_cuda, name, make, fmt):
    torch = torch_cuda
    v = make()
    print(f"case={name} fmt={fmt} tris={v.shape[0]}", flush=True)
    host = _lib.CwbvhHost(v, fmt)
    dbvh = host.upload()
    rays = workloads.random_rays(20000, 5, extent=4.0 if name == "cornell" else 1.5)
    g, gc = _trace_gpu(torch, dbvh, rays, counts=True)
    c, cc = oracle.cwbvh_trace(host.nodes, host.tris, fmt, rays, counts=True)
    assert (g[:, 3] != 0xFFFFFFFF).sum() > 1000
    np.testing.assert_array_equal(g, _as_u32(c))
    np.testing.assert_array_equal(gc, cc)
    # counting and plain variants agree
    np.testing.assert_array_equal(_trace_gpu(torch, dbvh, rays), g)


@pytest.mark.parametrize("name,make", CASES)
def test_any_hit(torch_cuda, name, make):
    torch = torch_cuda
    v = make()
    host = _lib.CwbvhHost(v, 0)
    dbvh = host.upload()
    rays = workloads.random_rays(20000, 9, extent=4.0 if name == "cornell" else 1.5)
    rays[:, 7] = np.random.default_rng(1).uniform(0.1, 3.0, rays.shape[0]).astype(np.float32)
    g = _any_gpu(torch, dbvh, rays)
    c = oracle.cwbvh_occluded(host.nodes, host.tris, 0, rays)
    np.testing.assert_array_equal(g, c)
    closest = oracle.cwbvh_trace(host.nodes, host.tris, 0, rays)
    np.testing.assert_array_equal(g, ((closest["prim"] != 0xFFFFFFFF) & (closest["t"] < rays[:, 7])).astype(np.uint32))


@pytest.mark.parametrize("fmt", [0, 1])
def test_edge_case_rays_bit_exact(torch_cuda, fmt):
    """The edge-case families of tests/edge_cases.py (shared vertices and
    edges, +0/-0 direction components, the tmax boundary, rays starting on
    the surface, a closed cube's corners and edges, edge-on rays, zero-area
    triangles): GPU closest hits and visit counts equal the oracle's bit for
    bit and any-hit flags equal it, in both triangle formats; for the
    watertight format the no-leak guarantees hold on the GPU results too."""
    import edge_cases as E

    torch = torch_cuda
    v = E.scene()
    host = _lib.CwbvhHost(v, fmt)
    dbvh = host.upload()
    hits = {}
    for name, rays in E.families().items():
        print(f"family={name} fmt={fmt} rays={rays.shape[0]}", flush=True)
        g, gc = _trace_gpu(torch, dbvh, rays, counts=True)
        c, cc = oracle.cwbvh_trace(host.nodes, host.tris, fmt, rays, counts=True)
        np.testing.assert_array_equal(g, _as_u32(c), err_msg=name)
        np.testing.assert_array_equal(gc, cc, err_msg=name)
        occ = _any_gpu(torch, dbvh, rays)
        np.testing.assert_array_equal(occ, oracle.cwbvh_occluded(host.nodes, host.tris, fmt, rays), err_msg=name)
        h = np.zeros(rays.shape[0], oracle.HIT_DTYPE)
        h["t"] = g[:, 0].view(np.float32)
        h["prim"] = g[:, 3]
        hits[name] = h
    if fmt == 0:
        E.check_properties(E.families(), hits)


def test_traversal_microbench_full_size(torch_cuda):
    """Full-size displaced sphere (2^20 triangles) with 2^21 primary, 2^21
    bounce and 2^21 shadow rays: every GPU result is checked against the
    oracle on a seeded 2^16-ray sample of each set and by size-independent
    properties on all rays (any-hit == closest-hit t < tmax)."""
    torch = torch_cuda
    sc = scenes.displaced_sphere()
    host = _lib.CwbvhHost(sc.verts, 0)
    dbvh = host.upload()
    prim = workloads.primary_rays(sc, 2048, 1024, seed=7)
    hp = _trace_gpu(torch, dbvh, prim)
    hits = np.zeros(prim.shape[0], oracle.HIT_DTYPE)
    hits["t"] = hp[:, 0].view(np.float32)
    hits["prim"] = hp[:, 3]
    assert (hp[:, 3] != 0xFFFFFFFF).mean() > 0.4
    bounce = workloads.bounce_rays(sc, prim, hits, seed=8)
    shadow = workloads.shadow_rays(sc, prim, hits, seed=9)
    hb = _trace_gpu(torch, dbvh, bounce)
    occ = _any_gpu(torch, dbvh, shadow)
    hs = _trace_gpu(torch, dbvh, shadow)
    shadow_closest = (hs[:, 3] != 0xFFFFFFFF) & (hs[:, 0].view(np.float32) < shadow[:, 7])
    np.testing.assert_array_equal(occ, shadow_closest.astype(np.uint32))
    rng = np.random.default_rng(10)
    for rays, g in ((prim, hp), (bounce, hb)):
        sel = np.sort(rng.choice(rays.shape[0], 1 << 16, replace=False))
        c = oracle.cwbvh_trace(host.nodes, host.tris, 0, np.ascontiguousarray(rays[sel]))
        np.testing.assert_array_equal(g[sel], _as_u32(c))
    sel = np.sort(rng.choice(shadow.shape[0], 1 << 16, replace=False))
    c = oracle.cwbvh_occluded(host.nodes, host.tris, 0, np.ascontiguousarray(shadow[sel]))
    np.testing.assert_array_equal(occ[sel], c)


def test_empty_batch_and_bad_args(torch_cuda):
    torch = torch_cuda
    host = _lib.CwbvhHost(scenes.random_soup(10, 1), 0)
    dbvh = host.upload()
    r = torch.zeros((1, 8), dtype=torch.float32, device="cuda")
    h = torch.zeros((1, 4), dtype=torch.int32, device="cuda")
    dbvh.trace_closest(r, h, 0)   # n == 0 is a no-op
    with pytest.raises(_lib.TfgError):
        _lib.check(_lib.lib().tfg_trace_closest(None, 0, r.data_ptr(), h.data_ptr(), None, 1, None))
    with pytest.raises(_lib.TfgError):
        _lib.check(_lib.lib().tfg_trace_closest(_lib.C.byref(dbvh.desc), 7, r.data_ptr(), h.data_ptr(), None, 1, None))


def test_ray_setup_reciprocal_is_the_ieee_division(torch_cuda):
    """The reciprocal of the ray set-up (v_rcp_f32 + one Newton step, the
    division for |d| >= 2^126 and non-normal d) equals 1.0f / d bit for bit
    on every normal float and on zeros, subnormals, infinities and NaN
    (tests/cpp/rcp_check.hip, built by __graft_entry__.build())."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "rcp_check")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 differ; " in p.stdout and p.stdout.rstrip().endswith("0 differ")


def test_node_test_scaled_slabs_match_unscaled(torch_cuda):
    """The node test reads the quantised bounds as FP16 subnormals (b * 2^-24)
    through v_fma_mix_f32 and compares slab distances scaled by 2^-24; on
    134M pseudo-random nodes and rays (about 10M of them hitting some child;
    origins on the node's planes, zero direction components and nonzero tmin
    among them) its hit masks equal those of the unscaled test with
    v_cvt_f32_ubyte bounds (tests/cpp/node_check.hip, built by
    __graft_entry__.build())."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "node_check")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.rstrip().endswith("; 0 differ")
