"""CPU checks of the CWBVH builder (product host code) against the oracle.

Pins: orc_cwbvh_validate (every decoded child box contains its triangles,
every triangle record appears exactly once), oracle traversal == brute
force over all triangles (bit-exact), any-hit == closest-hit t < tmax, and
the reference's BVH2 + Moller-Trumbore path (bvh.cpp, primitiva.h)
agreeing with the watertight ids.
"""
import numpy as np
import pytest

import oracle
from tfg_amd import _lib, scenes, workloads

MESHES = {
    "soup": lambda: scenes.random_soup(3000, 3),
    "cornell": lambda: scenes.cornell(level=3).verts,
    "sphere": lambda: scenes.displaced_sphere(128, 64).verts,
    "one": lambda: scenes.random_soup(1, 4),
    "two": lambda: scenes.random_soup(2, 5),
    "coplanar": lambda: scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)),
}


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("name", list(MESHES))
def test_builder_valid(name, fmt):
    v = MESHES[name]()
    h = _lib.CwbvhHost(v, fmt)
    rc, depth = oracle.validate(h.nodes, h.tris, fmt, v)
    assert rc == 0
    assert depth == h.max_depth
    assert h.num_tris == v.shape[0]
    # triangle records carry primIdx in T3.w
    prims = np.sort(h.tris[:, 15].view(np.uint32))
    np.testing.assert_array_equal(prims, np.arange(v.shape[0], dtype=np.uint32))


@pytest.mark.parametrize("opts", [dict(max_leaf=1), dict(max_leaf=2, bins=8), dict(node_cost=1.0, tri_cost=3.0),
                                  dict(node_cost=1.0, tri_cost=0.05, bins=64)])
def test_builder_options_give_valid_trees(opts):
    """Every builder setting gives a valid tree whose traversal equals brute
    force; leaf slots never hold more than max_leaf triangles."""
    v = MESHES["sphere"]()
    h = _lib.CwbvhHost(v, 0, **opts)
    rc, depth = oracle.validate(h.nodes, h.tris, 0, v)
    assert rc == 0 and depth == h.max_depth
    meta = h.nodes[:, 6:8].copy().view(np.uint8).reshape(-1, 8)
    imask = (h.nodes[:, 3].view(np.uint32) >> 24).astype(np.uint8)
    slot_internal = (imask[:, None] >> np.arange(8, dtype=np.uint8)) & 1
    leaf_counts = np.array([bin(m >> 5).count("1") for m in meta.ravel()]).reshape(-1, 8)
    assert leaf_counts[slot_internal == 0].max() <= opts.get("max_leaf", 3)
    rays = workloads.random_rays(3000, 17, extent=1.5)
    a = oracle.cwbvh_trace(h.nodes, h.tris, 0, rays)
    b = oracle.brute_trace(v, rays)
    np.testing.assert_array_equal(a["prim"], b["prim"])
    np.testing.assert_array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_builder_rejects_bad_options():
    v = MESHES["soup"]()
    for bad in (dict(max_leaf=0), dict(max_leaf=4), dict(bins=2), dict(bins=65), dict(tri_cost=0.0)):
        with pytest.raises(_lib.TfgError):
            _lib.CwbvhHost(v, 0, **bad)


@pytest.mark.parametrize("name", ["soup", "cornell", "sphere", "coplanar"])
def test_traversal_equals_brute_force(name):
    v = MESHES[name]()
    h = _lib.CwbvhHost(v, 0)
    ext = 4.0 if name == "cornell" else 1.5
    rays = workloads.random_rays(4000, 21, extent=ext)
    if name == "coplanar":
        rng = np.random.default_rng(2)
        rays[:, 0:3] = np.c_[rng.uniform(-0.2, 1.2, (4000, 2)), np.full(4000, 1.0)]
        rays[:, 4:7] = (0.0, 0.0, -1.0)
    a = oracle.cwbvh_trace(h.nodes, h.tris, 0, rays)
    b = oracle.brute_trace(v, rays)
    assert (b["prim"] != 0xFFFFFFFF).sum() > 100
    for f in ("t", "u", "v"):
        np.testing.assert_array_equal(a[f].view(np.uint32), b[f].view(np.uint32))
    np.testing.assert_array_equal(a["prim"], b["prim"])


def test_edge_case_rays_equal_brute_force():
    """Rays through shared vertices and edges, axis-aligned rays with +0 and
    -0 direction components, tmax equal to the hit distance and one ulp
    below, rays starting on the surface, rays leaving a closed cube through
    its corners and edges, edge-on rays and zero-area triangles
    (tests/edge_cases.py): the oracle traversal equals brute force over all
    triangles bit for bit, any-hit equals closest t < tmax, and the
    watertight guarantees hold (no ray leaks through a shared edge or
    vertex, the exact t at the tmax boundary, no degenerate triangle hit)."""
    import edge_cases as E

    v = E.scene()
    h = _lib.CwbvhHost(v, 0)
    rc, _ = oracle.validate(h.nodes, h.tris, 0, v)
    assert rc == 0
    fam = E.families()
    hits = {}
    for name, rays in fam.items():
        a = oracle.cwbvh_trace(h.nodes, h.tris, 0, rays)
        b = oracle.brute_trace(v, rays)
        for f in ("t", "u", "v"):
            np.testing.assert_array_equal(a[f].view(np.uint32), b[f].view(np.uint32), err_msg=name)
        np.testing.assert_array_equal(a["prim"], b["prim"], err_msg=name)
        occ = oracle.cwbvh_occluded(h.nodes, h.tris, 0, rays)
        np.testing.assert_array_equal(occ, ((a["prim"] != 0xFFFFFFFF) & (a["t"] < rays[:, 7])).astype(np.uint32),
                                      err_msg=name)
        hits[name] = a
    E.check_properties(fam, hits)


def test_any_hit_matches_closest():
    v = MESHES["sphere"]()
    h = _lib.CwbvhHost(v, 0)
    rays = workloads.random_rays(4000, 3)
    rays[:, 7] = np.random.default_rng(0).uniform(0.05, 3.0, 4000)
    occ = oracle.cwbvh_occluded(h.nodes, h.tris, 0, rays)
    c = oracle.cwbvh_trace(h.nodes, h.tris, 0, rays)
    np.testing.assert_array_equal(occ, ((c["prim"] != 0xFFFFFFFF) & (c["t"] < rays[:, 7])).astype(np.uint32))


def test_affine_format_agrees_with_watertight():
    v = MESHES["sphere"]()
    rays = workloads.random_rays(4000, 8)
    hw = _lib.CwbvhHost(v, 0)
    ha = _lib.CwbvhHost(v, 1)
    a = oracle.cwbvh_trace(ha.nodes, ha.tris, 1, rays)
    w = oracle.cwbvh_trace(hw.nodes, hw.tris, 0, rays)
    same = a["prim"] == w["prim"]
    assert same.mean() > 0.999
    hit = same & (w["prim"] != 0xFFFFFFFF)
    # affine rows are FP32-rounded: error is absolute at scene scale
    np.testing.assert_allclose(a["t"][hit], w["t"][hit], rtol=1e-5, atol=1e-6)


def test_reference_bvh2_moller_trumbore_agrees():
    """The reference's CPU path (bvh.cpp median BVH2 + primitiva.h MT with FP64
    vectors) gives the same ids on all but edge rays, and t within 1e-5."""
    sc = scenes.displaced_sphere(128, 64)
    h = _lib.CwbvhHost(sc.verts, 0)
    rays = workloads.primary_rays(sc, 128, 128, seed=3)
    w = oracle.cwbvh_trace(h.nodes, h.tris, 0, rays)
    ref = oracle.Bvh2(sc.verts).trace(rays)
    same = w["prim"] == ref["prim"]
    assert same.mean() > 0.999
    hit = same & (w["prim"] != 0xFFFFFFFF)
    assert hit.sum() > 1000
    np.testing.assert_allclose(w["t"][hit], ref["t"][hit], rtol=1e-5)
    # every disagreement is an edge ray: both triangles are hit at ~equal t
    for i in np.nonzero(~same)[0]:
        assert w["prim"][i] != 0xFFFFFFFF and ref["prim"][i] != 0xFFFFFFFF
        assert abs(w["t"][i] - ref["t"][i]) <= 1e-5 * abs(ref["t"][i])


def test_builder_rejects_bad_input():
    with pytest.raises(_lib.TfgError):
        _lib.CwbvhHost(np.zeros((0, 9), np.float32))
    bad = scenes.random_soup(10, 1)
    bad[3, 4] = np.nan
    with pytest.raises(_lib.TfgError):
        _lib.CwbvhHost(bad)
    with pytest.raises(_lib.TfgError):
        _lib.CwbvhHost(scenes.random_soup(10, 1), fmt=5)


def test_library_exports_declared_symbols():
    import ctypes
    import re

    L = ctypes.CDLL(_lib.LIB_PATH)
    hdr = open(_lib.os.path.join(_lib._HERE, "..", "include", "tfg.h")).read()
    names = set(re.findall(r"\b(tfg_[a-z0-9_]+)\s*\(", hdr))
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.declared_symbols()) <= names
