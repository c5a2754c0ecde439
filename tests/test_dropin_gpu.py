"""launchKernelRenderTiny drop-in (tfg_launch_render_tiny) on the reference's
own buffers (Primitiva, LuzPuntual, Camara, ImagenGPU, DatosMLP,
RegistroEntrenamiento), compared per path with the oracle integrator on the
same scene: >= 95% of pixels agree within 1e-3 relative (remaining paths
diverge on an edge or in a glass sphere)."""
import ctypes as C

import numpy as np
import pytest

import conditioning
import oracle
from tfg_amd import _lib, scenes

pytestmark = pytest.mark.gpu

V3 = (np.float64, 3)
COL = (np.float32, 3)
LUZ = np.dtype([("pos", V3), ("I", COL)], align=True)
DATOS = np.dtype([("pos", V3), ("dir", V3), ("nrm", V3), ("kd", COL), ("ks", COL), ("color", COL)], align=True)
GEOM = np.dtype([("pos", V3), ("dir", V3), ("nrm", V3), ("kd", COL), ("ks", COL)], align=True)
REG = np.dtype([("head", GEOM), ("tail", GEOM), ("luz", COL), ("thr", COL), ("norm", COL), ("valido", np.uint8)],
               align=True)


# Primitiva (primitiva.h:17-39) with the union's members at their offsets.
PRIM = np.dtype({"names": ["tipo", "emision", "difuso", "especular", "transmision", "ior", "centro", "radio",
                           "pnormal", "pdist", "tri"],
                 "formats": [np.int32, COL, COL, COL, COL, np.float32, V3, np.float32, V3, np.float32, (np.float64, 18)],
                 "offsets": [0, 4, 16, 28, 40, 52, 56, 80, 56, 80, 56], "itemsize": 200})


def _prims_from_scene(sc):
    """Reference Primitiva records: analytic list + per-primIdx mesh copy."""
    mats = sc.materials
    au = sc.analytic.view(np.uint32)
    an = np.zeros(sc.analytic.shape[0], PRIM)
    for i, rec in enumerate(sc.analytic):
        typ, m = int(au[i, 3]), mats[au[i, 11]]
        an[i]["tipo"] = typ
        an[i]["emision"], an[i]["difuso"], an[i]["especular"] = m[12:15], m[0:3], m[4:7]
        an[i]["transmision"], an[i]["ior"] = m[8:11], m[7]
        if typ == 0:
            an[i]["centro"], an[i]["radio"] = rec[0:3], rec[7]
        else:
            an[i]["pnormal"], an[i]["pdist"] = rec[0:3], rec[7]
    mesh = np.zeros(sc.n_tris, PRIM)
    mesh["tipo"] = 2
    m = mats[sc.mat_id]
    mesh["difuso"], mesh["especular"], mesh["transmision"], mesh["ior"] = m[:, 0:3], m[:, 4:7], m[:, 8:11], m[:, 7]
    mesh["tri"][:, 0:9] = sc.verts
    mesh["tri"][:, 9:18] = sc.normals
    return an, mesh


def _dev(torch, a):
    return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()).cuda()


def _call(torch, d, host, W, H, frame, usar, entrenar, img, inf, thr, recs, counter):
    L = _lib.lib()
    im = _lib.ImagenGpu(W, H, img.data_ptr())
    _lib.check(L.tfg_launch_render_tiny(
        d["cam"].data_ptr(), d["an"].data_ptr(), d["n_an"], d["luz"].data_ptr(), 1, None, 0, W, H, 1, frame,
        host.desc, d["mesh"].data_ptr(), d["n_mesh"], im, None, recs.data_ptr(), counter.data_ptr(), W * H,
        inf.data_ptr(), thr.data_ptr(), int(usar), int(entrenar), None), "tfg_launch_render_tiny")
    torch.cuda.synchronize()


def test_render_tiny_dropin(torch_cuda):
    torch = torch_cuda
    sc = scenes.reference_scene()
    W, H = 96, 64
    an, mesh = _prims_from_scene(sc)
    host = _lib.CwbvhHost(sc.verts)
    dbvh = host.upload()
    luz = np.zeros(1, LUZ)
    luz["pos"], luz["I"] = sc.lights[0, 0:3], sc.lights[0, 4:7]
    cam = torch.zeros(112, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().tfg_launch_inicializar_camara(cam.data_ptr(), W, H, None))
    d = {"cam": cam, "an": _dev(torch, an), "n_an": an.shape[0], "luz": _dev(torch, luz), "mesh": _dev(torch, mesh),
         "n_mesh": mesh.shape[0]}
    img = torch.zeros((H * W, 3), device="cuda")
    inf = torch.zeros(W * H * 112, dtype=torch.uint8, device="cuda")
    thr = torch.full((H * W, 3), 7.0, device="cuda")
    recs = torch.zeros(W * H * 232, dtype=torch.uint8, device="cuda")
    counter = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().tfg_dropin_configure(8, 5))
    # plain frame (both flags false, RenderKernel.cu:185-200)
    _call(torch, d, dbvh, W, H, 4, False, False, img, inf, thr, recs, counter)
    g = img.cpu().numpy()
    assert (thr.cpu().numpy() == 0).all()            # render.h:457
    assert counter.item() == 0
    R = oracle.Renderer(sc, host.nodes, host.tris, W, H)
    ref = R.render(4, 1, 8, 5)[0]
    ok = conditioning.agree(g, ref)
    n_bad, n_expl, base = conditioning.explain(R, g, ref, 4, 8, 5)
    print(f"plain agreement {ok.mean()}; {n_bad} disagreeing paths, {n_expl} ill-conditioned in the oracle "
          f"(base rate {base:.4f})", flush=True)
    assert ok.mean() >= 0.95
    assert n_expl >= 0.9 * n_bad
    # cache frame: 5% training paths, the rest queries
    _call(torch, d, dbvh, W, H, 5, True, True, img, inf, thr, recs, counter)
    n_rec = counter.item()
    o = oracle.Renderer(sc, host.nodes, host.tris, W, H).render_nrc(5, 8, 5, 3)
    heads = int(((o["flags"] & 3) == 3).sum())
    assert abs(n_rec - heads) <= max(3, 0.05 * heads)
    t = thr.cpu().numpy()
    need = ((o["flags"] & 4) != 0) | ((o["flags"] & 3) == 3)
    has = np.any(t != 0, axis=1)
    assert (has == need).mean() >= 0.95
    dd = np.frombuffer(inf.cpu().numpy().tobytes(), DATOS)
    sel = np.nonzero(has & need)[0]
    pos_ok = np.all(np.abs(dd["pos"][sel] - o["head"][sel, 0:3]) <= 1e-3, axis=1)
    assert pos_ok.mean() >= 0.95
    r = np.frombuffer(recs.cpu().numpy().tobytes(), REG)[:n_rec]
    assert (r["valido"] == 1).all() and np.isfinite(r["luz"]).all()
    # bad arguments are refused, nothing is launched
    with pytest.raises(_lib.TfgError):
        _lib.check(_lib.lib().tfg_launch_render_tiny(
            None, None, 0, None, 0, None, 0, W, H, 1, 0, dbvh.desc, d["mesh"].data_ptr(), d["n_mesh"],
            _lib.ImagenGpu(W, H, img.data_ptr()), None, None, None, 0, None, thr.data_ptr(), 0, 0, None))


def _foreign_desc(torch, host):
    """The host tree's node and record arrays in device buffers this library
    never saw (a tree from another builder, as tinybvh's upload makes one)."""
    nodes = torch.from_numpy(np.ascontiguousarray(host.nodes)).cuda()
    tris = torch.from_numpy(np.ascontiguousarray(host.tris)).cuda()
    return _lib.CwbvhGpu(nodes.data_ptr(), tris.data_ptr(), host.num_nodes, host.num_tris), (nodes, tris)


@pytest.mark.parametrize("source", ["uploaded_affine", "foreign_affine"])
def test_render_tiny_affine_tree_matches_watertight(torch_cuda, source):
    """A CWBVH in tinybvh's layout (80 B nodes, affine unit-triangle records
    with primIdx in T3.w, tinybvh_wrapper.cpp:3,53-75): built here in
    TFG_TRI_AFFINE and uploaded by tfg_cwbvh_upload, or the same arrays in
    device buffers this library never saw (then taken as tinybvh's layout,
    its depth found by walking the nodes).  The drop-in traces its triangles
    with the watertight test on the vertices of the Primitiva copy, so the
    image, the cache queries and throughputs and the record count equal
    those of the watertight-format tree of the same scene, bit for bit."""
    torch = torch_cuda
    sc = scenes.reference_scene()
    W, H = 96, 64
    an, mesh = _prims_from_scene(sc)
    hw = _lib.CwbvhHost(sc.verts, _lib.TFG_TRI_WATERTIGHT)
    ha = _lib.CwbvhHost(sc.verts, _lib.TFG_TRI_AFFINE)
    np.testing.assert_array_equal(hw.nodes, ha.nodes)      # same tree, other triangle records
    assert not np.array_equal(hw.tris, ha.tris)
    dw = hw.upload()
    if source == "uploaded_affine":
        da = ha.upload()
        keep = None
    else:
        desc_a, keep = _foreign_desc(torch, ha)
        da = type("ForeignTree", (), {"desc": desc_a})()
    luz = np.zeros(1, LUZ)
    luz["pos"], luz["I"] = sc.lights[0, 0:3], sc.lights[0, 4:7]
    cam = torch.zeros(112, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().tfg_launch_inicializar_camara(cam.data_ptr(), W, H, None))
    d = {"cam": cam, "an": _dev(torch, an), "n_an": an.shape[0], "luz": _dev(torch, luz), "mesh": _dev(torch, mesh),
         "n_mesh": mesh.shape[0]}
    _lib.check(_lib.lib().tfg_dropin_configure(8, 5))
    out = {}
    for name, tree in (("watertight", dw), ("affine", da)):
        res = []
        for frame, usar, entrenar in ((4, False, False), (5, True, True)):
            print(f"tree={name} source={source} frame={frame} flags=({int(usar)},{int(entrenar)})", flush=True)
            img = torch.zeros((H * W, 3), device="cuda")
            inf = torch.zeros(W * H * 112, dtype=torch.uint8, device="cuda")
            thr = torch.full((H * W, 3), 7.0, device="cuda")
            recs = torch.zeros(W * H * 232, dtype=torch.uint8, device="cuda")
            counter = torch.zeros(1, dtype=torch.int32, device="cuda")
            _call(torch, d, tree, W, H, frame, usar, entrenar, img, inf, thr, recs, counter)
            res.append((img.cpu().numpy(), thr.cpu().numpy(), inf.cpu().numpy(), counter.item()))
        out[name] = res
    for (iw, tw, fw, cw), (ia, ta, fa, ca) in zip(out["watertight"], out["affine"]):
        np.testing.assert_array_equal(iw, ia)
        np.testing.assert_array_equal(tw, ta)
        np.testing.assert_array_equal(fw, fa)
        assert cw == ca
    assert out["affine"][0][0].sum() > 0 and out["affine"][1][3] > 0
    del keep


def test_tracer_affine_scene_and_malformed_foreign_tree(torch_cuda):
    """tfg_pt_set_scene with an affine tree and its vertices renders the
    watertight tree's image bit for bit (plain and cache-capture frames);
    an affine tree without vertices, a format that contradicts the
    uploaded tree's, and a foreign tree whose child index leaves the node
    array are refused on the host before any launch."""
    from tfg_amd import pt

    torch = torch_cuda
    sc = scenes.instances(n_inst=8, sphere_res=(32, 16), box_res=4, seed=5)
    W, H = 96, 64
    ds = pt.DeviceScene(sc, torch)
    ha = _lib.CwbvhHost(sc.verts, _lib.TFG_TRI_AFFINE)
    da = ha.upload()
    verts = torch.from_numpy(np.ascontiguousarray(sc.verts, np.float32)).cuda()
    aff = _lib.PtScene()
    C.pointer(aff)[0] = ds.desc
    aff.bvh = da.desc
    aff.tri_format = _lib.TFG_TRI_AFFINE
    aff.d_verts = verts.data_ptr()
    aff.n_verts = sc.n_tris
    for mode in (0, 2):
        imgs = []
        for desc in (ds.desc, aff):
            tr = pt.PathTracer(W, H, 8, 3, mode=mode, use_graph=False)
            _lib.check(_lib.lib().tfg_pt_set_scene(tr._h, C.byref(desc)), "tfg_pt_set_scene")
            tr.set_camera(sc.camera.frame(W, H))
            img = torch.zeros((H * W, 3), device="cuda")
            tr.render(9, img)
            torch.cuda.synchronize()
            imgs.append((img.cpu().numpy(), tr.rays_last_frame()))
            tr.close()
        np.testing.assert_array_equal(imgs[0][0], imgs[1][0])
        assert imgs[0][1] == imgs[1][1]
    tr = pt.PathTracer(W, H, 8, 3, mode=0, use_graph=False)
    bad = _lib.PtScene()
    C.pointer(bad)[0] = aff
    bad.d_verts = None
    assert _lib.lib().tfg_pt_set_scene(tr._h, C.byref(bad)) == -1
    bad = _lib.PtScene()
    C.pointer(bad)[0] = ds.desc
    bad.tri_format = _lib.TFG_TRI_AFFINE          # the uploaded tree is watertight
    bad.d_verts = verts.data_ptr()
    bad.n_verts = sc.n_tris
    assert _lib.lib().tfg_pt_set_scene(tr._h, C.byref(bad)) == -1
    nodes = ha.nodes.copy()
    nodes.view(np.uint32)[0, 4] = ha.num_nodes + 5   # root's first child beyond the array
    foreign, keep = _foreign_desc(torch, type("H", (), {"nodes": nodes, "tris": ha.tris, "num_nodes": ha.num_nodes,
                                                        "num_tris": ha.num_tris})())
    bad = _lib.PtScene()
    C.pointer(bad)[0] = aff
    bad.bvh = foreign
    rc = _lib.lib().tfg_pt_set_scene(tr._h, C.byref(bad))
    assert rc == -1 and b"malformed" in _lib.lib().tfg_last_error()
    tr.close()
    del keep
