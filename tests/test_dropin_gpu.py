"""launchKernelRenderTiny drop-in (tfg_launch_render_tiny) on the reference's
own buffers (Primitiva, LuzPuntual, Camara, ImagenGPU, DatosMLP,
RegistroEntrenamiento), compared per path with the oracle integrator on the
same scene: >= 95% of pixels agree within 1e-3 relative (remaining paths
diverge on an edge or in a glass sphere)."""
import ctypes as C

import numpy as np
import pytest

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
    ref = oracle.Renderer(sc, host.nodes, host.tris, W, H).render(4, 1, 8, 5)[0]
    ok = np.all(np.abs(g - ref) <= 1e-3 * (1 + np.abs(ref)), axis=1)
    print("plain agreement", ok.mean(), flush=True)
    assert ok.mean() >= 0.95
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
