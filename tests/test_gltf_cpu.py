"""glTF 2.0 reader (tfg_amd.gltf), the replacement of the reference's Assimp
import (cargarModelo.h:83-224).

Pins: (1) synthetic files written here, with world coordinates worked out by
hand (node hierarchy with translation / quaternion rotation / scale and a
column-major `matrix`, interleaved buffers with byteStride, u8/u16/u32
indices, strips and fans, data-URI .gltf and .glb containers, normals by the
inverse transpose, face normals when NORMAL is absent); (2) the reference's
own model files when /root/reference is present: every primitive's decoded
POSITION bounds equal the accessor's declared min/max (float32, exact) and the
triangle count equals the index counts / 3; (3) `place` (cargarModelo.h:
21-81) against an independent extrinsic X, Y, Z rotation.
"""
import base64
import json
import math
import os
import struct

import numpy as np
import pytest

from tfg_amd import gltf, scenes

MODELS = "/root/reference/modelos"


def _doc(bin_bytes, nodes, meshes, accessors, views, uri=None):
    buf = {"byteLength": len(bin_bytes)}
    if uri is not None:
        buf["uri"] = uri
    return {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}], "nodes": nodes, "meshes": meshes,
            "accessors": accessors, "bufferViews": views, "buffers": [buf]}


def _write_gltf(path, doc_fn, bin_bytes):
    uri = "data:application/octet-stream;base64," + base64.b64encode(bin_bytes).decode()
    with open(path, "w") as f:
        json.dump(doc_fn(uri), f)


def _write_glb(path, doc_fn, bin_bytes):
    js = json.dumps(doc_fn(None)).encode()
    js += b" " * (-len(js) % 4)
    bb = bin_bytes + b"\0" * (-len(bin_bytes) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bb), 0x004E4942) + bb
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)


def _one_triangle_file(tmp_path, kind, root_node, with_normals=True, index_type=5123):
    # interleaved position + normal (stride 24), one triangle
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.tile(np.array([1.0, 1.0, 0.0], np.float32) / np.float32(math.sqrt(2)), (3, 1))
    inter = np.concatenate([pos, nrm], 1).astype(np.float32).tobytes()
    it = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[index_type]
    idx = np.array([0, 1, 2], it).tobytes()
    idx += b"\0" * (-len(idx) % 4)
    bin_bytes = inter + idx
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": len(inter), "byteStride": 24},
             {"buffer": 0, "byteOffset": len(inter), "byteLength": 3 * np.dtype(it).itemsize}]
    accessors = [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3", "min": [0, 0, 0], "max": [1, 1, 0]},
                 {"bufferView": 0, "byteOffset": 12, "componentType": 5126, "count": 3, "type": "VEC3"},
                 {"bufferView": 1, "componentType": index_type, "count": 3, "type": "SCALAR"}]
    attrs = {"POSITION": 0, "NORMAL": 1} if with_normals else {"POSITION": 0}
    meshes = [{"primitives": [{"attributes": attrs, "indices": 2}]}]
    nodes = [root_node, {"scale": [2.0, 1.0, 1.0], "mesh": 0}]
    path = tmp_path / f"tri.{kind}"
    fn = lambda uri: _doc(bin_bytes, nodes, meshes, accessors, views, uri)   # noqa: E731
    (_write_gltf if kind == "gltf" else _write_glb)(path, fn, bin_bytes)
    return str(path)


# root: translate (1, 2, 3), rotate 90 degrees about z; child: scale x by 2
ROOT_TRS = {"translation": [1.0, 2.0, 3.0], "rotation": [0.0, 0.0, math.sin(math.pi / 4), math.cos(math.pi / 4)],
            "children": [1]}
# the same root as a column-major matrix
ROOT_MATRIX = {"matrix": [0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 1, 2, 3, 1], "children": [1]}
# scale (2,1,1) gives (0,0,0),(2,0,0),(0,1,0); z-rotation (x,y)->(-y,x); then + (1,2,3)
WORLD = np.array([[1, 2, 3, 1, 4, 3, 0, 2, 3]], np.float64)


@pytest.mark.parametrize("kind", ["gltf", "glb"])
@pytest.mark.parametrize("root", [ROOT_TRS, ROOT_MATRIX], ids=["trs", "matrix"])
@pytest.mark.parametrize("index_type", [5121, 5123, 5125])
def test_hierarchy_transforms_and_containers(tmp_path, kind, root, index_type):
    v, n = gltf.load(_one_triangle_file(tmp_path, kind, root, index_type=index_type))
    np.testing.assert_allclose(v, WORLD, atol=1e-6)
    # normal (1,1,0)/sqrt2 through scale (2,1,1): inverse transpose gives (1/2, 1, 0); the z-rotation
    # then gives (-1, 1/2, 0), normalised
    e = np.array([-1.0, 0.5, 0.0]) / math.sqrt(1.25)
    np.testing.assert_allclose(n, np.tile(e, (1, 3)), atol=1e-6)


def test_face_normals_without_normal_attribute(tmp_path):
    v, n = gltf.load(_one_triangle_file(tmp_path, "glb", ROOT_TRS, with_normals=False))
    p = v.reshape(3, 3).astype(np.float64)
    fn = np.cross(p[1] - p[0], p[2] - p[0])
    fn /= np.linalg.norm(fn)
    np.testing.assert_allclose(n, np.tile(fn, (1, 3)), atol=1e-6)
    np.testing.assert_allclose(fn, [0, 0, 1], atol=1e-7)   # winding kept: still facing +z


def test_strips_and_fans():
    idx = np.arange(5)
    s = gltf._triangle_indices(5, idx)
    np.testing.assert_array_equal(s, [[0, 1, 2], [2, 1, 3], [2, 3, 4]])   # every second triangle flipped
    f = gltf._triangle_indices(6, idx)
    np.testing.assert_array_equal(f, [[0, 1, 2], [0, 2, 3], [0, 3, 4]])
    assert gltf._triangle_indices(1, idx).shape == (0, 3)   # lines: no triangles


def test_place_matches_extrinsic_rotation():
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(3)
    v = rng.normal(size=(50, 9)).astype(np.float32)
    n = rng.normal(size=(50, 9)).astype(np.float32)
    for rx, ry, rz in ((0, 0, 0), (30, 110, 0), (0, -90, 0), (0, 80, 30), (12, -45, 170)):
        pv, pn = gltf.place(v, n, 0.5, rx, ry, rz, (1.0, -2.8, -2.0))
        r = Rotation.from_euler("xyz", [rx, ry, rz], degrees=True)   # extrinsic: X, then Y, then Z
        ev = r.apply(v.reshape(-1, 3).astype(np.float64) * 0.5) + np.array([1.0, -2.8, -2.0])
        np.testing.assert_allclose(pv.reshape(-1, 3), ev, atol=2e-6)
        en = r.apply(n.reshape(-1, 3).astype(np.float64))
        en /= np.linalg.norm(en, axis=1, keepdims=True)
        np.testing.assert_allclose(pn.reshape(-1, 3), en, atol=2e-6)
    # no rotation: scale and translation only, exactly
    pv, _ = gltf.place(v, n, 1.0, 0, 0, 0, (0.0, 0.0, 0.0))
    np.testing.assert_array_equal(pv, v)


def _model_files():
    if not os.path.isdir(MODELS):
        return []
    out = []
    for d in sorted(os.listdir(MODELS)):
        for f in sorted(os.listdir(os.path.join(MODELS, d))):
            if f.endswith((".gltf", ".glb")):
                out.append(os.path.join(d, f))
    return out


@pytest.mark.skipif(not os.path.isdir(MODELS), reason="the reference's model files are not present")
@pytest.mark.parametrize("rel", _model_files())
def test_reference_models_decode_to_their_declared_bounds(rel):
    path = os.path.join(MODELS, rel)
    doc, bufs = gltf.read_document(path)
    want_tris = 0
    for node_i, _ in gltf.world_matrices(doc):
        node = doc["nodes"][node_i]
        if "mesh" not in node:
            continue
        for prim in doc["meshes"][node["mesh"]]["primitives"]:
            a = doc["accessors"][prim["attributes"]["POSITION"]]
            p = gltf.accessor(doc, bufs, prim["attributes"]["POSITION"])
            assert p.dtype == np.float32 and p.shape == (a["count"], 3)
            np.testing.assert_array_equal(p.min(0), np.float32(a["min"]))
            np.testing.assert_array_equal(p.max(0), np.float32(a["max"]))
            assert prim.get("mode", 4) == 4
            want_tris += (doc["accessors"][prim["indices"]]["count"] if "indices" in prim else a["count"]) // 3
    v, n = gltf.load(path)
    assert v.shape == (want_tris, 9) and n.shape == (want_tris, 9)
    assert np.all(np.isfinite(v))
    np.testing.assert_allclose(np.linalg.norm(n.reshape(-1, 3), axis=1), 1.0, atol=1e-5)


@pytest.mark.skipif(not os.path.isdir(MODELS), reason="the reference's model files are not present")
def test_reference_model_scene():
    sc = scenes.reference_models(MODELS)
    present = [m for m in scenes.REFERENCE_MODELS if os.path.exists(os.path.join(MODELS, m[0]))]
    assert len(present) == 6   # columna_suelo.glb and cabeza.glb are absent from the reference's tree
    assert sc.n_tris == sum(gltf.load(os.path.join(MODELS, m[0]))[0].shape[0] for m in present)
    assert sc.materials.shape[0] == int(sc.mat_id.max()) + 1
    assert sc.analytic is not None and sc.lights.shape[0] == 1
    # the scene builds into a valid CWBVH (every decoded box holds its triangles)
    import oracle
    from tfg_amd import _lib

    h = _lib.CwbvhHost(sc.verts, 0)
    rc, depth = oracle.validate(h.nodes, h.tris, 0, sc.verts)
    assert rc == 0 and depth == h.max_depth
