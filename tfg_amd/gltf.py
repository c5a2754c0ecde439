"""glTF 2.0 mesh reader (.gltf with external or data-URI buffers, and .glb).

Replaces the Assimp import the reference uses to load its meshes
(cargarModelo.h:83-224, flags Triangulate | GenNormals |
PreTransformVertices): every node that references a mesh contributes its
triangle primitives in world space (the node hierarchy's transforms applied,
skinning ignored as a pre-transform import ignores it), normals transformed
by the inverse transpose and normalised, and a face normal where a primitive
has no NORMAL attribute.  `place` then applies the reference's own model
transform: scale, rotations about X, Y, Z in degrees, translation
(cargarModelo.h:21-81,146-163), with its FP32 rounding of each rotated
coordinate.

Supported: triangle lists, strips and fans (modes 4, 5, 6), indexed or not,
u8/u16/u32 indices, float positions and normals with any byteStride, node
`matrix` or translation/rotation/scale.  Sparse accessors and
quantised (integer) positions are rejected with a ValueError.
"""
from __future__ import annotations

import base64
import json
import os
import struct

import numpy as np

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_GLB_MAGIC = 0x46546C67   # "glTF"
_CHUNK_JSON = 0x4E4F534A
_CHUNK_BIN = 0x004E4942


def read_document(path: str):
    """(glTF JSON dict, list of buffer bytes) of a .gltf or .glb file."""
    with open(path, "rb") as f:
        data = f.read()
    base = os.path.dirname(os.path.abspath(path))
    bin_chunk = None
    if len(data) >= 12 and struct.unpack_from("<I", data, 0)[0] == _GLB_MAGIC:
        _, version, length = struct.unpack_from("<III", data, 0)
        if version != 2:
            raise ValueError(f"{path}: GLB version {version}, expected 2")
        off, doc = 12, None
        while off + 8 <= min(length, len(data)):
            clen, ctype = struct.unpack_from("<II", data, off)
            chunk = data[off + 8: off + 8 + clen]
            if ctype == _CHUNK_JSON:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == _CHUNK_BIN and bin_chunk is None:
                bin_chunk = bytes(chunk)
            off += 8 + clen
        if doc is None:
            raise ValueError(f"{path}: GLB without a JSON chunk")
    else:
        doc = json.loads(data.decode("utf-8"))
    buffers = []
    for k, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError(f"{path}: buffer {k} has no uri and there is no GLB binary chunk")
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                buffers.append(f.read())
    return doc, buffers


def accessor(doc, buffers, index: int) -> np.ndarray:
    """Accessor `index` as a (count, components) array of its component type."""
    a = doc["accessors"][index]
    if "sparse" in a:
        raise ValueError(f"accessor {index}: sparse accessors are not supported")
    dt = np.dtype(_COMPONENT[a["componentType"]])
    nc = _NCOMP[a["type"]]
    count = a["count"]
    if "bufferView" not in a:
        return np.zeros((count, nc), dt)
    v = doc["bufferViews"][a["bufferView"]]
    buf = buffers[v["buffer"]]
    start = v.get("byteOffset", 0) + a.get("byteOffset", 0)
    elem = dt.itemsize * nc
    stride = v.get("byteStride", 0) or elem
    if count and start + stride * (count - 1) + elem > v.get("byteOffset", 0) + v["byteLength"]:
        raise ValueError(f"accessor {index} runs past its buffer view")
    raw = np.frombuffer(buf, np.uint8, count=stride * (count - 1) + elem if count else 0, offset=start)
    if stride == elem:
        return raw.view(dt).reshape(count, nc).copy()
    rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1))
    return np.ascontiguousarray(rows).view(dt).reshape(count, nc)


def _quat_matrix(q) -> np.ndarray:
    x, y, z, w = (float(c) for c in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def local_matrix(node) -> np.ndarray:
    """4x4 (row-major, column vectors) of a node: `matrix` (stored column-major) or T * R * S."""
    if "matrix" in node:
        return np.array(node["matrix"], np.float64).reshape(4, 4).T
    m = np.eye(4)
    s = np.array(node.get("scale", [1.0, 1.0, 1.0]), np.float64)
    m[:3, :3] = _quat_matrix(node.get("rotation", [0.0, 0.0, 0.0, 1.0])) * s[None, :]
    m[:3, 3] = node.get("translation", [0.0, 0.0, 0.0])
    return m


def world_matrices(doc):
    """[(node index, world matrix)] for every node reachable from the scene's roots, depth first."""
    scenes = doc.get("scenes")
    if scenes:
        roots = scenes[doc.get("scene", 0)]["nodes"]
    else:   # no scene: every node that is nobody's child is a root
        kids = {c for n in doc.get("nodes", []) for c in n.get("children", [])}
        roots = [i for i in range(len(doc.get("nodes", []))) if i not in kids]
    out, stack = [], [(r, np.eye(4)) for r in reversed(roots)]
    while stack:
        i, parent = stack.pop()
        w = parent @ local_matrix(doc["nodes"][i])
        out.append((i, w))
        for c in reversed(doc["nodes"][i].get("children", [])):
            stack.append((c, w))
    return out


def _triangle_indices(mode: int, idx: np.ndarray) -> np.ndarray:
    if mode == 4:
        n = idx.size // 3 * 3
        return idx[:n].reshape(-1, 3)
    if mode == 5:   # strip: (i, i+1, i+2), odd triangles flipped to keep the winding
        k = np.arange(max(idx.size - 2, 0))
        t = np.stack([idx[k], idx[k + 1], idx[k + 2]], 1)
        t[1::2, [0, 1]] = t[1::2, [1, 0]]
        return t
    if mode == 6:   # fan around the first index
        k = np.arange(1, max(idx.size - 1, 1))
        return np.stack([np.full(k.size, idx[0]), idx[k], idx[k + 1]], 1) if idx.size >= 3 else np.zeros((0, 3))
    return np.zeros((0, 3), np.int64)   # points and lines carry no triangles


def load(path: str):
    """World-space triangles of every mesh node: (verts (n, 9) float32,
    normals (n, 9) float32), in node order, then primitive order, then
    index order."""
    doc, buffers = read_document(path)
    verts, norms = [], []
    for node_i, w in world_matrices(doc):
        node = doc["nodes"][node_i]
        if "mesh" not in node:
            continue
        nmat = np.linalg.inv(w[:3, :3]).T
        for prim in doc["meshes"][node["mesh"]]["primitives"]:
            attrs = prim["attributes"]
            if "POSITION" not in attrs:
                continue
            if doc["accessors"][attrs["POSITION"]]["componentType"] != 5126:
                raise ValueError(f"{path}: quantised positions are not supported")
            pos = accessor(doc, buffers, attrs["POSITION"]).astype(np.float64)
            if "indices" in prim:
                idx = accessor(doc, buffers, prim["indices"]).reshape(-1).astype(np.int64)
            else:
                idx = np.arange(pos.shape[0], dtype=np.int64)
            tri = _triangle_indices(prim.get("mode", 4), idx)
            if tri.size == 0:
                continue
            p = pos @ w[:3, :3].T + w[:3, 3]
            v = p[tri].reshape(-1, 9)
            if "NORMAL" in attrs:
                nrm = accessor(doc, buffers, attrs["NORMAL"]).astype(np.float64) @ nmat.T
                nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
                nv = nrm[tri].reshape(-1, 9)
            else:   # face normal on all three vertices
                a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
                fn = np.cross(b - a, c - a)
                fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
                nv = np.tile(fn, (1, 3))
            verts.append(v.astype(np.float32))
            norms.append(nv.astype(np.float32))
    if not verts:
        raise ValueError(f"{path}: no triangles")
    return np.concatenate(verts), np.concatenate(norms)


def _rotate(p, deg, axis):
    """One of cargarModelo.h:21-66's rotations: angle in FP32 radians, FP32
    cosine and sine, the rotated coordinates rounded to FP32."""
    ang = np.float32(np.float64(np.float32(deg)) * np.pi / np.float64(np.float32(180.0)))
    c = np.float64(np.float32(np.cos(np.float64(ang))))
    s = np.float64(np.float32(np.sin(np.float64(ang))))
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    if axis == 0:     # [1 0 0; 0 c -s; 0 s c]
        out = (x, y * c - z * s, y * s + z * c)
    elif axis == 1:   # [c 0 s; 0 1 0; -s 0 c]
        out = (x * c + z * s, y, -x * s + z * c)
    else:             # [c -s 0; s c 0; 0 0 1]
        out = (x * c - y * s, x * s + y * c, z)
    return np.stack([np.asarray(o, np.float64).astype(np.float32).astype(np.float64) for o in out], 1)


def place(verts, normals, scale, rot_x, rot_y, rot_z, translation):
    """cargarModelo.h:146-183: scale, rotate about X then Y then Z (degrees),
    translate; normals rotated and normalised.  Returns float32 (n, 9) arrays."""
    p = np.asarray(verts, np.float32).reshape(-1, 3).astype(np.float64) * np.float64(np.float32(scale))
    for deg, axis in ((rot_x, 0), (rot_y, 1), (rot_z, 2)):
        p = _rotate(p, deg, axis)
    p = p + np.array([np.float64(np.float32(t)) for t in translation])
    n = np.asarray(normals, np.float32).reshape(-1, 3).astype(np.float64)
    for deg, axis in ((rot_x, 0), (rot_y, 1), (rot_z, 2)):
        n = _rotate(n, deg, axis)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return p.astype(np.float32).reshape(-1, 9), n.astype(np.float32).reshape(-1, 9)
