"""CPU oracle for the tfg_amd hot path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package; the product path never does.

``liboracle.so`` (oracle.cpp, built by ``make -C oracle``) restates the
reference's RNG, camera, CWBVH decode, BVH2 build and Moller-Trumbore test;
``mlp_oracle.py`` is the FP64 NumPy MLP.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TFG_ORACLE_LIB names another build of oracle.cpp (the planted-mistake
# tests load deliberately broken copies through it)
LIB_PATH = os.environ.get("TFG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")


class TfgRay(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


_lib = None

_SIGS = {
    "orc_cosine_local": (None, [C.c_double, C.c_double, C.c_void_p]),
    "orc_ggx_D": (C.c_double, [C.c_double, C.c_double]),
    "orc_ggx_G1": (C.c_double, [C.c_double, C.c_double]),
    "orc_ggx_eval": (None, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_ggx_vndf": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p]),
    "orc_analytic_t": (C.c_float, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_pcg32": (C.c_uint32, [C.POINTER(C.c_uint32)]),
    "orc_pcg32_float": (C.c_float, [C.POINTER(C.c_uint32)]),
    "orc_seed": (C.c_uint32, [C.c_int, C.c_int, C.c_int]),
    "orc_stream": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "orc_camera_default": (None, [C.c_int, C.c_int, C.c_void_p]),
    "orc_camera_make": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_void_p]),
    "orc_camera_ray": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p,
                              C.c_void_p]),
    "orc_cwbvh_trace": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                               C.c_int]),
    "orc_cwbvh_occluded": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                  C.c_void_p, C.c_int]),
    "orc_brute_trace": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
    "orc_tri_watertight": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "orc_bvh2_build": (C.c_void_p, [C.c_void_p, C.c_uint32]),
    "orc_bvh2_free": (None, [C.c_void_p]),
    "orc_bvh2_num_nodes": (C.c_uint32, [C.c_void_p]),
    "orc_bvh2_trace": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]),
    "orc_cwbvh_validate": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                     C.c_void_p]),
    "orc_render": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                          C.c_void_p, C.c_int]),
    "orc_render_rows": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_void_p, C.c_void_p, C.c_int]),
    "orc_render_nrc": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                              C.c_void_p, C.c_int]),
    "orc_render_pixel": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_void_p]),
    "orc_render_pixels_perturbed": (None, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_int, C.c_int,
                                           C.c_int, C.c_double, C.c_uint32, C.c_void_p, C.c_int]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.uint32)])


def rays_array(o: np.ndarray, d: np.ndarray, tmin=0.0, tmax=np.inf) -> np.ndarray:
    """(n, 8) float32 rays in tfg_ray layout."""
    n = o.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def cwbvh_trace(nodes, tris, fmt, rays, nthreads=os.cpu_count() or 1, counts=False):
    n = rays.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros((n, 2), np.uint32) if counts else None
    lib().orc_cwbvh_trace(_p(nodes), _p(tris), fmt, _p(rays), n, _p(hits), _p(cnt) if counts else None, nthreads)
    return (hits, cnt) if counts else hits


def cwbvh_occluded(nodes, tris, fmt, rays, nthreads=os.cpu_count() or 1, counts=False):
    n = rays.shape[0]
    occ = np.zeros(n, np.uint32)
    cnt = np.zeros((n, 2), np.uint32) if counts else None
    lib().orc_cwbvh_occluded(_p(nodes), _p(tris), fmt, _p(rays), n, _p(occ), _p(cnt) if counts else None, nthreads)
    return (occ, cnt) if counts else occ


def brute_trace(verts, rays):
    v = np.ascontiguousarray(verts, np.float32)
    n = rays.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    lib().orc_brute_trace(_p(v), v.shape[0], _p(rays), n, _p(hits))
    return hits


def validate(nodes, tris, fmt, verts):
    d = C.c_uint32()
    v = np.ascontiguousarray(verts, np.float32)
    rc = lib().orc_cwbvh_validate(_p(nodes), nodes.shape[0], _p(tris), tris.shape[0], fmt, _p(v), C.byref(d))
    return rc, d.value


PRIM_DTYPE = np.dtype([
    ("tipo", np.int32), ("emision", np.float32, 3), ("difuso", np.float32, 3), ("especular", np.float32, 3),
    ("transmision", np.float32, 3), ("ior", np.float32), ("g", np.float64, 18)])
assert PRIM_DTYPE.itemsize == 200


def primitivas_from_verts(verts, normals=None):
    """Primitiva records (primitiva.h:17-39) for a triangle list."""
    n = verts.shape[0]
    p = np.zeros(n, PRIM_DTYPE)
    p["tipo"] = 2
    p["ior"] = 1.0
    p["g"][:, 0:9] = verts.astype(np.float64)
    if normals is not None:
        p["g"][:, 9:18] = normals.astype(np.float64)
    return p


class Bvh2:
    """bvh.cpp median-split BVH2 traversed with primitiva.h Moller-Trumbore."""

    def __init__(self, verts):
        self.prims = primitivas_from_verts(np.asarray(verts, np.float32))
        self._h = lib().orc_bvh2_build(self.prims.ctypes.data, self.prims.shape[0])
        self.num_nodes = lib().orc_bvh2_num_nodes(self._h)

    def trace(self, rays, nthreads=os.cpu_count() or 1):
        n = rays.shape[0]
        hits = np.zeros(n, HIT_DTYPE)
        lib().orc_bvh2_trace(self._h, _p(rays), n, _p(hits), nthreads)
        return hits

    def __del__(self):
        try:
            lib().orc_bvh2_free(self._h)
        except Exception:
            pass


class OrcScene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tris", C.c_void_p), ("shading", C.c_void_p), ("materials", C.c_void_p),
                ("lights", C.c_void_p), ("n_lights", C.c_uint32), ("cam", C.c_void_p), ("analytic", C.c_void_p),
                ("n_analytic", C.c_uint32)]


class Renderer:
    """CPU restatement of the integrator (render.h) over a host CWBVH."""

    def __init__(self, scene, nodes, tris, width, height):
        self._keep = [np.ascontiguousarray(nodes, np.float32), np.ascontiguousarray(tris, np.float32),
                      np.ascontiguousarray(scene.shading_records()),
                      np.ascontiguousarray(scene.materials, np.float32),
                      np.ascontiguousarray(scene.lights, np.float32),
                      np.ascontiguousarray(scene.camera.frame(width, height), np.float32)]
        an = getattr(scene, "analytic", None)
        n_an = 0 if an is None else an.shape[0]
        self._keep.append(np.ascontiguousarray(an if n_an else np.zeros((1, 12), np.float32)))
        k = self._keep
        self.desc = OrcScene(_p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), scene.lights.shape[0], _p(k[5]),
                             _p(k[6]), n_an)
        self.width, self.height = width, height

    def render(self, frame0, spp, max_depth, rr_depth, nthreads=os.cpu_count() or 1):
        acc = np.zeros((self.height * self.width, 3), np.float64)
        rays = np.zeros(2, np.uint64)
        lib().orc_render(C.byref(self.desc), self.width, self.height, frame0, spp, max_depth, rr_depth, _p(acc),
                         _p(rays), nthreads)
        return acc, rays

    def render_rows(self, y0, y1, frame0, spp, max_depth, rr_depth, nthreads=os.cpu_count() or 1):
        acc = np.zeros(((y1 - y0) * self.width, 3), np.float64)
        rays = np.zeros(2, np.uint64)
        lib().orc_render_rows(C.byref(self.desc), self.width, self.height, y0, y1, frame0, spp, max_depth, rr_depth,
                              _p(acc), _p(rays), nthreads)
        return acc, rays

    def render_nrc(self, frame, max_depth, rr_depth, mode, train_frac=0.05, nthreads=os.cpu_count() or 1):
        """Per-pixel cache-mode outputs: dict of arrays (see orc_render_nrc)."""
        out = np.zeros((self.height * self.width, 64), np.float32)
        lib().orc_render_nrc(C.byref(self.desc), self.width, self.height, frame, max_depth, rr_depth, mode,
                             train_frac, _p(out), nthreads)
        return {"pixel": out[:, 0:3], "flags": out[:, 3].view(np.uint32), "head": out[:, 4:20],
                "tail": out[:, 20:36], "lsuf": out[:, 36:39], "tsuf": out[:, 39:42], "norm": out[:, 42:45]}

    def pixels_perturbed(self, xy, frame, max_depth, rr_depth, delta, seed, nthreads=os.cpu_count() or 1):
        """Paths of pixels xy ((n, 2) x, y) with every ray displaced by `delta`
        (relative) before its closest-hit query: a conditioning probe."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        out = np.zeros((xy.shape[0], 3), np.float32)
        if xy.shape[0]:
            lib().orc_render_pixels_perturbed(C.byref(self.desc), self.width, self.height, _p(xy), xy.shape[0], frame,
                                              max_depth, rr_depth, delta, seed, _p(out), nthreads)
        return out

    def pixel(self, x, y, frame, max_depth, rr_depth):
        out = np.zeros(3, np.float32)
        lib().orc_render_pixel(C.byref(self.desc), self.width, self.height, x, y, frame, max_depth, rr_depth,
                               _p(out))
        return out
