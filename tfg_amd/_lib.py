"""ctypes binding of libtfg_hip.so (the C ABI declared in include/tfg.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C tfg_amd/csrc``).  There is no fallback: if the library is missing,
importing anything that needs it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtfg_hip.so")

TFG_TRI_WATERTIGHT = 0
TFG_TRI_AFFINE = 1
TFG_MAX_TREE_DEPTH = 25


class TfgError(RuntimeError):
    pass


class CwbvhGpu(C.Structure):
    """tfg_cwbvh_gpu == TinyBVHD_GPU (tinybvh_wrapper.h:15-21)."""

    _fields_ = [("d_nodos", C.c_void_p), ("d_tris", C.c_void_p), ("num_nodos", C.c_uint32), ("num_tris", C.c_uint32)]


class ImagenGpu(C.Structure):
    """ImagenGPU (imagen_gpu.h:14-17), passed by value."""

    _fields_ = [("anchura", C.c_int32), ("altura", C.c_int32), ("datos", C.c_void_p)]


class SceneBounds(C.Structure):
    """SceneBounds (mlp_types.h:26-29), passed by value."""

    _fields_ = [("min", C.c_double * 3), ("max", C.c_double * 3)]


class PtConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("mode", C.c_uint32), ("train_fraction", C.c_float), ("use_graph", C.c_uint32)]


class PtScene(C.Structure):
    _fields_ = [("bvh", CwbvhGpu), ("d_shading", C.c_void_p), ("d_materials", C.c_void_p), ("d_lights", C.c_void_p),
                ("n_lights", C.c_uint32), ("d_analytic", C.c_void_p), ("n_analytic", C.c_uint32),
                ("tri_format", C.c_uint32), ("d_verts", C.c_void_p), ("n_verts", C.c_uint32)]


class PtProfile(C.Structure):
    _fields_ = [("ms_generate", C.c_float), ("ms_finalize", C.c_float), ("ms_total", C.c_float),
                ("ms_extend", C.c_float * 32), ("ms_shade", C.c_float * 32), ("ms_connect", C.c_float * 32),
                ("ext_rays", C.c_uint32 * 32), ("shadow_rays", C.c_uint32 * 32),
                ("ext_nodes", C.c_uint64), ("ext_tris", C.c_uint64), ("shadow_nodes", C.c_uint64),
                ("shadow_tris", C.c_uint64), ("at_extend", C.c_float * 32), ("at_shade", C.c_float * 32),
                ("at_connect", C.c_float * 32), ("at_finalize", C.c_float), ("ms_tail", C.c_float),
                ("at_tail", C.c_float)]


class PtSchedule(C.Structure):
    _fields_ = [("trace_block", C.c_uint32), ("grid_mult", C.c_uint32), ("tiles", C.c_uint32),
                ("side_streams", C.c_uint32), ("overlap", C.c_int32), ("use_graph", C.c_int32),
                ("xcd_group", C.c_uint32), ("octet_below", C.c_uint32), ("connect_per_cu", C.c_uint32),
                ("connect_batch", C.c_uint32), ("top_nodes", C.c_uint32), ("tail_from", C.c_uint32),
                ("defer_training", C.c_uint32), ("shadow_sort_from", C.c_uint32),
                ("ray_sort_from", C.c_uint32), ("refill_from", C.c_uint32), ("refill_idle", C.c_uint32),
                ("refill_per_cu", C.c_uint32)]


class NrcConfig(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("steps_per_frame", C.c_uint32), ("min_records", C.c_uint32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("ema_decay", C.c_float), ("loss_scale", C.c_float), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3), ("seed", C.c_uint32), ("precision", C.c_uint32)]


class NrcStats(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("skipped_nonfinite", C.c_uint32), ("skipped_no_data", C.c_uint32),
                ("seed_counter", C.c_uint32), ("last_loss", C.c_float)]


TFG_NRC_BF16 = 0
TFG_NRC_FP16 = 1
TFG_NRC_PARAMS = 21504

TFG_PT_INFER = 1
TFG_PT_TRAIN = 2
TFG_PT_NCOUNTERS = 72

_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "tfg_last_error": (C.c_char_p, []),
    "tfg_version": (C.c_int, []),
    "tfg_cwbvh_build": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tfg_cwbvh_build_ex": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_void_p)]),
    "tfg_cwbvh_default_build_options": (None, [C.c_void_p]),
    "tfg_cwbvh_host_info": (
        C.c_int,
        [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_double)],
    ),
    "tfg_cwbvh_host_nodes": (C.c_void_p, [C.c_void_p]),
    "tfg_cwbvh_host_tris": (C.c_void_p, [C.c_void_p]),
    "tfg_cwbvh_host_free": (None, [C.c_void_p]),
    "tfg_cwbvh_upload": (C.c_int, [C.c_void_p, C.POINTER(CwbvhGpu)]),
    "tfg_cwbvh_gpu_free": (C.c_int, [C.POINTER(CwbvhGpu)]),
    "tfg_cwbvh_gpu_download": (C.c_int, [C.POINTER(CwbvhGpu), C.c_void_p, C.c_void_p]),
    "tfg_cwbvh_refit_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "tfg_cwbvh_refit": (C.c_int, [C.c_void_p, C.POINTER(CwbvhGpu), C.c_void_p, C.c_uint32, C.c_void_p]),
    "tfg_cwbvh_refit_destroy": (C.c_int, [C.c_void_p]),
    "tfg_trace_closest": (
        C.c_int,
        [C.POINTER(CwbvhGpu), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    ),
    "tfg_trace_any": (
        C.c_int,
        [C.POINTER(CwbvhGpu), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    ),
    "tfg_pt_create": (C.c_int, [C.POINTER(PtConfig), C.POINTER(C.c_void_p)]),
    "tfg_pt_set_scene": (C.c_int, [C.c_void_p, C.POINTER(PtScene)]),
    "tfg_pt_set_camera": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_set_views": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "tfg_pt_render": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tfg_pt_counters": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_autotune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_get_schedule": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_set_schedule": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_ray_totals": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_profile_frame": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(PtProfile)]),
    "tfg_pt_reset_totals": (C.c_int, [C.c_void_p]),
    "tfg_pt_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]),
    "tfg_pt_destroy": (C.c_int, [C.c_void_p]),
    "tfg_pt_set_nrc": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_read_cache_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "tfg_pt_set_mode": (C.c_int, [C.c_void_p, C.c_uint32]),
    "tfg_pt_flush_training": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_pt_nrc_buffers": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p)]),
    "tfg_nrc_default_config": (C.c_int, [C.POINTER(NrcConfig)]),
    "tfg_nrc_create": (C.c_int, [C.POINTER(NrcConfig), C.POINTER(C.c_void_p)]),
    "tfg_nrc_destroy": (C.c_int, [C.c_void_p]),
    "tfg_nrc_set_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tfg_nrc_get_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tfg_nrc_get_stats": (C.c_int, [C.c_void_p, C.POINTER(NrcStats)]),
    "tfg_nrc_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "tfg_image_resolve": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "tfg_nrc_infer_queries": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tfg_nrc_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "tfg_nrc_last_gradient": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tfg_nrc_infer_datos": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "tfg_nrc_train_datos": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_void_p]),
    "tfg_nrc_infer_datos_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "tfg_launch_inicializar_camara": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "tfg_launch_composite": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "tfg_launch_calcular_targets": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "tfg_launch_preparar_inferencia_tail": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "tfg_launch_generar_indices": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_void_p]),
    "tfg_launch_preparar_entrenamiento": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                                    SceneBounds, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tfg_launch_preparar_inferencia": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, SceneBounds,
                                                 C.c_void_p]),
    "tfg_launch_guardar_salida": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "tfg_launch_render_tiny": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_int, CwbvhGpu, C.c_void_p, C.c_int,
                                         ImagenGpu, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "tfg_dropin_configure": (C.c_int, [C.c_uint32, C.c_uint32]),
    "tfg_dropin_invalidate": (C.c_int, []),
}


def lib():
    """Loads libtfg_hip.so (raises TfgError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TfgError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C tfg_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def declared_symbols():
    return list(_SIGS)


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().tfg_last_error().decode(errors="replace")
        raise TfgError(f"{what or 'tfg call'} failed ({rc}): {msg}")


def ptr(x) -> int:
    """Address of a numpy array or a torch tensor."""
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return x.data_ptr()


class BuildOptions(C.Structure):
    _fields_ = [("node_cost", C.c_float), ("tri_cost", C.c_float), ("max_leaf", C.c_uint32), ("bins", C.c_uint32)]


class CwbvhHost:
    """Host CWBVH built by tfg_cwbvh_build; nodes/tris exposed as numpy views.
    Builder settings (node_cost, tri_cost, max_leaf, bins) override the defaults."""

    def __init__(self, verts: np.ndarray, fmt: int = TFG_TRI_WATERTIGHT, **options):
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
        self._h = C.c_void_p()
        opts = BuildOptions()
        lib().tfg_cwbvh_default_build_options(C.byref(opts))
        for k, val in options.items():
            setattr(opts, k, val)
        check(lib().tfg_cwbvh_build_ex(v.ctypes.data, v.shape[0], fmt, C.byref(opts), C.byref(self._h)),
              "tfg_cwbvh_build")
        self.format = fmt
        nn, nt, dp = C.c_uint32(), C.c_uint32(), C.c_uint32()
        sah = C.c_double()
        check(lib().tfg_cwbvh_host_info(self._h, C.byref(nn), C.byref(nt), C.byref(dp), C.byref(sah)))
        self.num_nodes, self.num_tris, self.max_depth, self.sah_cost = nn.value, nt.value, dp.value, sah.value
        n_addr = lib().tfg_cwbvh_host_nodes(self._h)
        t_addr = lib().tfg_cwbvh_host_tris(self._h)
        self.nodes = np.ctypeslib.as_array((C.c_float * (20 * self.num_nodes)).from_address(n_addr)).reshape(-1, 20)
        self.tris = np.ctypeslib.as_array((C.c_float * (16 * self.num_tris)).from_address(t_addr)).reshape(-1, 16)

    def upload(self) -> "DeviceBvh":
        return DeviceBvh(self)

    def close(self):
        if self._h:
            lib().tfg_cwbvh_host_free(self._h)
            self._h = C.c_void_p()
            self.nodes = self.tris = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBvh:
    """Device copy of a CwbvhHost (tfg_cwbvh_upload)."""

    def __init__(self, host: CwbvhHost):
        self.desc = CwbvhGpu()
        self.format = host.format
        check(lib().tfg_cwbvh_upload(host._h, C.byref(self.desc)), "tfg_cwbvh_upload")

    def trace_closest(self, rays, hits, n: int, counts=None, stream=None):
        check(
            lib().tfg_trace_closest(
                C.byref(self.desc), self.format, ptr(rays), ptr(hits), ptr(counts) if counts is not None else None, n,
                stream,
            ),
            "tfg_trace_closest",
        )

    def trace_any(self, rays, occluded, n: int, counts=None, stream=None):
        check(
            lib().tfg_trace_any(
                C.byref(self.desc), self.format, ptr(rays), ptr(occluded), ptr(counts) if counts is not None else None,
                n, stream,
            ),
            "tfg_trace_any",
        )

    def download(self):
        """(nodes (n, 20), tris (m, 16)) float32 copies of the device tree."""
        nodes = np.zeros((self.desc.num_nodos, 20), np.float32)
        tris = np.zeros((self.desc.num_tris, 16), np.float32)
        check(lib().tfg_cwbvh_gpu_download(C.byref(self.desc), nodes.ctypes.data, tris.ctypes.data),
              "tfg_cwbvh_gpu_download")
        return nodes, tris

    def close(self):
        if self.desc.d_nodos:
            check(lib().tfg_cwbvh_gpu_free(C.byref(self.desc)), "tfg_cwbvh_gpu_free")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Refitter:
    """GPU refit of a device tree to moved vertices (tfg_cwbvh_refit): made
    once from the host tree the device tree was uploaded from."""

    def __init__(self, host: CwbvhHost):
        self._h = C.c_void_p()
        self.num_tris = host.num_tris
        check(lib().tfg_cwbvh_refit_create(host._h, C.byref(self._h)), "tfg_cwbvh_refit_create")

    def refit(self, dev: DeviceBvh, verts, stream=None):
        """verts: device tensor of 9 float32 per triangle, the build's input order."""
        assert verts.numel() == 9 * self.num_tris
        check(lib().tfg_cwbvh_refit(self._h, C.byref(dev.desc), ptr(verts), self.num_tris, stream), "tfg_cwbvh_refit")

    def close(self):
        if self._h:
            lib().tfg_cwbvh_refit_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
