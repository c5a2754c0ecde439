"""Host interface of the wavefront path tracer (tfg_pt_* in include/tfg.h).

``DeviceScene`` uploads a :class:`tfg_amd.scenes.Scene` (CWBVH, shading
records, materials, lights) to device memory; ``PathTracer`` renders frames
into caller-owned device images.  Device memory is held in torch tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


class DeviceScene:
    """Scene buffers in device memory; build_options go to the CWBVH builder."""

    def __init__(self, scene, torch, **build_options):
        self.scene = scene
        self.host = _lib.CwbvhHost(scene.verts, _lib.TFG_TRI_WATERTIGHT, **build_options)
        self.bvh = self.host.upload()
        self.shading = torch.from_numpy(np.ascontiguousarray(scene.shading_records()).view(np.int32)).cuda()
        self.materials = torch.from_numpy(np.ascontiguousarray(scene.materials, np.float32)).cuda()
        self.lights = torch.from_numpy(np.ascontiguousarray(scene.lights, np.float32)).cuda()
        self.desc = _lib.PtScene()
        self.desc.bvh = self.bvh.desc
        self.desc.d_shading = self.shading.data_ptr()
        self.desc.d_materials = self.materials.data_ptr()
        self.desc.d_lights = self.lights.data_ptr()
        self.desc.n_lights = scene.lights.shape[0]
        an = getattr(scene, "analytic", None)
        if an is not None and an.shape[0]:
            self.analytic = torch.from_numpy(np.ascontiguousarray(an).view(np.int32)).cuda()
            self.desc.d_analytic = self.analytic.data_ptr()
            self.desc.n_analytic = an.shape[0]


# tfg_pt_view: width, height, tfg_camera (16 floats)
VIEW_DTYPE = np.dtype([("width", np.uint32), ("height", np.uint32), ("camera", np.float32, 16)])
assert VIEW_DTYPE.itemsize == 72


class PathTracer:
    def __init__(self, width, height, max_depth, rr_depth, mode=0, train_fraction=0.05, use_graph=True):
        self.cfg = _lib.PtConfig(width, height, max_depth, rr_depth, mode, train_fraction, 1 if use_graph else 0)
        self._h = C.c_void_p()
        check(lib().tfg_pt_create(C.byref(self.cfg), C.byref(self._h)), "tfg_pt_create")
        self.width, self.height = width, height

    def set_scene(self, dscene: DeviceScene):
        self._scene = dscene
        check(lib().tfg_pt_set_scene(self._h, C.byref(dscene.desc)), "tfg_pt_set_scene")

    def set_camera(self, cam_frame: np.ndarray):
        c = np.ascontiguousarray(cam_frame, np.float32)
        assert c.size == 16
        check(lib().tfg_pt_set_camera(self._h, c.ctypes.data), "tfg_pt_set_camera")

    def set_views(self, views):
        """views: (camera frame (16 floats, tfg_camera), width, height) per view;
        their pixels, one view after the other, fill the tracer's canvas.
        Returns the canvas offset of each view."""
        rec = np.zeros(len(views), VIEW_DTYPE)
        bases, b = [], 0
        for k, (cam_frame, w, h) in enumerate(views):
            c = np.ascontiguousarray(cam_frame, np.float32)
            assert c.size == 16
            rec[k]["width"], rec[k]["height"], rec[k]["camera"] = w, h, c
            bases.append(b)
            b += w * h
        check(lib().tfg_pt_set_views(self._h, rec.ctypes.data if len(views) else None, len(views)),
              "tfg_pt_set_views")
        return bases

    def set_cache(self, cache):
        self._cache = cache
        check(lib().tfg_pt_set_nrc(self._h, cache.handle if cache is not None else None), "tfg_pt_set_nrc")

    def set_mode(self, mode: int):
        check(lib().tfg_pt_set_mode(self._h, mode), "tfg_pt_set_mode")

    def flush_training(self, stream=None):
        """With schedule defer_training: enqueue the last training frame's
        optimiser steps now (they otherwise run at the next render)."""
        check(lib().tfg_pt_flush_training(self._h, stream), "tfg_pt_flush_training")

    def cache_buffers(self):
        r, q, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().tfg_pt_nrc_buffers(self._h, C.byref(r), C.byref(q), C.byref(c)), "tfg_pt_nrc_buffers")
        return r.value, q.value, c.value

    def read_cache_data(self):
        """(records (n, 44), queries (m, 20)) of the last frame, float32."""
        n = self.width * self.height
        rec = np.zeros((n, 44), np.float32)
        q = np.zeros((n, 20), np.float32)
        nr, nq = C.c_uint32(), C.c_uint32()
        check(lib().tfg_pt_read_cache_data(self._h, rec.ctypes.data, n, q.ctypes.data, n, C.byref(nr), C.byref(nq)),
              "tfg_pt_read_cache_data")
        return rec[:nr.value], q[:nq.value]

    def render(self, frame: int, image, accum=None, stream=None):
        check(lib().tfg_pt_render(self._h, int(frame), image.data_ptr(),
                                  accum.data_ptr() if accum is not None else None, stream), "tfg_pt_render")

    def counters(self) -> np.ndarray:
        out = np.zeros(_lib.TFG_PT_NCOUNTERS, np.uint32)
        check(lib().tfg_pt_counters(self._h, out.ctypes.data), "tfg_pt_counters")
        return out

    def rays_last_frame(self):
        c = self.counters()
        return int(c[8:40].sum()), int(c[40:72].sum())

    def ray_totals(self):
        out = np.zeros(2, np.uint64)
        check(lib().tfg_pt_ray_totals(self._h, out.ctypes.data), "tfg_pt_ray_totals")
        return int(out[0]), int(out[1])

    def profile_frame(self, frame: int, image) -> dict:
        pr = _lib.PtProfile()
        check(lib().tfg_pt_profile_frame(self._h, int(frame), image.data_ptr(), C.byref(pr)), "tfg_pt_profile_frame")
        d = self.cfg.max_depth
        return {
            "ms_generate": pr.ms_generate, "ms_finalize": pr.ms_finalize, "ms_total": pr.ms_total,
            "ms_extend": list(pr.ms_extend[:d]), "ms_shade": list(pr.ms_shade[:d]),
            "ms_connect": list(pr.ms_connect[:d]), "ext_rays": list(pr.ext_rays[:d]),
            "shadow_rays": list(pr.shadow_rays[:d]), "ext_nodes": pr.ext_nodes, "ext_tris": pr.ext_tris,
            "shadow_nodes": pr.shadow_nodes, "shadow_tris": pr.shadow_tris,
            "at_extend": list(pr.at_extend[:d]), "at_shade": list(pr.at_shade[:d]),
            "at_connect": list(pr.at_connect[:d]), "at_finalize": pr.at_finalize, "ms_tail": pr.ms_tail,
            "at_tail": pr.at_tail,
        }

    def reset_totals(self):
        check(lib().tfg_pt_reset_totals(self._h), "tfg_pt_reset_totals")

    def autotune(self, stream=None):
        check(lib().tfg_pt_autotune(self._h, stream), "tfg_pt_autotune")
        return self.launch_shape()

    def schedule(self) -> dict:
        sc = _lib.PtSchedule()
        check(lib().tfg_pt_get_schedule(self._h, C.byref(sc)), "tfg_pt_get_schedule")
        return {"trace_block": sc.trace_block, "grid_mult": sc.grid_mult, "tiles": sc.tiles,
                "side_streams": sc.side_streams, "overlap": bool(sc.overlap), "use_graph": bool(sc.use_graph),
                "xcd_group": sc.xcd_group, "octet_below": sc.octet_below, "connect_per_cu": sc.connect_per_cu,
                "connect_batch": sc.connect_batch, "top_nodes": sc.top_nodes, "tail_from": sc.tail_from,
                "defer_training": bool(sc.defer_training), "shadow_sort_from": sc.shadow_sort_from,
                "ray_sort_from": sc.ray_sort_from, "refill_from": sc.refill_from, "refill_idle": sc.refill_idle,
                "refill_per_cu": sc.refill_per_cu}

    def set_schedule(self, **kw):
        """Change any of trace_block, grid_mult, tiles, side_streams, overlap, xcd_group,
        octet_below, connect_per_cu, connect_batch, top_nodes, tail_from, defer_training, shadow_sort_from,
        ray_sort_from, refill_from, refill_idle, refill_per_cu (use_graph is fixed when the tracer is
        created)."""
        cur = self.schedule()
        unknown = set(kw) - set(cur)
        if unknown:
            raise TypeError(f"unknown schedule fields {sorted(unknown)}")
        cur.update(kw)
        sc = _lib.PtSchedule(cur["trace_block"], cur["grid_mult"], cur["tiles"], cur["side_streams"],
                             1 if cur["overlap"] else 0, 1 if cur["use_graph"] else 0, cur["xcd_group"],
                             cur["octet_below"], cur["connect_per_cu"], cur["connect_batch"], cur["top_nodes"],
                             cur["tail_from"], 1 if cur["defer_training"] else 0, cur["shadow_sort_from"],
                             cur["ray_sort_from"], cur["refill_from"], cur["refill_idle"], cur["refill_per_cu"])
        check(lib().tfg_pt_set_schedule(self._h, C.byref(sc)), "tfg_pt_set_schedule")

    def launch_shape(self):
        b, g, s = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().tfg_pt_get_launch(self._h, C.byref(b), C.byref(g), C.byref(s)))
        out = {"trace_grid": g.value, "shade_grid": s.value}
        out.update(self.schedule())
        return out

    def set_overlap(self, enable: bool):
        self.set_schedule(overlap=bool(enable))

    def set_tiles(self, tiles: int):
        self.set_schedule(tiles=int(tiles))

    def set_sides(self, sides: int):
        self.set_schedule(side_streams=int(sides))

    def set_launch(self, trace_block: int, grid_mult: int):
        self.set_schedule(trace_block=int(trace_block), grid_mult=int(grid_mult))

    def close(self):
        if self._h:
            lib().tfg_pt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
