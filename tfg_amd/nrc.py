"""Host side of the radiance cache (tfg_nrc_* in include/tfg.h).

``RadianceCache`` owns the device network (FP32 master weights, Adam and EMA
state, 16-bit matrix-core operands).  Attached to a ``PathTracer`` created
with a cache mode, its inference and training run inside the tracer's frame.
``save`` / ``load`` replace ColorMLP::save_model / load_model
(mlp.cpp:173-276) with a small versioned binary file.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _lib
from ._lib import check, lib

_MAGIC = b"TFGNRC01"


def default_config(**over) -> _lib.NrcConfig:
    cfg = _lib.NrcConfig()
    check(lib().tfg_nrc_default_config(C.byref(cfg)))
    for k, v in over.items():
        if k in ("bounds_min", "bounds_max"):
            getattr(cfg, k)[:] = list(v)
        else:
            setattr(cfg, k, v)
    return cfg


class RadianceCache:
    def __init__(self, cfg: _lib.NrcConfig | None = None, **over):
        self.cfg = cfg if cfg is not None else default_config(**over)
        self._h = C.c_void_p()
        check(lib().tfg_nrc_create(C.byref(self.cfg), C.byref(self._h)), "tfg_nrc_create")

    @property
    def handle(self):
        return self._h

    def weights(self):
        m = np.zeros(_lib.TFG_NRC_PARAMS, np.float32)
        e = np.zeros(_lib.TFG_NRC_PARAMS, np.float32)
        check(lib().tfg_nrc_get_weights(self._h, m.ctypes.data, e.ctypes.data), "tfg_nrc_get_weights")
        return m, e

    def set_weights(self, master, ema=None):
        m = np.ascontiguousarray(master, np.float32)
        e = np.ascontiguousarray(ema, np.float32) if ema is not None else None
        assert m.size == _lib.TFG_NRC_PARAMS
        check(lib().tfg_nrc_set_weights(self._h, m.ctypes.data, e.ctypes.data if e is not None else None),
              "tfg_nrc_set_weights")

    def stats(self) -> dict:
        st = _lib.NrcStats()
        check(lib().tfg_nrc_get_stats(self._h, C.byref(st)), "tfg_nrc_get_stats")
        return {"steps": st.steps, "skipped_nonfinite": st.skipped_nonfinite, "skipped_no_data": st.skipped_no_data,
                "seed_counter": st.seed_counter, "last_loss": st.last_loss}

    def eval(self, verts_dev, n, out_dev, dump_dev=None, train_weights=False, stream=None):
        check(lib().tfg_nrc_eval(self._h, verts_dev.data_ptr(), n, out_dev.data_ptr(),
                                 dump_dev.data_ptr() if dump_dev is not None else None, 1 if train_weights else 0,
                                 stream), "tfg_nrc_eval")

    def infer_queries(self, queries_ptr, count_ptr, image_ptr, stream=None):
        """Cache inference + composite over raw device pointers (the tracer's
        query buffer and counter, an image)."""
        check(lib().tfg_nrc_infer_queries(self._h, queries_ptr, count_ptr, image_ptr, stream),
              "tfg_nrc_infer_queries")

    def train_ptr(self, records_ptr, count_ptr, steps=1, stream=None):
        check(lib().tfg_nrc_train(self._h, records_ptr, count_ptr, steps, stream), "tfg_nrc_train")

    def train(self, records_dev, n_records_dev, steps=1, stream=None):
        check(lib().tfg_nrc_train(self._h, records_dev.data_ptr(), n_records_dev.data_ptr(), steps, stream),
              "tfg_nrc_train")

    def last_gradient(self):
        g = np.zeros(_lib.TFG_NRC_PARAMS, np.float32)
        check(lib().tfg_nrc_last_gradient(self._h, g.ctypes.data), "tfg_nrc_last_gradient")
        return g

    def save(self, path):
        m, e = self.weights()
        with open(path, "wb") as f:
            f.write(_MAGIC)
            f.write(struct.pack("<II", _lib.TFG_NRC_PARAMS, int(self.cfg.precision)))
            f.write(m.tobytes())
            f.write(e.tobytes())

    def load(self, path):
        with open(path, "rb") as f:
            if f.read(8) != _MAGIC:
                raise ValueError(f"{path}: not a radiance-cache weight file")
            n, _prec = struct.unpack("<II", f.read(8))
            if n != _lib.TFG_NRC_PARAMS:
                raise ValueError(f"{path}: {n} parameters, expected {_lib.TFG_NRC_PARAMS}")
            m = np.frombuffer(f.read(4 * n), np.float32)
            e = np.frombuffer(f.read(4 * n), np.float32)
        self.set_weights(m, e)

    def close(self):
        if self._h:
            lib().tfg_nrc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
