"""tfg_nrc_infer_datos_noalloc (the allocation-free variant of
tfg_nrc_infer_datos, scratch in a static device array) against
tfg_nrc_infer_datos on the same DatosMLP queries: bit-identical colours on
two calls in a row, the second with fewer queries, and a clear error for a
call beyond the scratch capacity."""
import numpy as np
import pytest

from tfg_amd import _lib, nrc

pytestmark = pytest.mark.gpu

MAX_QUERIES = 1 << 16
DATOS = np.dtype([("pos", np.float64, 3), ("dir", np.float64, 3), ("nrm", np.float64, 3), ("kd", np.float32, 3),
                  ("ks", np.float32, 3), ("color", np.float32, 3)], align=True)
assert DATOS.itemsize == 112


def _datos(n, seed):
    rng = np.random.default_rng(seed)
    d = np.zeros(n, DATOS)
    d["pos"] = rng.uniform((-6, -4, -10), (6, 6, 2), (n, 3))
    d["dir"] = rng.uniform(0, 1, (n, 3))
    d["nrm"] = rng.uniform(0, 1, (n, 3))
    d["kd"] = rng.uniform(0, 1, (n, 3))
    d["ks"] = rng.uniform(0, 0.3, (n, 3))
    return d


def test_noalloc_inference_matches_allocating_entry_point(torch_cuda):
    torch = torch_cuda
    L = _lib.lib()
    cache = nrc.RadianceCache(seed=21)
    stream = torch.cuda.current_stream().cuda_stream
    for k, n in enumerate((MAX_QUERIES, 1000)):
        print(f"call {k}: {n} queries", flush=True)
        d = torch.from_numpy(np.frombuffer(_datos(n, 30 + k).tobytes(), np.uint8).copy()).cuda()
        ref = torch.full((n, 3), -1.0, device="cuda")
        out = torch.full((n, 3), -2.0, device="cuda")
        _lib.check(L.tfg_nrc_infer_datos(cache.handle, d.data_ptr(), n, ref.data_ptr(), stream), "infer_datos")
        _lib.check(L.tfg_nrc_infer_datos_noalloc(cache.handle, d.data_ptr(), n, out.data_ptr(), stream),
                   "infer_datos_noalloc")
        torch.cuda.synchronize()
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert (r >= 0).all() and np.isfinite(r).all()
        np.testing.assert_array_equal(o.view(np.uint32), r.view(np.uint32))
    big = torch.zeros((MAX_QUERIES + 1) * 112, dtype=torch.uint8, device="cuda")
    res = torch.zeros((MAX_QUERIES + 1, 3), device="cuda")
    rc = L.tfg_nrc_infer_datos_noalloc(cache.handle, big.data_ptr(), MAX_QUERIES + 1, res.data_ptr(), stream)
    assert rc == -4   # TFG_E_LIMIT
    assert b"scratch" in L.tfg_last_error()
    cache.close()
