"""The reference-layout launchers (RenderKernel.h / NrcKernel.h counterparts)
against NumPy restatements of RenderKernel.cu / NrcKernel.cu on the same
inputs, special values included.  Bar: bit-exact for integer work (indices),
FP32 results equal to the restatement evaluated in FP32 (within 1 ulp)."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as M
from tfg_amd import _lib

pytestmark = pytest.mark.gpu

V3 = (np.float64, 3)
COL = (np.float32, 3)
DATOS = np.dtype([("pos", V3), ("dir", V3), ("nrm", V3), ("kd", COL), ("ks", COL), ("color", COL)], align=True)
GEOM = np.dtype([("pos", V3), ("dir", V3), ("nrm", V3), ("kd", COL), ("ks", COL)], align=True)
REG = np.dtype([("head", GEOM), ("tail", GEOM), ("luz", COL), ("thr", COL), ("norm", COL), ("valido", np.uint8)],
               align=True)
CAM = np.dtype([("pos", V3), ("front", V3), ("up", V3), ("left", V3), ("w", np.float32), ("h", np.float32),
                ("f", np.float32)], align=True)
assert DATOS.itemsize == 112 and REG.itemsize == 232 and CAM.itemsize == 112


def _dev(torch, a):
    return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()).cuda()


def _host(t, dtype):
    return np.frombuffer(t.cpu().numpy().tobytes(), dtype)


def _ok(rc):
    _lib.check(rc)


def test_camera(torch_cuda):
    torch = torch_cuda
    d = torch.zeros(112, dtype=torch.uint8, device="cuda")
    _ok(_lib.lib().tfg_launch_inicializar_camara(d.data_ptr(), 1920, 1080, None))
    c = _host(d, CAM)[0]
    np.testing.assert_array_equal(c["pos"], [0, -0.75, 4.25])
    np.testing.assert_array_equal(c["front"], [0, 0, -1])
    np.testing.assert_array_equal(c["up"], [0, 1, 0])
    np.testing.assert_array_equal(c["left"], [1, 0, 0])    # front x up (camara.h:36)
    h = np.float32(2 * np.tan(np.float32(np.float32(45 * np.pi / 180) / 2)))
    assert abs(c["h"] - h) <= 1e-6 and abs(c["w"] - h * np.float32(1920 / 1080)) <= 1e-6 and c["f"] == 1


def test_indices_bit_exact(torch_cuda):
    torch = torch_cuda
    idx = torch.zeros(16384, dtype=torch.int32, device="cuda")
    for seed in (0, 5, 4096):
        _ok(_lib.lib().tfg_launch_generar_indices(idx.data_ptr(), 65533, 16384, seed, None))
        np.testing.assert_array_equal(idx.cpu().numpy(), M.batch_indices(seed, 16384, 65533))


def test_targets_tail_and_composite(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(0)
    n = 1000
    reg = np.zeros(n, REG)
    for part in ("head", "tail"):
        for f in ("pos", "dir", "nrm"):
            reg[part][f] = rng.normal(size=(n, 3))
        reg[part]["kd"] = rng.uniform(size=(n, 3))
        reg[part]["ks"] = rng.uniform(size=(n, 3))
    reg["luz"] = rng.uniform(0, 3, (n, 3))
    reg["thr"] = rng.uniform(0, 1, (n, 3))
    reg["norm"] = rng.uniform(0, 1, (n, 3))
    reg["valido"] = rng.integers(0, 2, n)
    reg["luz"][:4] = [[np.inf, -np.inf, np.nan], [1e9, 0, 0], [0, 0, 0], [1, 1, 1]]
    reg["norm"][3] = 0.0
    pred = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    d_reg = _dev(torch, reg)
    d_pred = torch.from_numpy(pred).cuda()
    out = torch.zeros(n * 112, dtype=torch.uint8, device="cuda")
    _ok(_lib.lib().tfg_launch_calcular_targets(d_reg.data_ptr(), d_pred.data_ptr(), out.data_ptr(), n, None))
    o = _host(out, DATOS)
    v = reg["valido"] == 1
    recs = np.zeros((n, 11, 4), np.float32)
    recs[:, 8, :3], recs[:, 9, :3], recs[:, 10, :3] = reg["luz"], reg["thr"], reg["norm"]
    with np.errstate(all="ignore"):
        t32 = (reg["luz"] + reg["thr"] * pred) / np.fmax(reg["norm"], np.float32(1e-8))
        t32 = np.fmin(t32, np.float32(1e5))
        t32[np.isnan(t32) | np.isinf(t32)] = 0
    np.testing.assert_allclose(o["color"][v], t32[v], rtol=2e-7)
    np.testing.assert_array_equal(o["color"][~v], 0.0)
    np.testing.assert_array_equal(o["pos"][v], reg["head"]["pos"][v])
    np.testing.assert_allclose(M.targets(recs.reshape(n, 44)[v], pred[v]), t32[v], rtol=1e-6)
    # tail inputs (RenderKernel.cu:139-166)
    _ok(_lib.lib().tfg_launch_preparar_inferencia_tail(d_reg.data_ptr(), out.data_ptr(), n, None))
    o = _host(out, DATOS)
    np.testing.assert_array_equal(o["pos"][v], reg["tail"]["pos"][v])
    np.testing.assert_array_equal(o["pos"][~v], 0.0)
    # composite (RenderKernel.cu:49-66)
    img = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    thr = rng.uniform(-0.5, 1, (n, 3)).astype(np.float32)
    thr[:10] = 0
    d_img = torch.from_numpy(img.copy()).cuda()
    _ok(_lib.lib().tfg_launch_composite(d_img.data_ptr(), d_pred.data_ptr(), torch.from_numpy(thr).cuda().data_ptr(),
                                        n, 1, None))
    any_pos = (thr > 0).any(1)
    # the kernel may fuse the multiply-add (as nvcc does by default): the
    # reference is the fused result (exact product, one rounding), and the
    # bound is one rounding of the larger operand, since negative throughput
    # entries cancel
    ref = img.copy()
    prod = thr.astype(np.float64) * pred.astype(np.float64)
    ref[any_pos] = (img[any_pos].astype(np.float64) + prod[any_pos]).astype(np.float32)
    bound = np.where(any_pos[:, None], 1.2e-7 * (np.abs(img) + np.abs(prod)), 0.0)
    assert np.all(np.abs(d_img.cpu().numpy().astype(np.float64) - ref) <= bound)
    np.testing.assert_array_equal(d_img.cpu().numpy()[~any_pos], img[~any_pos])


def test_pack_and_save_output(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(1)
    n = 513
    d = np.zeros(n, DATOS)
    d["pos"] = rng.uniform(-8, 4, (n, 3))
    d["dir"] = rng.uniform(size=(n, 3))
    d["nrm"] = rng.uniform(size=(n, 3))
    d["kd"] = rng.uniform(size=(n, 3))
    d["ks"] = rng.uniform(size=(n, 3))
    d["color"] = rng.uniform(0, 2e5, (n, 3))
    b = _lib.SceneBounds((-6.0, -4.0, -10.0), (6.0, 6.0, 2.0))
    d_d = _dev(torch, d)
    idx = rng.integers(0, n, 128).astype(np.int32)
    inp = torch.zeros((128, 15), device="cuda")
    tgt = torch.zeros((128, 3), device="cuda")
    _ok(_lib.lib().tfg_launch_preparar_entrenamiento(d_d.data_ptr(), torch.from_numpy(idx).cuda().data_ptr(), 128,
                                                     15, 3, b, inp.data_ptr(), tgt.data_ptr(), None))
    mn = np.float32([-6, -4, -10])
    mx = np.float32([6, 6, 2])
    pos = d["pos"][idx].astype(np.float32)
    ref_pos = np.clip((pos - mn) / (mx - mn), 0, 1)
    got = inp.cpu().numpy()
    np.testing.assert_allclose(got[:, 0:3], ref_pos, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(got[:, 3:6], d["dir"][idx].astype(np.float32))
    np.testing.assert_array_equal(got[:, 9:12], d["kd"][idx])
    np.testing.assert_array_equal(tgt.cpu().numpy(), np.minimum(d["color"][idx], np.float32(1e5)))
    inp2 = torch.zeros((n, 15), device="cuda")
    _ok(_lib.lib().tfg_launch_preparar_inferencia(d_d.data_ptr(), n, 15, inp2.data_ptr(), b, None))
    np.testing.assert_array_equal(inp2.cpu().numpy()[idx], got)
    raw = np.array([[-1, np.nan, 2e5], [np.inf, -np.inf, 3.0]], np.float32)
    col = torch.zeros((2, 3), device="cuda")
    _ok(_lib.lib().tfg_launch_guardar_salida(torch.from_numpy(raw).cuda().data_ptr(), col.data_ptr(), 2, None))
    np.testing.assert_array_equal(col.cpu().numpy(), M.sanitize(raw).astype(np.float32))
