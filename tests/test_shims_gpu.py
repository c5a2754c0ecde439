"""The reference-named C++ shims (include/tfg_ref) driven by a C++ caller
(tests/cpp/shim_run.cpp, built by tests/cpp/Makefile with hipcc and linked
against libtfg_hip.so) through the frame main.cpp:1016-1056 runs, every
output checked against the oracle:

  * plain frame (launchKernelRenderTiny, flags 0/0, the drop-in's default
    depth 20 and Russian roulette after 10): >= 95% of the paths agree with
    the oracle integrator within 1e-3 relative, and >= 90% of those that do
    not are ill-conditioned in the oracle (tests/conditioning.py);
  * cache frame (flags 1/1): record count within 5% of the oracle's
    training heads, queries (nonzero buffer_throughput) where the oracle
    makes them for >= 95% of the pixels;
  * ColorMLP::inference on buffer_inference, the tail inputs and after
    save_model / load_model: >= 99% of the predictions within 2e-2 relative
    of the FP64 oracle network (BF16 operand rounding) with the weights the
    model files hold;
  * launchKernelComposite: image = prefix + throughput x prediction where
    the throughput is nonzero (RenderKernel.cu:49-66), to 1e-6;
  * launchKernelPrepararInferenciaTail and launchKernelCalcularTargets on
    the records: the tail inputs and targets of RenderKernel.cu:92-166
    restated in NumPy, to 1e-6;
  * ColorMLP::train_step: a finite positive loss and changed weights.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import conditioning
import oracle
import test_dropin_gpu as D
from oracle import mlp_oracle as M
from tfg_amd import _lib, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "shim_run")
BMIN, BMAX = (-6.0, -4.0, -10.0), (6.0, 6.0, 2.0)


def _model(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"TFGNRC01"
        n, _ = struct.unpack("<II", f.read(8))
        m = np.frombuffer(f.read(4 * n), np.float32)
        e = np.frombuffer(f.read(4 * n), np.float32)
    return m, e


def _verts_of_datos(d):
    """DatosMLP -> the cache's vertex record (roughness 0), as ColorMLP::inference
    converts it before the network."""
    v = np.zeros((d.shape[0], 16), np.float32)
    v[:, 0:3] = d["pos"]
    v[:, 4:7] = d["dir"]
    v[:, 7] = d["kd"][:, 0]
    v[:, 8:11] = d["nrm"]
    v[:, 11] = d["kd"][:, 1]
    v[:, 12] = d["kd"][:, 2]
    v[:, 13:16] = d["ks"]
    return v


def _net(ema, d):
    x = M.encode(_verts_of_datos(d), BMIN, BMAX)
    return M.sanitize(M.forward(ema, x, emulate="bf16")[0][5][:, :3])


def _agree(got, expect, rtol, atol):
    return np.all(np.abs(got - expect) <= atol + rtol * np.abs(expect), axis=1)


def test_reference_shims_drive_the_frame(torch_cuda, tmp_path):
    assert os.path.exists(EXE), f"{EXE} missing: build it with __graft_entry__.build() (make -C tests/cpp)"
    sc = scenes.reference_scene()
    W, H, f_plain, f_cache = 96, 64, 3, 7
    an, mesh = D._prims_from_scene(sc)
    luz = np.zeros(1, D.LUZ)
    luz["pos"], luz["I"] = sc.lights[0, 0:3], sc.lights[0, 4:7]
    inp, out = tmp_path / "in", tmp_path / "out"
    inp.mkdir()
    out.mkdir()
    (inp / "params.txt").write_text(f"{W} {H} {f_plain} {f_cache}\n")
    (inp / "prims.bin").write_bytes(an.tobytes())
    (inp / "mesh.bin").write_bytes(mesh.tobytes())
    (inp / "luces.bin").write_bytes(luz.tobytes())
    print(f"running {EXE}", flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    n = W * H
    rd = lambda name, dt: np.fromfile(out / name, dt)
    counter, n_train, loss, loaded, n_nodes, n_prims = (out / "scalars.txt").read_text().split()
    counter, n_train, loss = int(counter), int(n_train), float(loss)
    assert int(loaded) == 1 and int(n_prims) == sc.n_tris and int(n_nodes) > 0

    # plain frame vs the oracle integrator (the shim's CWBVH: the library's build
    # of the same triangles, so the oracle traces a tree built the same way)
    host = _lib.CwbvhHost(sc.verts)
    R = oracle.Renderer(sc, host.nodes, host.tris, W, H)
    plain = rd("img_plain.bin", np.float32).reshape(n, 3)
    ref = R.render(f_plain, 1, 20, 10)[0]
    ok = _agree(plain, ref, 1e-3, 1e-3)
    n_bad, n_expl, base = conditioning.explain(R, plain, ref, f_plain, 20, 10)
    print(f"plain frame agreement {ok.mean()}; {n_bad} paths outside 1e-3 relative, {n_expl} of them "
          f"ill-conditioned in the oracle (base rate {base:.4f})", flush=True)
    assert ok.mean() >= 0.95
    assert n_expl >= 0.9 * n_bad

    # cache frame: records and queries
    o = R.render_nrc(f_cache, 20, 10, 3)
    heads = int(((o["flags"] & 3) == 3).sum())
    assert abs(counter - heads) <= max(3, 0.05 * heads) and counter > 0
    thr = rd("thr.bin", np.float32).reshape(n, 3)
    need = ((o["flags"] & 4) != 0) | ((o["flags"] & 3) == 3)
    has = np.any(thr != 0, axis=1)
    assert (has == need).mean() >= 0.95

    # inference + composite
    w0, e0 = _model(out / "model_before.bin")
    inf = np.frombuffer((out / "inference.bin").read_bytes(), D.DATOS)
    pred = rd("pred.bin", np.float32).reshape(n, 3)
    pred_ok = _agree(pred, _net(e0, inf), 2e-2, 1e-3)
    print("inference agreement", pred_ok.mean(), flush=True)
    assert pred_ok.mean() >= 0.99
    prefix = rd("img_prefix.bin", np.float32).reshape(n, 3)
    comp = rd("img_composite.bin", np.float32).reshape(n, 3)
    expect = prefix + np.where(has[:, None], thr * pred, 0.0).astype(np.float32)
    np.testing.assert_allclose(comp, expect, rtol=1e-6, atol=1e-7)

    # tail inputs, tail inference and targets (RenderKernel.cu:92-166)
    assert n_train == min(counter, 65536, n) // 4 * 4 and n_train > 0
    regs = np.frombuffer((out / "registros.bin").read_bytes(), D.REG)
    tail_in = np.frombuffer((out / "tail_in.bin").read_bytes(), D.DATOS)
    valid = regs["valido"] != 0
    assert valid.all()                            # the drop-in writes only valid records (render.h:461-464)
    for f_in, f_reg in (("pos", "pos"), ("dir", "dir"), ("nrm", "nrm"), ("kd", "kd"), ("ks", "ks")):
        np.testing.assert_array_equal(tail_in[f_in][valid], regs["tail"][f_reg][valid])
        assert not tail_in[f_in][~valid].any()
    tail_pred = rd("tail_pred.bin", np.float32).reshape(n_train, 3)
    assert _agree(tail_pred, _net(e0, tail_in), 2e-2, 1e-3).mean() >= 0.99
    train = np.frombuffer((out / "train_final.bin").read_bytes(), D.DATOS)
    luzs = regs["luz"].astype(np.float64)
    with np.errstate(all="ignore"):
        tgt = (luzs + regs["thr"].astype(np.float64) * M.sanitize(tail_pred)) / np.fmax(regs["norm"].astype(np.float64),
                                                                                       1e-8)
        tgt = np.fmin(tgt, 1e5)
        tgt[np.isnan(tgt) | np.isinf(tgt)] = 0.0
    np.testing.assert_allclose(train["color"], tgt, rtol=1e-6, atol=1e-7)
    for f_in in ("pos", "dir", "nrm", "kd", "ks"):
        np.testing.assert_array_equal(train[f_in], regs["head"][f_in])

    # training step, save / load
    assert np.isfinite(loss) and loss > 0
    w1, e1 = _model(out / "model_after.bin")
    assert not np.array_equal(w1, w0) and np.isfinite(w1).all()
    reloaded = rd("pred_reloaded.bin", np.float32).reshape(n, 3)
    assert _agree(reloaded, _net(e1, inf), 2e-2, 1e-3).mean() >= 0.99
