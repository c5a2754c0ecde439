"""FP64 NumPy oracle of the radiance-cache network (test infrastructure only).

Restates, in float64:
  * the batch-index draw of generarIndicesAleatorios (NrcKernel.cu:10-16):
    idx[i] = pcg32(seed ^ i) % n (rng.h:12-17);
  * the training target of kernelCalcularTargets (RenderKernel.cu:92-137);
  * output sanitising of guardarSalidaInferencia (NrcKernel.cu:99-126);
  * position normalisation of normalize_coord (NrcKernel.cu:5-8);
  * the network of mlp.cpp:22-70 as built by tfg_amd: 64 encoded inputs,
    5 hidden ReLU layers of 64, 3 linear outputs, no biases;
  * the RelativeL2Luminance loss (luminance of the prediction held
    constant) and its gradient;
  * Adam (Kingma & Ba 2015) and the EMA of the weights.

The input-slot permutation and feature meanings match tfg_amd/csrc/nrc_device.h:
slot (c, g, j) of the 64 inputs is feature index perm(c, g, j).

Parity with tiny-cuda-nn itself (absent from the reference tree) is
unpinned; this oracle is pinned by finite differences (gradients), closed
forms (encoding, Adam) and by the reference's own integer recurrences.
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 5 * 64 * 64 + 16 * 64
LUM = np.array([0.299, 0.587, 0.114])


def perm(c: int, g: int, j: int) -> int:
    return 32 * c + (4 * g + j if j < 4 else 16 + 4 * g + j - 4)


def pcg32(state: np.ndarray) -> np.ndarray:
    """rng.h:12-17 on a uint32 array (returns the output; state is advanced in place)."""
    s = (state.astype(np.uint64) * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xFFFFFFFF)
    state[...] = s.astype(np.uint32)
    word = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & np.uint64(0xFFFFFFFF)
    return ((word >> np.uint64(22)) ^ word).astype(np.uint32)


def batch_indices(seed: int, batch: int, n: int) -> np.ndarray:
    """NrcKernel.cu:10-16."""
    st = (np.uint32(seed) ^ np.arange(batch, dtype=np.uint32)).astype(np.uint32)
    return (pcg32(st) % np.uint32(n)).astype(np.int64)


def sanitize(x):
    """NrcKernel.cu:109-123."""
    x = np.array(x, np.float64, copy=True)
    x[x < 0] = 0.0
    x[np.isnan(x)] = 0.0
    x[x > 1e5] = 1e5
    x[np.isinf(x)] = 1e5
    return x


def targets(rec: np.ndarray, pred_tail: np.ndarray) -> np.ndarray:
    """RenderKernel.cu:113-126 with records in the tfg layout (11 x 4 floats)."""
    r = rec.reshape(-1, 11, 4).astype(np.float64)
    ls, ts, nm = r[:, 8, :3], r[:, 9, :3], r[:, 10, :3]
    with np.errstate(all="ignore"):
        t = (ls + ts * sanitize(pred_tail)) / np.fmax(nm, 1e-8)
        t = np.fmin(t, 1e5)          # C fminf: a NaN operand yields the other one
        t[np.isnan(t) | np.isinf(t)] = 0.0
    return t


def _oneblob(s, g):
    d = (s - (g + 0.5) * 0.25) * 4.0
    return np.exp(-0.5 * d * d)


def _sph(m):
    d = 2.0 * m - 1.0
    th = np.arccos(np.clip(d[:, 2], -1.0, 1.0)) / np.pi
    ph = np.arctan2(d[:, 1], d[:, 0]) / (2.0 * np.pi) + 0.5
    return th, ph


def encode(verts: np.ndarray, bmin, bmax) -> np.ndarray:
    """(n, 16) vertex records -> (n, 64) features (index = logical input)."""
    v = verts.reshape(-1, 16).astype(np.float64)
    bmin = np.asarray(bmin, np.float64)
    ext = np.asarray(bmax, np.float64) - bmin
    inv = (1.0 / ext.astype(np.float32)).astype(np.float64)   # the kernel stores 1/extent in FP32
    p = np.clip((v[:, 0:3] - bmin) * inv, 0.0, 1.0)
    rough = v[:, 3]
    dth, dph = _sph(v[:, 4:7])
    nth, nph = _sph(v[:, 8:11])
    kd = np.stack([v[:, 7], v[:, 11], v[:, 12]], 1)
    ks = v[:, 13:16]
    n = v.shape[0]
    x = np.zeros((n, 64))

    def s(d, f):
        return np.sin((2.0 ** f) * np.pi * p[:, d])

    for g in range(4):
        c0 = [s(0, g), s(0, 4 + g), s(0, 8 + g), s(1, g), s(1, 4 + g), s(1, 8 + g), s(2, g), s(2, 4 + g)]
        c1 = [s(2, 8 + g), _oneblob(dth, g), _oneblob(dph, g), _oneblob(nth, g), _oneblob(nph, g),
              _oneblob(rough, g), kd[:, g] if g < 3 else np.ones(n), ks[:, g] if g < 3 else np.zeros(n)]
        for j in range(8):
            x[:, perm(0, g, j)] = c0[j]
            x[:, perm(1, g, j)] = c1[j]
    return x


def split(params: np.ndarray):
    """Flat parameter vector -> [W0..W4 (64x64), W5 (16x64)]."""
    p = np.asarray(params, np.float64)
    Ws = [p[l * 4096:(l + 1) * 4096].reshape(64, 64) for l in range(5)]
    Ws.append(p[5 * 4096:].reshape(16, 64))
    return Ws


def round_to(x, dtype):
    """Round float64 values through bf16 or fp16 (round to nearest even)."""
    if dtype == "fp16":
        return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)
    if dtype == "bf16":
        f = np.asarray(x, np.float32)
        u = f.view(np.uint32).astype(np.uint64)
        rounding = ((u >> 16) & 1) + 0x7FFF
        u = ((u + rounding) >> 16) << 16
        out = u.astype(np.uint32).view(np.float32).astype(np.float64)
        nan = np.isnan(f)
        out[nan] = np.nan
        return out
    return np.asarray(x, np.float64)


def forward(params, x, emulate=None):
    """Returns (pre-activations per layer [6 x (n, rows)], layer inputs [6 x (n, 64)]).

    emulate = "bf16" / "fp16" rounds weights, inputs and inter-layer
    activations like the matrix-core kernels do (accumulation stays FP64)."""
    Ws = split(params)
    if emulate:
        Ws = [round_to(W, emulate) for W in Ws]
    h = round_to(x, emulate) if emulate else np.asarray(x, np.float64)
    pre, ins = [], []
    for l, W in enumerate(Ws):
        ins.append(h)
        z = h @ W.T
        pre.append(z)
        if l < 5:
            h = np.maximum(z, 0.0)
            if emulate:
                h = round_to(h, emulate)
    return pre, ins


def loss_and_grad(params, x, target, emulate=None):
    """RelativeL2Luminance mean over samples and the 3 channels, with the
    luminance of the prediction treated as a constant; returns (loss, grad)."""
    Ws = split(params)
    if emulate:
        Ws = [round_to(W, emulate) for W in Ws]
    pre, ins = forward(params, x, emulate)
    out = pre[5][:, :3]
    n = out.shape[0]
    lum = out @ LUM
    inv = 1.0 / (lum * lum + 0.01)
    d = out - target
    loss = float(np.sum(d * d * inv[:, None]) / (3 * n))
    dy = np.zeros((n, 16))
    dy[:, :3] = 2.0 * d * inv[:, None] / (3 * n)
    grads = [None] * 6
    for l in range(5, -1, -1):
        grads[l] = dy.T @ ins[l]
        if l > 0:
            dx = dy @ Ws[l]
            dy = dx * (ins[l] > 0)
    g = np.concatenate([grads[l].ravel() for l in range(6)])
    g[5 * 4096 + 3 * 64:] = 0.0
    return loss, g


def loss_only(params, x, target, lum_params=None):
    """Loss with the luminance denominator computed from lum_params (for
    finite differences with the denominator held fixed)."""
    pre, _ = forward(params, x)
    out = pre[5][:, :3]
    lp = forward(lum_params if lum_params is not None else params, x)[0][5][:, :3]
    lum = lp @ LUM
    d = out - target
    return float(np.sum(d * d / (lum * lum + 0.01)[:, None]) / (3 * out.shape[0]))


def adam_step(w, m, v, ema, g, t, lr, b1, b2, eps, decay):
    """One Adam step at step number t (1-based) and the EMA update."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    w = w - lr * mh / (np.sqrt(vh) + eps)
    ema = decay * ema + (1 - decay) * w
    return w, m, v, ema
