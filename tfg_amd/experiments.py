"""Equal-time comparison of the cache-accelerated and the plain path tracer.

The frame schedule restates main.cpp:958-1089: ``warmup`` frames with 100%
training paths and no inference, then frames in which 5% of the pixels
train and the rest query the cache at their first diffuse vertex with a
large enough footprint.  The reference image is plain path tracing with FP64
accumulation; relMSE = mean over pixels and channels of
(I - R)^2 / (R^2 + 0.01).
"""
from __future__ import annotations

import time

import numpy as np

from . import nrc, pt


def relmse(img: np.ndarray, ref: np.ndarray) -> float:
    return float(np.mean((img - ref) ** 2 / (ref * ref + 0.01)))


def _sync_time(torch):
    torch.cuda.synchronize()
    return time.perf_counter()


def _run_budget(torch, render, budget_s, first):
    """render(first), render(first + 1), ... until budget_s of wall time have
    passed, with one frame in flight: after submitting frame n the host waits
    for frame n - 1 and checks the clock, so the GPU never idles between
    frames while the host enqueues the next one.  The time returned runs to
    the end of the last submitted frame.  Returns (frames, seconds)."""
    evs = [torch.cuda.Event(), torch.cuda.Event()]
    t0 = _sync_time(torch)
    n = 0
    while True:
        render(first + n)
        evs[n % 2].record()
        if n > 0:
            evs[(n - 1) % 2].synchronize()
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    return n, _sync_time(torch) - t0


def _graph(schedule):
    return bool(schedule.get("use_graph", True)) if schedule else True


def _apply(tr, schedule):
    if schedule:
        tr.set_schedule(**{k: v for k, v in schedule.items() if k in tr.schedule()})


def reference_image(torch, dscene, W, H, max_depth, rr_depth, spp, frame0=1_000_000, schedule=None):
    tr = pt.PathTracer(W, H, max_depth, rr_depth, mode=0, use_graph=_graph(schedule))
    tr.set_scene(dscene)
    _apply(tr, schedule)
    tr.set_camera(dscene.scene.camera.frame(W, H))
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    for k in range(spp):
        tr.render(frame0 + k, img, acc)
    torch.cuda.synchronize()
    return (acc / spp).cpu().numpy()


def plain_run(torch, dscene, W, H, max_depth, rr_depth, budget_s, frame0=0, schedule=None):
    """Plain FP32 PT for budget_s seconds of wall time; returns (mean image, frames, seconds)."""
    tr = pt.PathTracer(W, H, max_depth, rr_depth, mode=0, use_graph=_graph(schedule))
    tr.set_scene(dscene)
    _apply(tr, schedule)
    tr.set_camera(dscene.scene.camera.frame(W, H))
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    tr.render(frame0, img)          # capture the graph outside the timed region
    n, secs = _run_budget(torch, lambda f: tr.render(f, img, acc), budget_s, frame0 + 1)
    return (acc / n).cpu().numpy(), n, secs


def nrc_run(torch, dscene, W, H, max_depth, rr_depth, budget_s, warmup=48, precision=0, train_fraction=0.05,
            frame0=0, cache_cfg=None, count_warmup=True, schedule=None, cache_tail_from=None, defer_training=True):
    """Cache-accelerated PT: `warmup` training-only frames, then train+infer
    frames until the budget (wall time including warmup when
    count_warmup) is spent.  Returns (mean image, frames, seconds, cache stats).

    Train+infer frames end most paths at their first cache query, so depths
    >= 2 hold little more than the training paths; they run with the
    schedule's other settings but ``tail_from=cache_tail_from`` (default
    CACHE_TAIL_FROM; negative keeps the schedule's value).  With direct
    launches and defer_training, each frame's optimiser steps run beside the
    next frame's tracing (schedule defer_training: same images and weights);
    the last frame's steps are flushed after the clock stops."""
    tr = pt.PathTracer(W, H, max_depth, rr_depth, mode=pt_mode_train(), train_fraction=train_fraction,
                       use_graph=_graph(schedule))
    tr.set_scene(dscene)
    _apply(tr, schedule)
    if defer_training and not _graph(schedule):
        tr.set_schedule(defer_training=True)
    tr.set_camera(dscene.scene.camera.frame(W, H))
    sc = dscene.scene
    lo = sc.verts.reshape(-1, 3).min(0) - 0.5
    hi = sc.verts.reshape(-1, 3).max(0) + 0.5
    kw = dict(precision=precision, bounds_min=tuple(map(float, lo)), bounds_max=tuple(map(float, hi)))
    if cache_cfg:
        kw.update(cache_cfg)
    cache = nrc.RadianceCache(**kw)
    tr.set_cache(cache)
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    t0 = _sync_time(torch)
    for k in range(warmup):
        tr.render(frame0 + k, img)
    t_w = _sync_time(torch)
    tr.set_mode(pt_mode_train() | pt_mode_infer())
    cache_tail_from = CACHE_TAIL_FROM if cache_tail_from is None else cache_tail_from
    if cache_tail_from >= 0:
        tr.set_schedule(tail_from=cache_tail_from)
    spent = t_w - t0 if count_warmup else 0.0
    n, secs = _run_budget(torch, lambda f: tr.render(f, img, acc), budget_s - spent, frame0 + warmup)
    tr.flush_training()   # the last frame's steps, after the clock
    torch.cuda.synchronize()
    return (acc / n).cpu().numpy(), n, secs + spent, cache.stats(), t_w - t0


# Train+infer frames end most paths at their first cache query, so the late
# depths hold little more than the 5% training paths and the per-depth launch
# chain is bound by latency there.  On the 64-instance scene in its closed
# room at 1080p the queues of depths 2..7 hold 126K, 63K, 37K, 3.5K, 1.7K and
# 0.9K rays; one tail launch for depths 5.. takes a train+infer frame from
# 2.81 ms (no tail launch; 3.10 ms with the tail from depth 2) to 2.76 ms.
# Plain and training-only frames, whose late depths are full, keep the tuned
# value (0).
CACHE_TAIL_FROM = 5


def pt_mode_train():
    return 2


def pt_mode_infer():
    return 1


def equal_time(torch, scene, W, H, max_depth, rr_depth, ref_spp, budget_s, warmup=48, precision=0, schedule=None,
               repeats=1, seed_stride=100_000):
    """Both methods (and the reference) run with the same frame schedule.

    With repeats > 1 the pair of runs is repeated with frame seeds offset by
    seed_stride (a fresh cache each time) against the same reference: the
    relMSE of a 1-spp-per-frame image is dominated by rare bright samples, so
    one pair of runs scatters widely (plain relMSE 1.4e-4 to 2.0e-4 between
    seeds at 1080p, 2 s).  The headline ratio is the pooled one, mean plain
    relMSE / mean cache relMSE; the per-run values are listed too."""
    dscene = pt.DeviceScene(scene, torch)
    ref = reference_image(torch, dscene, W, H, max_depth, rr_depth, ref_spp, schedule=schedule)
    runs = []
    for k in range(repeats):
        f0 = k * seed_stride
        p_img, p_n, p_t = plain_run(torch, dscene, W, H, max_depth, rr_depth, budget_s, frame0=f0,
                                    schedule=schedule)
        n_img, n_n, n_t, st, w_t = nrc_run(torch, dscene, W, H, max_depth, rr_depth, budget_s, warmup=warmup,
                                           precision=precision, frame0=f0, schedule=schedule)
        rp = relmse(p_img, ref)
        rn = relmse(n_img, ref)
        runs.append({"frame0": f0,
                     "plain": {"relmse": rp, "frames": p_n, "seconds": p_t},
                     "nrc": {"relmse": rn, "frames": n_n, "seconds": n_t, "warmup_frames": warmup, "warmup_s": w_t,
                             "cache": st},
                     "ratio": rp / rn if rn > 0 else None})
    mp = float(np.mean([r["plain"]["relmse"] for r in runs]))
    mn = float(np.mean([r["nrc"]["relmse"] for r in runs]))
    return {
        "ref_spp": ref_spp,
        "budget_s": budget_s,
        "repeats": repeats,
        "plain": {"relmse": mp, "frames": float(np.mean([r["plain"]["frames"] for r in runs]))},
        "nrc": {"relmse": mn, "frames": float(np.mean([r["nrc"]["frames"] for r in runs])), "warmup_frames": warmup},
        "runs": runs,
        "relmse_ratio_plain_over_nrc": mp / mn if mn > 0 else None,
    }


def reconstruct(torch, dscene, W, H, max_depth, rr_depth, model_path, spp, frame0=0, png_path=None, tonemap=True,
                precision=None, schedule=None, bounds=None):
    """Reconstruction mode (renderizarModoReconstruccion, main.cpp:454-668):
    a trained cache is loaded from a weight file (TFGNRC01, RadianceCache.save)
    and `spp` frames are rendered with cache inference and no training
    (usar_red = true, entrenar_red = false, main.cpp:566-574): every path ends
    at its first cacheable vertex, whose cached radiance is composited into
    the pixel (main.cpp:596-597), and the frames (seeds frame0 + sample, as
    main.cpp:568 passes the sample index) are accumulated in FP64 on the GPU.
    The mean goes through the ACES filmic curve (main.cpp:640) into an 8-bit
    PNG at png_path (optional).

    bounds: the cache's scene bounds (min, max); default: the scene's vertex
    box widened by 0.5, as the training routines here choose them.  Returns
    (mean image (H*W, 3) float64, rgb8 bytes (H*W, 3) uint8)."""
    from . import image

    tr = pt.PathTracer(W, H, max_depth, rr_depth, mode=pt_mode_infer(), use_graph=_graph(schedule))
    tr.set_scene(dscene)
    _apply(tr, schedule)
    tr.set_camera(dscene.scene.camera.frame(W, H))
    if bounds is None:
        v = dscene.scene.verts.reshape(-1, 3)
        bounds = (v.min(0) - 0.5, v.max(0) + 0.5)
    kw = dict(bounds_min=tuple(map(float, bounds[0])), bounds_max=tuple(map(float, bounds[1])))
    if precision is not None:
        kw["precision"] = precision
    cache = nrc.RadianceCache(**kw)
    cache.load(model_path)
    tr.set_cache(cache)
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    for k in range(spp):
        tr.render(frame0 + k, img, acc)
    rgb8 = image.resolve(torch, acc, spp, tonemap=tonemap)
    torch.cuda.synchronize()
    out = rgb8.cpu().numpy()
    if png_path:
        image.write_png(png_path, out, W, H)
    mean = (acc / max(spp, 1)).cpu().numpy()
    tr.close()
    cache.close()
    return mean, out


# gridsearch.h:54-61: the reference's search space for the optimiser settings
# (its network width/depth, hash-grid size and one-blob bins have no
# counterpart: this cache's network and encoding are fixed)
SWEEP_GRID = {
    "lr": (1e-2, 5e-3, 1e-3),
    "ema_decay": (0.95, 0.99, 0.999),
    "steps_per_frame": (1, 4),
    "precision": (0, 1),
}


def grid_search(torch, scene, W, H, max_depth, rr_depth, grid=None, frames=100, score_from=10, schedule=None,
                frame0=0):
    """Hyper-parameter sweep of the radiance cache, restating
    gridsearch.h:39-314: for every combination of `grid` (default
    SWEEP_GRID), a fresh cache trained by `frames` training-only frames
    (every pixel a training path, gridsearch.h:178-186); the score is the mean
    of the frames' training losses from frame `score_from` on
    (gridsearch.h:196-205), and the lowest mean wins (gridsearch.h:214-217;
    non-finite or zero scores never win).  Every configuration sees the same
    frame seeds.  Returns the results sorted by score, best first."""
    import itertools

    grid = grid or SWEEP_GRID
    dscene = pt.DeviceScene(scene, torch)
    lo = scene.verts.reshape(-1, 3).min(0) - 0.5
    hi = scene.verts.reshape(-1, 3).max(0) + 0.5
    keys = sorted(grid)
    # direct launches unless the schedule asks for graph replay: the sweep
    # swaps the cache once per configuration, and every swap re-captures a
    # replayed frame (a host segmentation fault inside the HIP runtime ended
    # one such run of capture / destroy cycles; direct launches have none)
    use_graph = bool(schedule.get("use_graph", False)) if schedule else False
    tr = pt.PathTracer(W, H, max_depth, rr_depth, mode=pt_mode_train(), use_graph=use_graph)
    tr.set_scene(dscene)
    _apply(tr, schedule)
    tr.set_camera(scene.camera.frame(W, H))
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    results = []
    for combo in itertools.product(*(grid[k] for k in keys)):
        hp = dict(zip(keys, combo))
        cache = nrc.RadianceCache(bounds_min=tuple(map(float, lo)), bounds_max=tuple(map(float, hi)), **hp)
        tr.set_cache(cache)
        losses = []
        t0 = _sync_time(torch)
        for f in range(frames):
            tr.render(frame0 + f, img)
            st = cache.stats()   # synchronises: the frame's last loss
            if f >= score_from:
                losses.append(st["last_loss"])
        t = _sync_time(torch) - t0
        score = float(np.mean(losses)) if losses else float("inf")
        if not np.isfinite(score) or score <= 1e-7:
            score = float("inf")
        results.append({**hp, "score": score, "final_loss": losses[-1] if losses else None, "seconds": t,
                        "steps": cache.stats()["steps"]})
        tr.set_cache(None)
        del cache
    tr.close()
    return sorted(results, key=lambda r: r["score"])
