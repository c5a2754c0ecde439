"""Benchmark: prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pt1080|pt|traversal]

Workload ``pt1080`` (default, the metric's configuration): the plain FP32
wavefront path tracer at 1920x1080 on the 64-instance scene (value = its
Mrays/s over K frames with the tuned schedule), followed by the equal-time relMSE
comparison of the cache-accelerated tracer against it.

Workload ``pt``: the Mrays/s part alone.

Workload ``nrc``: the cache's inference and training kernels alone, on the
data of a 1080p cache frame (also part of the default run's JSON line as
``nrc_kernels``).

Workload ``copy``: a plain device-to-device copy of 1 GiB (torch ``copy_``),
the calibration run for the HBM byte counters (FETCH_SIZE / WRITE_SIZE per
dispatch of a kernel that moves a known number of bytes).

Workload ``traversal``: the traversal microbenchmark on a 2^20-triangle
Perlin-displaced sphere (CWBVH built by this library).  One step = 2^21
primary closest-hit rays + 2^21 cosine bounce closest-hit rays + 2^21
shadow any-hit rays, all resident in HBM before timing starts.

The roofline object is for the dominant kernel (bounce closest-hit):
algorithmic bytes per launch = sum over rays of 32 B ray + 16 B hit +
80 B per node decoded + 48 B per triangle tested, with the per-ray counts
taken from the counting variant of the same kernel on the same rays; its
launch time is measured with HIP events on the stream it runs on.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# a fatal signal prints the Python stack of every thread before the process ends
faulthandler.enable(all_threads=True)

METRIC = "Mrays/s + equal-time relMSE, NRC vs plain PT, 1080p, 1 MI355X"
HBM_PEAK_GBS = 8000.0

# HBM bytes between L2 and memory (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# runs of this file's workloads, scaled by the same counters on a 1 GiB copy:
# FETCH_SIZE x 2.0, WRITE_SIZE x 1.0), from the committed summaries in
# profiles/ named in "traffic_source": per launch for the traversal
# microbench; per frame (the 8 launches of the stage in the roofline frames,
# overlap off, saved schedule) for the path tracer's stages.
MEASURED_TRAFFIC = {
    "k_trace_bounce": (366.2e6, "profiles/k_trace_bounce_displaced_sphere_1M.md"),
    "k_extend": (2010.0e6, "profiles/k_extend_k_connect_pt_1080p_instances.md"),
    "k_connect": (3327.0e6, "profiles/k_extend_k_connect_pt_1080p_instances.md"),
}


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_traversal(args, torch):
    from tfg_amd import _lib, scenes, workloads

    t0 = time.time()
    sc = scenes.displaced_sphere()
    host = _lib.CwbvhHost(sc.verts, 0)
    dev = host.upload()
    _log(f"scene {sc.n_tris} tris, {host.num_nodes} nodes, depth {host.max_depth}, built in {time.time() - t0:.1f}s")
    W, H = 2048, 1024
    prim = workloads.primary_rays(sc, W, H, seed=7)
    n = prim.shape[0]
    d_prim = torch.from_numpy(prim).cuda()
    d_hp = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    dev.trace_closest(d_prim, d_hp, n)
    torch.cuda.synchronize()
    hp = d_hp.cpu().numpy().view(np.uint32)
    hits = np.zeros(n, [("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.uint32)])
    hits["t"] = hp[:, 0].view(np.float32)
    hits["prim"] = hp[:, 3]
    bounce = workloads.bounce_rays(sc, prim, hits, seed=8)
    shadow = workloads.shadow_rays(sc, prim, hits, seed=9)
    d_b = torch.from_numpy(bounce).cuda()
    d_s = torch.from_numpy(shadow).cuda()
    d_hb = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    d_occ = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.empty((n, 2), dtype=torch.int32, device="cuda")

    # algorithmic bytes of each launch from the counting variant
    def algo_bytes(fn, rays_dev):
        fn(rays_dev, cnt)
        torch.cuda.synchronize()
        c = cnt.cpu().numpy().view(np.uint32).astype(np.int64)
        nodes = c[:, 0] & 0x7FFFFFFF
        assert not (c[:, 0] >> 31).any(), "traversal stack overflow"
        return int((32 + 16) * n + 80 * nodes.sum() + 48 * c[:, 1].sum()), float(nodes.mean()), float(c[:, 1].mean())

    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    launches = {
        "primary_closest": lambda c=None: dev.trace_closest(d_prim, d_hp, n, counts=c, stream=sh),
        "bounce_closest": lambda c=None: dev.trace_closest(d_b, d_hb, n, counts=c, stream=sh),
        "shadow_any": lambda c=None: dev.trace_any(d_s, d_occ, n, counts=c, stream=sh),
    }
    bytes_per = {}
    visits = {}
    for k, f in launches.items():
        b, nv, nt = algo_bytes(lambda r, c, f=f: f(c), None)
        bytes_per[k] = b
        visits[k] = (nv, nt)
    _log(f"per-ray visits (nodes, tris): {visits}")

    for _ in range(args.warmup):
        for f in launches.values():
            f()
    torch.cuda.synchronize()
    ev = {k: [] for k in launches}
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        for k, f in launches.items():
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            f()
            b.record(stream)
            ev[k].append((a, b))
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    total_ms = start.elapsed_time(stop)
    per_kernel_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
    ms_step = total_ms / args.steps
    rays_step = 3 * n
    value = rays_step / (ms_step * 1e-3) / 1e6
    dom = "bounce_closest"
    achieved = bytes_per[dom] / (per_kernel_ms[dom] * 1e-3) / 1e9
    res = {
        "value": value,
        "ms_per_step": ms_step,
        "per_kernel_ms": per_kernel_ms,
        "mrays_per_kernel": {k: n / (v * 1e-3) / 1e6 for k, v in per_kernel_ms.items()},
        "roofline": {
            "kernel": "k_trace<watertight, closest> (bounce rays)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC["k_trace_bounce"][0],
            "traffic_source": MEASURED_TRAFFIC["k_trace_bounce"][1],
            "algorithmic_bytes": bytes_per[dom],
        },
        "config": {
            "workload": "traversal microbench: displaced sphere 1,048,576 tris (Perlin amp 0.05, seed 1), "
                        "2^21 primary + 2^21 cosine bounce closest-hit + 2^21 shadow any-hit rays",
            "tris": sc.n_tris,
            "cwbvh_nodes": host.num_nodes,
            "tree_depth": host.max_depth,
            "rays_per_step": rays_step,
            "visits_per_ray": visits,
            "seeds": {"scene": 1, "primary": 7, "bounce": 8, "shadow": 9},
        },
        "wall_s": wall,
        "dtype": "fp32",
    }
    res["_cpu_inputs"] = (host, prim, bounce, shadow)
    return res


# The schedule tfg_pt_autotune picked for the 1080p workload below on MI355X
# (plain frames, direct launches).  The default run uses it as saved, so the
# timed frames, the roofline frames and every counter pass of a profile run
# describe the same launches (the tuner's pick between near-equal candidates
# varies from run to run); --tune re-tunes instead.
TUNED_SCHEDULE_1080P = {"trace_block": 64, "grid_mult": 4, "tiles": 1, "side_streams": 2, "overlap": True,
                        "xcd_group": 0, "octet_below": 131072, "connect_per_cu": 0, "connect_batch": 0,
                        "top_nodes": 1, "tail_from": 0, "shadow_sort_from": 1, "ray_sort_from": 0,
                        "refill_from": 1, "refill_idle": 32, "refill_per_cu": 14}


def _stage_roofline(prof, frames, stage, n_launch):
    """Algorithmic bytes and summed launch time of one traversal stage over
    `frames` profiled frames (tfg_pt_profile_frame dicts)."""
    if stage == "extend":
        rays = sum(sum(p["ext_rays"]) for p in prof)
        nodes = sum(p["ext_nodes"] for p in prof)
        tris = sum(p["ext_tris"] for p in prof)
        ms = sum(sum(p["ms_extend"]) for p in prof)
    else:
        rays = sum(sum(p["shadow_rays"]) for p in prof)
        nodes = sum(p["shadow_nodes"] for p in prof)
        tris = sum(p["shadow_tris"] for p in prof)
        ms = sum(sum(p["ms_connect"]) for p in prof)
    byt = 48 * rays + 80 * nodes + 48 * tris
    return {"rays_per_frame": rays / frames, "bytes_per_frame": byt / frames, "ms_per_frame": ms / frames,
            "bytes_per_launch": byt / frames / n_launch, "avg_launch_ms": ms / frames / n_launch,
            "gbs": byt / (ms * 1e-3) / 1e9, "nodes_per_ray": nodes / rays, "tris_per_ray": tris / rays}


def bench_pt(args, torch):
    """Plain FP32 wavefront path tracer, 64-instance scene in its closed room,
    1920x1080, 1 spp per frame, max depth 8, Russian roulette after depth 3;
    frames with the saved tuned schedule (or a fresh tune with --tune),
    replayed as per-frame hipGraphs (the schedule is tuned, profiled and
    compared on direct launches; the A/B of both is part of the line).  One
    step = one frame.

    Roofline: the traversal stages are then timed on `steps` further frames
    of the same schedule with the shadow stage on the main stream (overlap
    off), so that every launch has the GPU to itself and its HIP-event
    duration is the kernel's own; the dominant stage is the one with the
    larger total.  Algorithmic bytes per launch = (48 B per ray + 80 B per
    node decoded + 48 B per triangle tested) / launches, the node and
    triangle counts from the counting pass of the same frames."""
    from tfg_amd import pt, scenes

    t0 = time.time()
    sc = scenes.instances()
    t_scene = time.time() - t0
    t0 = time.time()
    ds = pt.DeviceScene(sc, torch)
    t_build = time.time() - t0
    _log(f"scene {sc.n_tris} tris, {ds.host.num_nodes} nodes, depth {ds.host.max_depth} "
         f"(scene {t_scene:.1f}s, CWBVH build + upload {t_build:.1f}s)")
    W, H, MD, RR = 1920, 1080, 8, 3
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    accum = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    # direct launches for tuning and the profiled passes; the timed frames
    # replay the same schedule as per-frame hipGraphs
    tracer = pt.PathTracer(W, H, MD, RR, mode=0, use_graph=False)
    tracer.set_scene(ds)
    tracer.set_camera(sc.camera.frame(W, H))
    if args.tune:
        tracer.autotune()
        tuned = "tfg_pt_autotune in this run"
    else:
        tracer.set_schedule(**TUNED_SCHEDULE_1080P)
        tuned = "saved (bench.TUNED_SCHEDULE_1080P, an earlier tfg_pt_autotune pick on MI355X)"
    shape = tracer.launch_shape()
    _log(f"schedule ({tuned}) {shape}")
    timed = pt.PathTracer(W, H, MD, RR, mode=0, use_graph=True)
    timed.set_scene(ds)
    timed.set_camera(sc.camera.frame(W, H))
    timed.set_schedule(**{k: v for k, v in shape.items() if k in timed.schedule() and k != "use_graph"})
    timed.render(0, img, accum)   # captures the frame's graphs (set-up, whatever the warmup count)
    for f in range(args.warmup):
        timed.render(f, img, accum)
    torch.cuda.synchronize()
    timed.reset_totals()
    stream = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    a.record(stream)
    for k in range(args.steps):
        timed.render(1000 + k, img, accum, stream=stream.cuda_stream)
    b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_frame = a.elapsed_time(b) / args.steps
    ext, shd = timed.ray_totals()
    rays_frame = (ext + shd) / args.steps
    value = rays_frame / (ms_frame * 1e-3) / 1e6
    timed_shape = timed.launch_shape()
    _log(f"timed {args.steps} frames (graph replay): {ms_frame:.4f} ms per frame, {value:.1f} Mrays/s")
    for f in range(args.warmup):
        tracer.render(f, img, accum)
    torch.cuda.synchronize()
    # the same frames replayed as per-frame hipGraphs (tfg_pt_render with the
    # graph option), against direct launches, in the timed schedule and with
    # the shadow stage on the main stream (no fork in the graph); each graph
    # tracer captures its frame once and is closed only after the device is
    # idle (host faults inside the HIP runtime were seen in runs that dropped
    # and re-captured graphs many times)
    graph_ab = _graph_vs_direct(torch, tracer, timed, ds, sc, W, H, MD, RR, img, accum, stream, args.steps)
    torch.cuda.synchronize()
    timed.close()
    # one profiled frame in the timed schedule (stages overlapped as timed)
    prof = tracer.profile_frame(1000, img)
    # roofline frames: the same schedule serialised (overlap off)
    tracer.set_schedule(overlap=False)
    n_roof = max(3, args.steps)
    rprof = [tracer.profile_frame(1000 + k, img) for k in range(n_roof)]
    tracer.set_schedule(overlap=shape["overlap"])
    # GPU refit of the tree to its vertices (the animated-geometry path; with
    # the build's own vertices the tree is unchanged), HIP events on the stream
    from tfg_amd import _lib

    refitter = _lib.Refitter(ds.host)
    d_verts = torch.from_numpy(np.ascontiguousarray(sc.verts, np.float32).reshape(-1, 9)).cuda()
    refitter.refit(ds.bvh, d_verts, stream=stream.cuda_stream)
    ra, rb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ra.record(stream)
    for _ in range(5):
        refitter.refit(ds.bvh, d_verts, stream=stream.cuda_stream)
    rb.record(stream)
    torch.cuda.synchronize()
    refit_ms = ra.elapsed_time(rb) / 5
    refitter.close()
    # launches per frame and tile: one per depth, except that connect_batch
    # k > 0 traces the shadow rays of depths k..MD-1 in one launch
    cb = shape.get("connect_batch", 0)
    launches = {"extend": MD * shape.get("tiles", 1),
                "connect": (MD if cb == 0 or cb >= MD else cb + 1) * shape.get("tiles", 1)}
    serial = {k: _stage_roofline(rprof, n_roof, k, launches[k]) for k in launches}
    in_sched = {k: _stage_roofline([prof], 1, k, launches[k]) for k in launches}
    dom = max(serial, key=lambda k: serial[k]["ms_per_frame"])
    d = serial[dom]
    names = {"extend": "k_extend (closest-hit stage)", "connect": "k_connect (shadow any-hit stage)"}
    traffic = MEASURED_TRAFFIC.get("k_" + dom)
    tr_launch = traffic[0] / launches[dom] if traffic else None
    ser_ms = {"generate": float(np.mean([p["ms_generate"] for p in rprof])),
              "extend": serial["extend"]["ms_per_frame"],
              "shade": float(np.mean([sum(p["ms_shade"]) for p in rprof])),
              "connect": serial["connect"]["ms_per_frame"],
              "finalize": float(np.mean([p["ms_finalize"] for p in rprof])),
              "frame": float(np.mean([p["ms_total"] for p in rprof]))}
    algo_frame = serial["extend"]["bytes_per_frame"] + serial["connect"]["bytes_per_frame"]
    res = {
        "value": value,
        "ms_per_step": ms_frame,
        "per_kernel_ms": {"generate": prof["ms_generate"], "extend": sum(prof["ms_extend"]),
                          "shade": sum(prof["ms_shade"]), "connect": sum(prof["ms_connect"]),
                          "finalize": prof["ms_finalize"], "profiled_frame": prof["ms_total"],
                          "note": "one frame of the timed schedule: stages on side streams overlap, so their "
                                  "sums can exceed the frame"},
        "per_kernel_ms_serial": ser_ms,
        "mrays_per_kernel": {k: serial[k]["rays_per_frame"] / (serial[k]["ms_per_frame"] * 1e-3) / 1e6
                             for k in serial},
        "roofline": {
            "kernel": names[dom],
            "bound": "hbm",
            "achieved": d["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["gbs"] / HBM_PEAK_GBS,
            "traffic": tr_launch,
            "traffic_per_frame": traffic[0] if traffic else None,
            "traffic_source": traffic[1] if traffic else None,
            "achieved_measured_traffic": tr_launch / (d["avg_launch_ms"] * 1e-3) / 1e9 if traffic else None,
            "frac_measured_traffic": tr_launch / (d["avg_launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
            "algorithmic_bytes_per_launch": int(d["bytes_per_launch"]),
            "launches_per_frame": launches[dom],
            "avg_launch_ms": d["avg_launch_ms"],
            "timing": f"HIP events around each launch on its stream, {n_roof} frames of the timed schedule with "
                      "overlap off (each launch alone on the GPU)",
            "nodes_per_ray": d["nodes_per_ray"], "tris_per_ray": d["tris_per_ray"],
            "stages_serial": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                              for k, v in serial.items()},
            "stages_in_timed_schedule": {k: {"avg_launch_ms": round(v["avg_launch_ms"], 4), "gbs": round(v["gbs"], 1)}
                                         for k, v in in_sched.items()},
            "frame_level": {"algorithmic_bytes_per_frame": int(algo_frame),
                            "gbs_over_timed_frame": algo_frame / (ms_frame * 1e-3) / 1e9,
                            "frac": algo_frame / (ms_frame * 1e-3) / 1e9 / HBM_PEAK_GBS},
        },
        "config": {
            "workload": "plain FP32 wavefront PT: 64 seeded instances (32 displaced spheres + 32 boxes, "
                        f"{sc.n_tris} tris) in a closed 12-triangle room, 4 point lights, Lambertian+GGX; 1920x1080, "
                        "1 spp/frame, max depth 8, RR after depth 3; per-frame hipGraph replay, schedule in launch",
            "tris": sc.n_tris,
            "cwbvh_nodes": ds.host.num_nodes,
            "rays_per_frame": rays_frame,
            "ext_rays_per_depth": prof["ext_rays"],
            "shadow_rays_per_depth": prof["shadow_rays"],
            "ms_per_depth_serial": {
                "extend": [round(float(np.mean([p["ms_extend"][i] for p in rprof])), 4) for i in range(MD)],
                "shade": [round(float(np.mean([p["ms_shade"][i] for p in rprof])), 4) for i in range(MD)],
                "connect": [round(float(np.mean([p["ms_connect"][i] for p in rprof])), 4) for i in range(MD)]},
            "launch": timed_shape,
            "schedule_source": tuned,
            "cwbvh_build_s": round(t_build, 2),
            "cwbvh_refit_ms": round(refit_ms, 3),
            "seeds": {"scene": sc.seed, "frames": [1000, 1000 + args.steps - 1]},
        },
        "graph_replay": graph_ab,
        "wall_s": wall,
        "dtype": "fp32",
        "_cpu_inputs": (sc, ds.host, W, H, MD, RR),
    }
    return res


def _graph_vs_direct(torch, direct, graph_on, ds, sc, W, H, MD, RR, img, accum, stream, steps):
    """ms per frame of per-frame hipGraph replay and of direct launches over
    the same `steps` frames, in the direct tracer's schedule (overlap on:
    graph_on, the timed tracer, already captured with the same image and
    accumulation buffers) and with the shadow stage on the main stream
    (overlap off: a second graph tracer, given its schedule before its first
    frame, so that each graph tracer captures once)."""
    from tfg_amd import pt

    shape = direct.launch_shape()
    graph_off = pt.PathTracer(W, H, MD, RR, mode=0, use_graph=True)
    graph_off.set_scene(ds)
    graph_off.set_camera(sc.camera.frame(W, H))
    graph_off.set_schedule(**{k: v for k, v in shape.items() if k in graph_off.schedule() and k != "use_graph"})
    graph_off.set_schedule(overlap=False)
    out = {}
    for overlap in (True, False):
        ms = {}
        direct.set_schedule(overlap=overlap)
        for name, tr in (("direct", direct), ("graph", graph_on if overlap else graph_off)):
            for f in range(3):
                tr.render(f, img, accum, stream=stream.cuda_stream)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for k in range(steps):
                tr.render(1000 + k, img, accum, stream=stream.cuda_stream)
            b.record(stream)
            torch.cuda.synchronize()
            ms[name] = a.elapsed_time(b) / steps
        out["overlap_on" if overlap else "overlap_off"] = {
            "direct_ms_per_frame": round(ms["direct"], 4), "graph_ms_per_frame": round(ms["graph"], 4),
            "graph_over_direct": round(ms["graph"] / ms["direct"], 4)}
    direct.set_schedule(overlap=shape["overlap"])
    torch.cuda.synchronize()
    graph_off.close()
    _log(f"graph replay vs direct: {out}")
    return out


def bench_pt1080(args, torch):
    """The metric's configuration: the plain FP32 wavefront PT of bench_pt at
    1920x1080 (value = its Mrays/s), then the equal-time comparison of the
    cache-accelerated tracer (BF16 MLP, 48 warmup frames counted in its time)
    against the plain one, both measured by relMSE against a plain reference
    accumulated in FP64; the pair of runs is repeated --repeats times with
    distinct frame seeds and the ratio pooled over them."""
    from tfg_amd import experiments, scenes

    res = bench_pt(args, torch)
    sc = res["_cpu_inputs"][0]
    W, H, MD, RR = 1920, 1080, 8, 3
    t0 = time.time()
    # the timed schedule on direct launches: per-frame graph replay and direct
    # launches take the same time (graph_replay), the cache frames then train
    # one frame behind beside the next frame's tracing (defer_training), and
    # the runs create, re-mode and destroy a dozen tracers without capturing
    # and dropping graphs each time
    eq_sched = dict(res["config"]["launch"], use_graph=False)
    cmp = experiments.equal_time(torch, sc, W, H, MD, RR, ref_spp=args.ref_spp, budget_s=args.budget,
                                 warmup=48, precision=0, schedule=eq_sched, repeats=args.repeats)
    _log(f"equal-time comparison: {json.dumps(cmp)} ({time.time() - t0:.1f}s)")
    res["relmse"] = cmp
    t0 = time.time()
    res["nrc_kernels"] = bench_nrc_kernels(torch, sc, res["config"]["launch"])
    _log(f"cache kernels: {json.dumps(res['nrc_kernels'])} ({time.time() - t0:.1f}s)")
    res["config"]["workload"] = ("1080p path tracer (64 instances, 2.57M tris, in a closed room, 4 lights, "
                                 "Lambertian+GGX, depth 8, RR after 3): plain FP32 wavefront PT Mrays/s; equal-time relMSE of "
                                 "NRC-BF16 vs plain against a plain FP64-accumulated reference")
    res["config"]["relmse_reference_spp"] = args.ref_spp
    res["config"]["equal_time_budget_s"] = args.budget
    res["config"]["equal_time_repeats"] = args.repeats
    res["config"]["cache_frame_tail_from"] = experiments.CACHE_TAIL_FROM
    res["config"]["equal_time_launch"] = "direct launches of the timed schedule; cache frames train deferred"
    return res


MFMA_BF16_PEAK_TFLOPS = 2500.0
NRC_INFER_FLOP = 41_344            # 2 x (5 x 64 x 64 + 64 x 3) per query
NRC_TRAIN_FLOP = 115_840 + 41_344  # per training sample: forward + dX + dW, plus the tail inference


def bench_nrc_kernels(torch, sc, schedule, reps=20):
    """The cache's two matrix-core kernels on the frame's own data: a cache
    tracer (BF16, 5% training paths) runs 48 training-only frames and 4
    train+infer frames; then the last frame's queries are inferred and
    composited (k_nrc_infer) `reps` times and 4 optimiser steps on its
    records (k_nrc_train + reduction + Adam) are run `reps` times, each timed
    with HIP events on the launch stream."""
    from tfg_amd import experiments, nrc, pt

    W, H, MD, RR = 1920, 1080, 8, 3
    ds = pt.DeviceScene(sc, torch)
    tr = pt.PathTracer(W, H, MD, RR, mode=experiments.pt_mode_train(), use_graph=False)
    tr.set_scene(ds)
    tr.set_camera(sc.camera.frame(W, H))
    tr.set_schedule(**{k: v for k, v in schedule.items() if k in tr.schedule() and k != "use_graph"})
    lo = sc.verts.reshape(-1, 3).min(0) - 0.5
    hi = sc.verts.reshape(-1, 3).max(0) + 0.5
    cache = nrc.RadianceCache(precision=0, bounds_min=tuple(map(float, lo)), bounds_max=tuple(map(float, hi)))
    tr.set_cache(cache)
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    for f in range(48):
        tr.render(f, img)
    tr.set_mode(experiments.pt_mode_train() | experiments.pt_mode_infer())
    tr.set_schedule(tail_from=experiments.CACHE_TAIL_FROM)
    for f in range(4):
        tr.render(48 + f, img)
    torch.cuda.synchronize()
    rec, qry = tr.read_cache_data()
    n_rec, n_q = rec.shape[0], qry.shape[0]
    r_ptr, q_ptr, c_ptr = tr.cache_buffers()
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    ms_inf = timed(lambda: cache.infer_queries(q_ptr, c_ptr + 4 * 4, img.data_ptr(), stream=sh))
    # the configuration's query count (~2M per frame): the frame's queries tiled
    # to 2^21 (their pixel ids stay inside the image)
    n_big = 1 << 21
    big = np.ascontiguousarray(np.resize(qry, (n_big, qry.shape[1])))
    d_big = torch.from_numpy(big).cuda()
    d_cnt = torch.tensor([n_big], dtype=torch.int32, device="cuda")
    ms_big = timed(lambda: cache.infer_queries(d_big.data_ptr(), d_cnt.data_ptr(), img.data_ptr(), stream=sh))
    big_tf = n_big * NRC_INFER_FLOP / (ms_big * 1e-3) / 1e12
    steps = int(cache.cfg.steps_per_frame)
    batch = int(cache.cfg.batch)
    ms_train = timed(lambda: cache.train_ptr(r_ptr, c_ptr + 4 * 3, steps=steps, stream=sh))
    inf_tf = n_q * NRC_INFER_FLOP / (ms_inf * 1e-3) / 1e12
    tr_tf = steps * batch * NRC_TRAIN_FLOP / (ms_train * 1e-3) / 1e12
    out = {
        "precision": "bf16 operands, fp32 accumulation",
        "inference": {"kernel": "k_nrc_infer (+ composite)", "queries": int(n_q), "ms": ms_inf,
                      "flop_per_query": NRC_INFER_FLOP, "achieved_tflops": inf_tf,
                      "mfma_frac": inf_tf / MFMA_BF16_PEAK_TFLOPS},
        "inference_2m": {"kernel": "k_nrc_infer (+ composite)", "queries": n_big, "ms": ms_big,
                         "achieved_tflops": big_tf, "mfma_frac": big_tf / MFMA_BF16_PEAK_TFLOPS,
                         "note": "the frame's queries tiled to 2^21"},
        "training": {"kernels": "k_nrc_train + k_nrc_grad_reduce + k_nrc_adam", "steps": steps, "batch": batch,
                     "records_available": int(n_rec), "ms": ms_train, "flop_per_sample": NRC_TRAIN_FLOP,
                     "achieved_tflops": tr_tf, "mfma_frac": tr_tf / MFMA_BF16_PEAK_TFLOPS},
        "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
    }
    tr.close()
    cache.close()
    return out


def bench_copy(args, torch):
    """1 GiB device-to-device copy; one step = one copy (1 GiB read + 1 GiB
    written)."""
    n = 1 << 28
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    stream = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    a.record(stream)
    for _ in range(args.steps):
        dst.copy_(src)
    b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    assert torch.equal(dst[-4:], src[-4:])
    ms = a.elapsed_time(b) / args.steps
    gbs = 2 * 4 * n / (ms * 1e-3) / 1e9
    return {"metric": "HBM copy bandwidth (byte-counter calibration)", "value": gbs, "unit": "GB/s",
            "ms_per_step": ms, "bytes_read": 4 * n, "bytes_written": 4 * n, "wall_s": wall}


def bench_c1(args, torch):
    """C1: the 12-quad room with a box and the level-4 icosphere (5,144
    triangles) and one point light, 256x256, 16 spp (16 frames, frame seed =
    base seed + sample), max depth 5, Russian roulette after depth 10 (never
    reached).  One step = one 16-spp image on the GPU; the CPU baseline runs
    the oracle integrator over the whole 16-spp image on all host threads
    and on one thread."""
    import oracle
    from tfg_amd import pt, scenes

    sc = scenes.cornell(level=4)
    ds = pt.DeviceScene(sc, torch)
    W, H, MD, RR, SPP, BASE = 256, 256, 5, 10, 16, 500
    tracer = pt.PathTracer(W, H, MD, RR, mode=0, use_graph=False)
    tracer.set_scene(ds)
    tracer.set_camera(sc.camera.frame(W, H))
    tracer.autotune()   # small frames: launch latency and the late depths' tail decide the schedule
    shape = tracer.launch_shape()
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    accum = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()

    def time_images(tr):
        for f in range(max(1, args.warmup)):
            tr.render(BASE + f, img, accum, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        tr.reset_totals()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for k in range(args.steps):
            for j in range(SPP):
                tr.render(BASE + j, img, accum, stream=stream.cuda_stream)
        b.record(stream)
        torch.cuda.synchronize()
        ext, shd = tr.ray_totals()
        return a.elapsed_time(b) / args.steps, (ext + shd) / args.steps

    # the tuned schedule, timed with direct launches and with per-frame graph
    # replay (at 65,536 paths a frame is ~50 short launches, so launch cost
    # matters); the faster is the step
    ms_direct, rays_image = time_images(tracer)
    tuned = {k: v for k, v in tracer.schedule().items() if k != "use_graph"}
    tracer.close()
    gtracer = pt.PathTracer(W, H, MD, RR, mode=0, use_graph=True)
    gtracer.set_scene(ds)
    gtracer.set_camera(sc.camera.frame(W, H))
    gtracer.set_schedule(**tuned)
    ms_graph, rays_graph = time_images(gtracer)
    torch.cuda.synchronize()
    gtracer.close()
    assert rays_graph == rays_image, (rays_graph, rays_image)
    use_graph = ms_graph < ms_direct
    ms_image = min(ms_direct, ms_graph)
    shape = dict(shape, use_graph=use_graph)
    cpu = None
    if not args.no_cpu:
        cores = _host_cores()
        R = oracle.Renderer(sc, ds.host.nodes, ds.host.tris, W, H)
        t0 = time.perf_counter()
        _, rays = R.render(BASE, SPP, MD, RR, nthreads=cores)
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        _, rays1 = R.render(BASE, SPP, MD, RR, nthreads=1)
        dt1 = time.perf_counter() - t1
        cpu = {"value": float(rays.sum()) / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"the whole 16-spp image ({int(rays[0])} extension + {int(rays[1])} shadow rays) through "
                         f"the oracle integrator on {cores} threads, {dt:.2f} s",
               "single_thread": {"value": float(rays1.sum()) / dt1 / 1e6, "unit": "Mrays/s",
                                 "sample": f"the same image on 1 thread, {dt1:.2f} s"}}
    return {"metric": "C1 Mrays/s (Cornell room + L4 icosphere, 256x256, 16 spp, depth 5)",
            "value": rays_image / (ms_image * 1e-3) / 1e6, "unit": "Mrays/s", "ms_per_step": ms_image,
            "dtype": "fp32", "data": "synthetic (seeded procedural scene)", "cpu_baseline": cpu,
            "config": {"workload": "C1: 12-quad room + box + level-4 icosphere (5,144 tris), 1 point light, "
                                   "256x256, 16 spp per step (frame seeds 500..515), max depth 5, RR after 10; "
                                   "plain FP32 wavefront PT, schedule from tfg_pt_autotune in this run, the faster of "
                                   "direct launches and per-frame graph replay",
                       "rays_per_image": rays_image, "tris": sc.n_tris, "launch": shape,
                       "ms_per_image": {"direct": ms_direct, "graph_replay": ms_graph}}}


def _host_cores():
    """Threads for the CPU baseline: the cores this process may run on, capped
    by OMP_NUM_THREADS when that is set (the host's share for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline_pt(sc, host, W, H, MD, RR):
    """The oracle integrator on one full frame of the same workload (all host
    threads), plus a single-thread band of rows."""
    import oracle

    cores = _host_cores()
    R = oracle.Renderer(sc, host.nodes, host.tris, W, H)
    t0 = time.perf_counter()
    _, rays = R.render_rows(0, H, 1000, 1, MD, RR, nthreads=cores)
    dt = time.perf_counter() - t0
    y0, y1 = 500, 580
    t1 = time.perf_counter()
    _, rays1 = R.render_rows(y0, y1, 1000, 1, MD, RR, nthreads=1)
    dt1 = time.perf_counter() - t1
    return {
        "value": float(rays.sum()) / dt / 1e6,
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"one full 1080p frame of the same workload (frame seed 1000, {int(rays[0])} extension + "
                  f"{int(rays[1])} shadow rays) through the oracle integrator on {cores} threads, {dt:.2f} s",
        "single_thread": {"value": float(rays1.sum()) / dt1 / 1e6, "unit": "Mrays/s",
                          "sample": f"rows {y0}-{y1} of the same frame on 1 thread, {dt1:.2f} s"},
    }


def cpu_baseline_traversal(host, prim, bounce, shadow):
    """The oracle CWBVH traversal on the full ray sets (all host threads), plus
    a single-thread sample of 2^16 rays per set."""
    import oracle

    cores = _host_cores()

    def run(sets, nthreads):
        total = 0
        t0 = time.perf_counter()
        for rays, kind in sets:
            if kind == "closest":
                oracle.cwbvh_trace(host.nodes, host.tris, 0, rays, nthreads=nthreads)
            else:
                oracle.cwbvh_occluded(host.nodes, host.tris, 0, rays, nthreads=nthreads)
            total += rays.shape[0]
        return total, time.perf_counter() - t0

    full = ((prim, "closest"), (bounce, "closest"), (shadow, "any"))
    total, dt = run(full, cores)
    rng = np.random.default_rng(0)
    sample = 1 << 16
    sub = tuple((np.ascontiguousarray(r[rng.choice(r.shape[0], sample, replace=False)]), k) for r, k in full)
    total1, dt1 = run(sub, 1)
    return {
        "value": total / dt / 1e6,
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the full step (2^21 primary + 2^21 bounce closest-hit + 2^21 shadow any-hit rays), "
                  f"oracle CWBVH traversal on {cores} threads, {dt:.2f} s",
        "single_thread": {"value": total1 / dt1 / 1e6, "unit": "Mrays/s",
                          "sample": f"3 x 2^16 rays drawn from the same sets on 1 thread, {dt1:.2f} s"},
    }


def run_record(torch):
    """The machine the run measured, recorded with every measurement: host
    CPU model, logical CPUs and the cores this process
    may use, GPU model, compute units and memory, and the device-to-device
    copy bandwidth measured now (1 GiB read + 1 GiB written per copy, HIP
    events, best of 5 after 2 warmups)."""
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    prop = torch.cuda.get_device_properties(0)
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    best = None
    for _ in range(5):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return {
        "host_cpu": cpu,
        "host_logical_cpus": os.cpu_count(),
        "host_cores_usable": _host_cores(),
        "gpu": torch.cuda.get_device_name(0),
        "gpu_compute_units": prop.multi_processor_count,
        "gpu_memory_gib": round(prop.total_memory / 2 ** 30, 1),
        "gpu_arch": getattr(prop, "gcnArchName", ""),
        "copy_gbs_measured": round(2 * 4 * n / (best * 1e-3) / 1e9, 1),
        "torch": torch.__version__,
        "hip": getattr(torch.version, "hip", None),
    }


def _children():
    """Child processes of this process still alive (names), for the run record."""
    try:
        import psutil

        return [c.name() for c in psutil.Process().children(recursive=True)]
    except Exception as e:   # psutil missing or a process gone while listing
        return [f"unknown ({type(e).__name__})"]


def _pick_cache_tail(torch, render, tr, frame0, frames=3):
    """The fastest of tail_from 0, 2 and experiments.CACHE_TAIL_FROM for a
    tracer's train+infer frames, timed over `frames` frames each (outside any
    timed region): the tail launch wins where late depths hold few paths and
    the per-depth stages win where they hold many (a large canvas)."""
    from tfg_amd import experiments

    best = None
    for tail in sorted({0, 2, experiments.CACHE_TAIL_FROM}):
        tr.set_schedule(tail_from=tail)
        render(frame0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            render(frame0 + 1 + f)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, tail)
    tr.set_schedule(tail_from=best[1])
    return best[1]


def bench_multiview(args, torch):
    """16 seeded views (256x256 ... 1920x1080, scenes.views) of the 64-instance
    scene sharing one online-trained radiance cache (BF16), measured two ways:
      batched: all views on one canvas (tfg_pt_set_views), one chain of
        launches per frame; per-frame latency p50/p95/p99 (HIP events on the
        launch stream; a view's pixels are done when the frame is) and
        aggregate Mrays/s = the canvas's rays / the frames' time;
      per view: one tracer per view, all on one stream one after the other
        (the shared cache's training steps keep view order); each view's frame
        latency p50/p95/p99 and Mrays/s.
    Each tracer's train+infer frames run with the faster of tail_from 0 and
    experiments.CACHE_TAIL_FROM (_pick_cache_tail, before the timed frames);
    the canvas tracers, each the only user of its cache, defer their
    training to the next frame (schedule defer_training).
    Then BF16 vs FP16 cache inference on the batched canvas: a fresh cache per
    precision trained over the same frames, each view's image accumulated over
    the same frames and scored by relMSE against a plain reference of that
    view (FP64 accumulation)."""
    from tfg_amd import experiments, nrc, pt, scenes

    sc = scenes.instances()
    ds = pt.DeviceScene(sc, torch)
    views = scenes.views()
    frames_v = [(cam.frame(W, H), W, H) for cam, W, H in views]
    total = sum(W * H for _, W, H in views)
    bases = np.cumsum([0] + [W * H for _, W, H in views])
    MD, RR = 8, 3
    lo = sc.verts.reshape(-1, 3).min(0) - 0.5
    hi = sc.verts.reshape(-1, 3).max(0) + 0.5
    bounds = dict(bounds_min=tuple(map(float, lo)), bounds_max=tuple(map(float, hi)))
    train, both = experiments.pt_mode_train(), experiments.pt_mode_train() | experiments.pt_mode_infer()
    stream = torch.cuda.current_stream()
    warm, timed = 16, max(args.steps, 16)

    def canvas(precision):
        tr = pt.PathTracer(total, 1, MD, RR, mode=train, train_fraction=0.05, use_graph=False)
        tr.set_scene(ds)
        tr.set_views(frames_v)
        cache = nrc.RadianceCache(precision=precision, **bounds)
        tr.set_cache(cache)
        return tr, cache

    # batched: schedule tuned on plain frames of the canvas
    tr, cache = canvas(0)
    tr.set_mode(0)
    tr.autotune()
    sched = {k: v for k, v in tr.schedule().items() if k != "use_graph"}
    _log(f"multiview: canvas of {total} pixels, schedule {sched}")
    img = torch.zeros((total, 3), dtype=torch.float32, device="cuda")
    tr.set_mode(train)
    for f in range(warm):
        tr.render(f, img, stream=stream.cuda_stream)
    tr.set_mode(both)
    tr.set_schedule(defer_training=True)   # one tracer owns this cache
    tail_b = _pick_cache_tail(torch, lambda f: tr.render(f, img, stream=stream.cuda_stream), tr, 500_000)
    for f in range(2):
        tr.render(warm + f, img, stream=stream.cuda_stream)
    tr.reset_totals()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(timed)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f, (a, b) in enumerate(ev):
        a.record(stream)
        tr.render(warm + 2 + f, img, stream=stream.cuda_stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat_b = np.array([a.elapsed_time(b) for a, b in ev])
    agg = float(sum(tr.ray_totals()) / wall / 1e6)
    tr.close()
    del cache
    batched = {"mrays_per_s": round(agg, 1), "ms_per_frame": round(wall / timed * 1e3, 4), "tail_from": tail_b,
               **{f"p{q}_ms": round(float(np.percentile(lat_b, q)), 4) for q in (50, 95, 99)}}
    _log(f"multiview batched: {batched}")

    # per view: one tracer per view on one stream, the same schedule
    cache = nrc.RadianceCache(precision=0, **bounds)
    trs, imgs = [], []
    for cam, W, H in frames_v:
        t = pt.PathTracer(W, H, MD, RR, mode=train, train_fraction=0.05, use_graph=False)
        t.set_scene(ds)
        t.set_schedule(**sched)
        t.set_camera(cam)
        t.set_cache(cache)
        trs.append(t)
        imgs.append(torch.zeros((H * W, 3), dtype=torch.float32, device="cuda"))
    for r in range(warm + 2):
        if r == warm:
            for t, im in zip(trs, imgs):
                t.set_mode(both)
                _pick_cache_tail(torch, lambda f: t.render(f, im, stream=stream.cuda_stream), t, 500_000)
        for t, im in zip(trs, imgs):
            t.render(r, im, stream=stream.cuda_stream)
    for t in trs:
        t.reset_totals()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in views]
          for _ in range(timed)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(timed):
        for v, (t, im) in enumerate(zip(trs, imgs)):
            ev[r][v][0].record(stream)
            t.render(warm + 2 + r, im, stream=stream.cuda_stream)
            ev[r][v][1].record(stream)
    torch.cuda.synchronize()
    wall_v = time.perf_counter() - t0
    lat = np.array([[a.elapsed_time(b) for a, b in row] for row in ev])
    rays = np.array([sum(t.ray_totals()) for t in trs], np.float64)
    per_view = []
    for v, (_, W, H) in enumerate(views):
        per_view.append({"view": v, "width": W, "height": H, "tail_from": trs[v].schedule()["tail_from"],
                         **{f"p{q}_ms": round(float(np.percentile(lat[:, v], q)), 4) for q in (50, 95, 99)},
                         "mrays_per_s": round(float(rays[v] / timed / (np.median(lat[:, v]) * 1e-3) / 1e6), 1)})
    sequential = {"mrays_per_s": round(float(rays.sum() / wall_v / 1e6), 1),
                  "ms_per_round": round(wall_v / timed * 1e3, 4)}
    _log(f"multiview per view: {sequential}")
    for t in trs:
        t.close()
    del cache

    # BF16 vs FP16 inference on the batched canvas
    ref_spp, acc_frames = min(args.ref_spp, 512), 32
    tp = pt.PathTracer(total, 1, MD, RR, mode=0, use_graph=False)
    tp.set_scene(ds)
    tp.set_schedule(**sched)
    tp.set_views(frames_v)
    acc = torch.zeros((total, 3), dtype=torch.float64, device="cuda")
    for k in range(ref_spp):
        tp.render(1_000_000 + k, img, acc)
    torch.cuda.synchronize()
    ref = (acc / ref_spp).cpu().numpy()
    tp.close()
    rel = {}
    for name, prec in (("bf16", 0), ("fp16", 1)):
        tr, cache = canvas(prec)
        tr.set_schedule(**{**sched, "defer_training": True})
        for f in range(warm):
            tr.render(f, img)
        tr.set_mode(both)
        tr.set_schedule(tail_from=tail_b)
        acc.zero_()
        for f in range(acc_frames):
            tr.render(warm + f, img, acc)
        torch.cuda.synchronize()
        mean = (acc / acc_frames).cpu().numpy()
        rel[name] = [experiments.relmse(mean[bases[v]:bases[v + 1]], ref[bases[v]:bases[v + 1]])
                     for v in range(len(views))]
        tr.close()
        del cache
    for v in range(len(views)):
        per_view[v]["relmse_bf16"] = rel["bf16"][v]
        per_view[v]["relmse_fp16"] = rel["fp16"][v]
    return {
        "metric": "C5 batched multi-view: aggregate Mrays/s of 16 views on one canvas sharing one online-trained "
                  "cache",
        "value": batched["mrays_per_s"],
        "unit": "Mrays/s",
        "ms_per_step": batched["ms_per_frame"],
        "batched": batched,
        "per_view_tracers": sequential,
        "views": per_view,
        "relmse_mean": {k: float(np.mean(v)) for k, v in rel.items()},
        "relmse_fp16_over_bf16_mean": float(np.mean(np.array(rel["fp16"]) / np.array(rel["bf16"]))),
        "dtype": "fp32 tracing, bf16 cache (fp16 compared)",
        "data": "synthetic (seeded scene and camera poses)",
        "config": {"workload": "16 views (256x256 ... 1920x1080, scenes.views seeds) of the 64-instance scene "
                               "(2.57M tris, 4 lights), depth 8, RR after 3, one shared cache, train+infer frames",
                   "canvas_pixels": int(total), "schedule": sched, "frames": timed, "warmup_frames": warm,
                   "relmse_reference_spp": ref_spp, "relmse_frames": acc_frames},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pt1080", choices=["pt1080", "pt", "traversal", "copy", "nrc", "multiview", "sweep", "c1"])
    ap.add_argument("--ref-spp", type=int, default=4096, help="reference spp of the relMSE comparison")
    ap.add_argument("--budget", type=float, default=2.0, help="equal-time budget in seconds per method")
    ap.add_argument("--repeats", type=int, default=5,
                    help="equal-time runs per method (distinct frame seeds, fresh cache); the ratio is pooled")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="re-tune the path tracer's schedule (tfg_pt_autotune) instead of the saved one")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (frames would shard by pixels or views, "
              "the cache by one gradient all-reduce per optimiser step)", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.workload == "nrc":
        from tfg_amd import scenes

        sched = {"trace_block": 256, "grid_mult": 4, "tiles": 1, "side_streams": 2, "overlap": True,
                 "xcd_group": 4096}
        res = bench_nrc_kernels(torch, scenes.instances(), sched, reps=args.steps)
        print(json.dumps({"metric": "cache kernel throughput (MFMA)", **res}), flush=True)
        return
    if args.workload == "c1":
        res = bench_c1(args, torch)
        res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None})
        print(json.dumps(res), flush=True)
        return
    if args.workload == "multiview":
        res = bench_multiview(args, torch)
        res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None})
        print(json.dumps(res), flush=True)
        return
    if args.workload == "sweep":
        from tfg_amd import experiments, scenes

        t0 = time.time()
        res = experiments.grid_search(torch, scenes.instances(), 1920, 1080, 8, 3, frames=100, score_from=10)
        print(json.dumps({"metric": "radiance-cache hyper-parameter sweep (mean training loss, frames 10-99)",
                          "best": res[0], "results": res, "grid": experiments.SWEEP_GRID,
                          "config": {"workload": "64-instance scene, 1920x1080, training-only cache frames",
                                     "frames": 100, "seconds": round(time.time() - t0, 1)}}), flush=True)
        return
    if args.workload == "copy":
        res = bench_copy(args, torch)
        res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True})
        print(json.dumps(res), flush=True)
        return
    record = run_record(torch)
    _log(f"run record: {json.dumps(record)}")
    if args.workload in ("pt", "pt1080"):
        res = bench_pt1080(args, torch) if args.workload == "pt1080" else bench_pt(args, torch)
        cpu_in = res.pop("_cpu_inputs")
        cpu = None if args.no_cpu else cpu_baseline_pt(*cpu_in)
    else:
        res = bench_traversal(args, torch)
        cpu_in = res.pop("_cpu_inputs")
        cpu = None if args.no_cpu else cpu_baseline_traversal(*cpu_in)
    out = {
        "metric": METRIC,
        "value": round(res["value"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": res["dtype"],
        "data": "synthetic (seeded procedural scene and rays)",
        "config": res["config"],
        "per_kernel_ms": res["per_kernel_ms"],
        "mrays_per_kernel": res["mrays_per_kernel"],
        "per_kernel_ms_serial": res.get("per_kernel_ms_serial"),
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
    }
    if "graph_replay" in res:
        out["graph_replay"] = res["graph_replay"]
    if "relmse" in res:
        out["relmse"] = res["relmse"]
    if "nrc_kernels" in res:
        out["nrc_kernels"] = res["nrc_kernels"]
    record["child_processes_at_end"] = _children()
    out["run"] = record
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
