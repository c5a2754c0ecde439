// pt.hip - host side of the wavefront path tracer (tfg_pt_* C ABI).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <vector>

#include "host_common.h"
#include "nrc_iface.h"
#include "pt_kernels.h"

namespace tfg {
uint32_t known_tree_depth(const void* d_nodes);
uint32_t known_tree_format(const void* d_nodes);
uint32_t walk_tree_depth(const void* d_nodes, uint32_t num_nodes, int* rc);
}

struct tfg_pt {
    tfg_pt_config cfg{};
    tfg::PtParams P{};
    std::vector<void*> allocs;
    bool scene_set = false;
    bool camera_set = false;
    int n_cu = 0;
    // launch shapes
    int trace_block = 128;
    int trace_mult = 1;
    int trace_grid = 0;
    int refill_grid = 0;   // resident grid of the lane-refill closest-hit kernel
    int refill_cap = 0;    // at most this many of its workgroups per CU (0: as many as fit)
    int connect_per_cu = 0;   // shadow-stage workgroups per CU (0: the closest-hit grid)
    uint32_t connect_batch = 0;   // depths >= this (if > 0) share one shadow-stage launch
    uint32_t top_nodes = 9;       // leading tree nodes each traversal workgroup stages in LDS
    uint32_t tail_from = 0;       // depths >= this (if > 0) run in the one-launch tail kernel
    uint32_t scene_nodes = 0;     // nodes of the current scene's tree
    // watertight copy of an affine (tinybvh-layout) scene's triangle records
    float4* wt_tris = nullptr;
    size_t wt_cap = 0;            // records
    // pixel tiles: each runs its own chain of stages on its own streams
    int tiles = 1;
    tfg::PtParams TP[tfg::kMaxTiles]{};
    hipStream_t tstream[tfg::kMaxTiles] = {};   // generate, extend, shade, finalize of a tile
    hipStream_t tside[tfg::kMaxTiles][tfg::kMaxSide] = {};   // the tile's shadow-ray stage (connect)
    // Streams used only for graph capture, never for direct launches: a stream
    // that took part in a capture is not reused for direct work (the graphs
    // are replayed on tstream / stream).
    hipStream_t cap_tile[tfg::kMaxTiles] = {};
    hipStream_t cap_side[tfg::kMaxTiles][tfg::kMaxSide] = {};
    hipStream_t cap_post = nullptr;
    int sides = 1;                               // side streams per tile: connect(d) runs on side d % sides
    hipEvent_t ev_shaded[tfg::kMaxTiles][tfg::kMaxDepth] = {};   // shade(d) of a tile done
    hipEvent_t ev_cjoin[tfg::kMaxTiles][tfg::kMaxSide] = {};   // every connect of a side stream done
    hipEvent_t ev_tdone[tfg::kMaxTiles] = {};   // a tile finalized
    hipEvent_t ev_start = nullptr;              // frame counters reset
    bool overlap = true;
    bool run_graph = false;   // replay captured graphs (cfg.use_graph, fixed at creation)
    // graphs: one per tile (its stream chain and side stream) and one for the
    // cache work after the tiles (single stream)
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t graph_tile[tfg::kMaxTiles] = {};
    hipGraphExec_t exec_tile[tfg::kMaxTiles] = {};
    hipGraph_t graph_post = nullptr;
    hipGraphExec_t exec_post = nullptr;
    bool graphs_ready = false;
    const void* key_image = nullptr;
    const void* key_accum = nullptr;
    const void* key_nrc = nullptr;
    uint32_t key_mode = 0;
    // radiance cache attached to the frame (nrc.hip)
    tfg_nrc* nrc = nullptr;
    tfg_nrc_frame nrc_frame{};
    bool allow_no_cache = false;   // drop-in export frames: queries leave the frame uncomposited
    // Deferred training (schedule defer_training, direct launches only): a
    // frame's optimiser steps run at the start of the next frame on cstream,
    // beside its tracing, on the samples they draw, prepared with their
    // targets at the end of the frame; the next cache inference waits for
    // them, so every inference sees the weights it would see with the steps
    // at the end of their own frame.
    bool defer_train = false;
    bool train_pending = false;   // prepared samples wait for their steps
    bool wait_trained = false;    // the next cache inference waits for ev_trained
    float4* train_buf = nullptr;
    uint32_t* train_nrec = nullptr;
    size_t train_cap = 0;         // float4 entries of train_buf
    hipStream_t cstream = nullptr;
    hipEvent_t ev_gathered = nullptr, ev_trained = nullptr;
};

namespace tfg {
namespace {

template <typename T>
int dalloc(tfg_pt* pt, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = count * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(path tracer state)");
    pt->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return TFG_OK;
}

// Destroys the captured graphs.  Every entry point that calls this may follow
// an asynchronous tfg_pt_render, so replays still in flight (tile graphs on
// the tile streams, the cache graph on the tracer's stream) finish first.
void drop_graph(tfg_pt* pt) {
    bool any = pt->exec_post != nullptr;
    for (uint32_t t = 0; t < kMaxTiles; ++t) any = any || pt->exec_tile[t] != nullptr;
    if (any) {
        for (uint32_t t = 0; t < kMaxTiles; ++t)
            if (pt->tstream[t]) (void)hipStreamSynchronize(pt->tstream[t]);
        if (pt->stream) (void)hipStreamSynchronize(pt->stream);
    }
    for (uint32_t t = 0; t < kMaxTiles; ++t) {
        if (pt->exec_tile[t]) (void)hipGraphExecDestroy(pt->exec_tile[t]);
        if (pt->graph_tile[t]) (void)hipGraphDestroy(pt->graph_tile[t]);
        pt->exec_tile[t] = nullptr;
        pt->graph_tile[t] = nullptr;
    }
    if (pt->exec_post) (void)hipGraphExecDestroy(pt->exec_post);
    if (pt->graph_post) (void)hipGraphDestroy(pt->graph_post);
    pt->exec_post = nullptr;
    pt->graph_post = nullptr;
    pt->graphs_ready = false;
}

constexpr uint32_t kTopNodes = 73;   // at most root + 8 + 64 nodes staged in LDS

size_t trace_lds(const tfg_pt* pt, int block) {
    return (size_t)pt->P.n_top * 80 + (size_t)pt->P.stack_depth * block * 8;
}

// Grid of the per-path kernels (generate, shade, finalize, accumulate;
// 256-thread grid-stride loops) for n paths: one path per thread, between 8
// and 64 workgroups per CU.  Memory latency bounds these kernels, and more
// resident waves hide it better than a loop over several paths per thread
// (1080p shading stage: 1% shorter frame than 8 workgroups per CU).
uint32_t path_grid(const tfg_pt* pt, uint32_t n) {
    const uint32_t lo = (uint32_t)pt->n_cu * 8u, hi = (uint32_t)pt->n_cu * 64u;
    const uint32_t g = (n + 255u) / 256u;
    return g < lo ? lo : (g > hi ? hi : g);
}

template <int B>
int occupancy_extend(const tfg_pt* pt, int* blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_extend<B>, B, trace_lds(pt, B)) == hipSuccess ? 0 : -1;
}

template <int B>
int occupancy_refill(const tfg_pt* pt, int* blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_extend_refill<B>, B, trace_lds(pt, B)) == hipSuccess
               ? 0
               : -1;
}

int set_trace_shape(tfg_pt* pt, int block, int mult) {
    int per_cu = 0, refill_per_cu = 0;
    if (block == 64) occupancy_extend<64>(pt, &per_cu), occupancy_refill<64>(pt, &refill_per_cu);
    else if (block == 128) occupancy_extend<128>(pt, &per_cu), occupancy_refill<128>(pt, &refill_per_cu);
    else occupancy_extend<256>(pt, &per_cu), occupancy_refill<256>(pt, &refill_per_cu);
    if (per_cu <= 0) per_cu = 1;
    if (refill_per_cu <= 0) refill_per_cu = 1;
    if (pt->refill_cap > 0 && refill_per_cu > pt->refill_cap) refill_per_cu = pt->refill_cap;
    pt->trace_block = block;
    pt->trace_mult = mult;
    pt->trace_grid = pt->n_cu * per_cu * mult;
    // the refill kernel's waves stay for the whole launch: one resident grid
    pt->refill_grid = pt->n_cu * refill_per_cu;
    return TFG_OK;
}

template <bool kStats>
void launch_extend(tfg_pt* pt, hipStream_t s, const PtParams& T, uint32_t q, uint32_t d) {
    const size_t lds = trace_lds(pt, pt->trace_block);
    if (T.refill_from != 0 && d >= T.refill_from) {
        const dim3 g(pt->refill_grid);
        switch (pt->trace_block) {
            case 64: hipLaunchKernelGGL((k_extend_refill<64, kStats>), g, dim3(64), lds, s, T, q, d); break;
            case 256: hipLaunchKernelGGL((k_extend_refill<256, kStats>), g, dim3(256), lds, s, T, q, d); break;
            default: hipLaunchKernelGGL((k_extend_refill<128, kStats>), g, dim3(128), lds, s, T, q, d); break;
        }
        return;
    }
    const dim3 g(pt->trace_grid);
    switch (pt->trace_block) {
        case 64: hipLaunchKernelGGL((k_extend<64, kStats>), g, dim3(64), lds, s, T, q, d); break;
        case 256: hipLaunchKernelGGL((k_extend<256, kStats>), g, dim3(256), lds, s, T, q, d); break;
        default: hipLaunchKernelGGL((k_extend<128, kStats>), g, dim3(128), lds, s, T, q, d); break;
    }
}

template <int B, bool kStats>
void launch_tail_b(tfg_pt* pt, hipStream_t s, const PtParams& T, uint32_t q, uint32_t d0) {
    const size_t lds = trace_lds(pt, B) + (size_t)pt->P.stack_depth * (B / 8) * 8;
    if (pt->P.mode != 0) hipLaunchKernelGGL((k_tail<B, true, kStats>), dim3(pt->trace_grid), dim3(B), lds, s, T, q, d0);
    else hipLaunchKernelGGL((k_tail<B, false, kStats>), dim3(pt->trace_grid), dim3(B), lds, s, T, q, d0);
}

template <bool kStats>
void launch_tail(tfg_pt* pt, hipStream_t s, const PtParams& T, uint32_t q, uint32_t d0) {
    switch (pt->trace_block) {
        case 64: launch_tail_b<64, kStats>(pt, s, T, q, d0); break;
        case 256: launch_tail_b<256, kStats>(pt, s, T, q, d0); break;
        default: launch_tail_b<128, kStats>(pt, s, T, q, d0); break;
    }
}

template <bool kStats>
void launch_connect(tfg_pt* pt, hipStream_t s, const PtParams& T, uint32_t d_lo, uint32_t d_hi) {
    // a bounded persistent grid leaves CU slots to the closest-hit chain
    const dim3 g(pt->connect_per_cu > 0 ? pt->n_cu * pt->connect_per_cu : pt->trace_grid);
    const size_t lds = trace_lds(pt, pt->trace_block);
    switch (pt->trace_block) {
        case 64: hipLaunchKernelGGL((k_connect<64, kStats>), g, dim3(64), lds, s, T, d_lo, d_hi); break;
        case 256: hipLaunchKernelGGL((k_connect<256, kStats>), g, dim3(256), lds, s, T, d_lo, d_hi); break;
        default: hipLaunchKernelGGL((k_connect<128, kStats>), g, dim3(128), lds, s, T, d_lo, d_hi); break;
    }
}

void launch_sort(hipStream_t s, const SortJob& J) {
    hipLaunchKernelGGL(k_sort_count, dim3(kSortBlocks), dim3(kSortThreads), 0, s, J);
    hipLaunchKernelGGL(k_sort_scan_bins, dim3(kSortBins / kSortTileBins), dim3(256), 0, s, J);
    hipLaunchKernelGGL(k_sort_scatter, dim3(kSortBlocks), dim3(kSortThreads), 0, s, J);
}

// The counting sort of depth d's shadow entries (before its shadow stage).
void launch_shadow_sort(hipStream_t s, const PtParams& T, uint32_t d) {
    const size_t o = (size_t)d * T.depth_stride;
    launch_sort(s, SortJob{T.skey + o, T.counters + CNT_SHQ + d, T.sperm + o, T.sort_scr + kSortScr * d,
                           T.n_paths >> 4});
}

// The counting sort of depth d's extension rays (before its closest-hit stage).
void launch_ray_sort(hipStream_t s, const PtParams& T, uint32_t d) {
    launch_sort(s, SortJob{T.rkey, T.counters + CNT_QD + d, T.rperm, T.ray_scr, T.n_paths >> 4});
}

// Watertight triangle records (v0, v1, v2, primIdx in T3.w) for the affine
// records of a tinybvh-layout tree: record k keeps its position (the nodes
// address it as triBase + 4k) and its primIdx (T3.w), and takes that
// triangle's vertices from verts (9 floats per primIdx).  A primIdx beyond
// the vertex array gives a zero-area record, which no ray hits.
__global__ __launch_bounds__(256) void k_watertight_records(const float4* __restrict__ aff, uint32_t n_rec,
                                                            const float* __restrict__ verts, uint32_t n_verts,
                                                            float4* __restrict__ out) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n_rec; k += gridDim.x * blockDim.x) {
        const float pw = aff[4 * (size_t)k + 3].w;
        const uint32_t prim = __float_as_uint(pw);
        float4 v[3] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f),
                       make_float4(0.f, 0.f, 0.f, 0.f)};
        if (prim < n_verts) {
            const float* p = verts + 9 * (size_t)prim;
            for (int j = 0; j < 3; ++j) v[j] = make_float4(p[3 * j], p[3 * j + 1], p[3 * j + 2], 0.0f);
        }
        out[4 * (size_t)k] = v[0];
        out[4 * (size_t)k + 1] = v[1];
        out[4 * (size_t)k + 2] = v[2];
        out[4 * (size_t)k + 3] = make_float4(0.f, 0.f, 0.f, pw);
    }
}

// Split the frame's paths into pt->tiles contiguous pixel ranges (multiples
// of 256 pixels) and derive each tile's parameters: slices of every per-path
// and queue array, its own counter block, the shared frame counter block and
// frame-wide cache I/O.
void make_tiles(tfg_pt* pt) {
    const PtParams& P = pt->P;
    const uint32_t n = P.n_paths;
    const uint32_t k = (uint32_t)pt->tiles;
    for (uint32_t t = 0; t < k; ++t) {
        const uint32_t b = t == 0 ? 0u : (uint32_t)(((uint64_t)n * t / k) & ~255ull);
        const uint32_t e = t + 1 == k ? n : (uint32_t)(((uint64_t)n * (t + 1) / k) & ~255ull);
        PtParams T = P;
        T.n_paths = e - b;
        T.pix_base = b;
        T.depth_stride = n;
        T.counters = P.counters + (size_t)CNT_ALLOC * (1 + t);
        T.gcounters = P.counters;
        T.rays[0] = P.rays[0] + 2 * (size_t)b;
        T.rays[1] = P.rays[1] + 2 * (size_t)b;
        T.qpath[0] = P.qpath[0] + b;
        T.qpath[1] = P.qpath[1] + b;
        T.hits = P.hits + b;
        T.shadow = P.shadow + 4 * (size_t)b;
        T.skey = P.skey + b;
        T.sperm = P.sperm + b;
        T.sort_scr = P.sort_scr + kSortScr * P.max_depth * t;
        T.rkey = P.rkey + b;
        T.rperm = P.rperm + b;
        T.ray_scr = P.ray_scr + kSortScr * t;
        T.rcur = P.rcur + kRefillCurWords * t;
        T.nee = P.nee + b;
        T.neemask = P.neemask + b;
        T.thr = P.thr + b;
        T.rad = P.rad + b;
        T.rng = P.rng + b;
        T.flags = P.flags + b;
        if (P.mode || P.foot) {
            T.foot = P.foot ? P.foot + b : nullptr;
            T.pre = P.pre ? P.pre + b : nullptr;
            T.tthr = P.tthr ? P.tthr + b : nullptr;
            T.head = P.head ? P.head + (size_t)kVertF4 * b : nullptr;
            T.tail = P.tail ? P.tail + (size_t)kVertF4 * b : nullptr;
            T.tsuf = P.tsuf ? P.tsuf + b : nullptr;
        }
        pt->TP[t] = T;
    }
}

// Per-stage events of a profiled frame (nullptr: plain frame): a start and
// an end event around each launch, recorded on the stream it runs on.
struct StageEvents {
    hipEvent_t frame[2] = {};
    hipEvent_t gen[kMaxTiles][2] = {}, fin[kMaxTiles][2] = {};
    hipEvent_t ext[kMaxTiles][kMaxDepth][2] = {}, shd[kMaxTiles][kMaxDepth][2] = {},
               con[kMaxTiles][kMaxDepth][2] = {}, tail[kMaxTiles][2] = {};
    int create(uint32_t tiles, uint32_t depth) {
        hipError_t e = hipSuccess;
        auto mk = [&](hipEvent_t* p) {
            if (e == hipSuccess) e = hipEventCreate(p);
        };
        for (int k = 0; k < 2; ++k) {
            mk(&frame[k]);
            for (uint32_t t = 0; t < tiles; ++t) {
                mk(&gen[t][k]);
                mk(&fin[t][k]);
                mk(&tail[t][k]);
                for (uint32_t d = 0; d < depth; ++d) {
                    mk(&ext[t][d][k]);
                    mk(&shd[t][d][k]);
                    mk(&con[t][d][k]);
                }
            }
        }
        return e == hipSuccess ? TFG_OK : hip_fail(e, "hipEventCreate(profile)");
    }
    ~StageEvents() {
        auto rm = [](hipEvent_t e) {
            if (e) (void)hipEventDestroy(e);
        };
        for (int k = 0; k < 2; ++k) {
            rm(frame[k]);
            for (uint32_t t = 0; t < kMaxTiles; ++t) {
                rm(gen[t][k]);
                rm(fin[t][k]);
                rm(tail[t][k]);
                for (uint32_t d = 0; d < kMaxDepth; ++d) {
                    rm(ext[t][d][k]);
                    rm(shd[t][d][k]);
                    rm(con[t][d][k]);
                }
            }
        }
    }
};

// depths traced by the per-depth stages; the tail kernel runs the rest
inline uint32_t staged_depths(const tfg_pt* pt) {
    return pt->tail_from > 0 && pt->tail_from < pt->cfg.max_depth ? pt->tail_from : pt->cfg.max_depth;
}

// whether depth d's shadow stage has a launch of its own (the first depth of a
// batch stands for the batch; depths of the tail kernel have none)
inline bool connect_launched(const tfg_pt* pt, uint32_t d) {
    return d < staged_depths(pt) && (pt->connect_batch == 0 || d <= pt->connect_batch);
}

inline void mark(hipEvent_t e, hipStream_t s) {
    if (e) (void)hipEventRecord(e, s);
}

// One tile's chain; every launch reads its work size from device counters.
// On the tile stream ts: generate, then extend(d) -> shade(d) per depth,
// then finalize.  The tile's connect(d) runs on its side stream after
// shade(d), overlapping extend/shade of the following depths (without
// overlap it runs on ts); ts waits for the last connect before finalize.
int enqueue_tile(tfg_pt* pt, int t, hipStream_t ts, hipStream_t* side, tfg_color* image, double* accum,
                 StageEvents* pe, bool stats) {
    const bool nrc = pt->P.mode != 0;
    const int sides = pt->overlap ? pt->sides : 0;
    const PtParams& T = pt->TP[t];
    mark(pe ? pe->gen[t][0] : nullptr, ts);
    const dim3 pg(path_grid(pt, T.n_paths));
    hipLaunchKernelGGL(k_generate, pg, dim3(256), 0, ts, T);
    mark(pe ? pe->gen[t][1] : nullptr, ts);
    const uint32_t staged_d = staged_depths(pt);
    for (uint32_t d = 0; d < staged_d; ++d) {
        const uint32_t q = d & 1u;
        if (T.ray_sort_from && d >= T.ray_sort_from) launch_ray_sort(ts, T, d);
        mark(pe ? pe->ext[t][d][0] : nullptr, ts);
        if (stats) launch_extend<true>(pt, ts, T, q, d);
        else launch_extend<false>(pt, ts, T, q, d);
        mark(pe ? pe->ext[t][d][1] : nullptr, ts);
        mark(pe ? pe->shd[t][d][0] : nullptr, ts);
        if (nrc) hipLaunchKernelGGL(k_shade<true>, pg, dim3(256), 0, ts, T, q, d);
        else hipLaunchKernelGGL(k_shade<false>, pg, dim3(256), 0, ts, T, q, d);
        mark(pe ? pe->shd[t][d][1] : nullptr, ts);
        // the shadow stage of depth d, or of depths batch..max_depth-1 at once
        const uint32_t batch = pt->connect_batch;
        const bool batched = batch > 0 && d >= batch;
        if (batched && d + 1 < staged_d) continue;
        const uint32_t d_lo = batched ? batch : d;
        hipStream_t cs = sides ? side[d_lo % sides] : ts;
        if (sides) {
            TFG_HIP_CHECK(hipEventRecord(pt->ev_shaded[t][d], ts));
            TFG_HIP_CHECK(hipStreamWaitEvent(cs, pt->ev_shaded[t][d], 0));
        }
        if (T.sort_from)
            for (uint32_t k = d_lo > T.sort_from ? d_lo : T.sort_from; k <= d; ++k) launch_shadow_sort(cs, T, k);
        mark(pe ? pe->con[t][d_lo][0] : nullptr, cs);
        if (stats) launch_connect<true>(pt, cs, T, d_lo, d);
        else launch_connect<false>(pt, cs, T, d_lo, d);
        mark(pe ? pe->con[t][d_lo][1] : nullptr, cs);
    }
    if (staged_d < pt->cfg.max_depth) {   // the remaining depths in one launch
        mark(pe ? pe->tail[t][0] : nullptr, ts);
        if (stats) launch_tail<true>(pt, ts, T, staged_d & 1u, staged_d);
        else launch_tail<false>(pt, ts, T, staged_d & 1u, staged_d);
        mark(pe ? pe->tail[t][1] : nullptr, ts);
    }
    for (int k = 0; k < sides && k < (int)staged_d; ++k) {
        TFG_HIP_CHECK(hipEventRecord(pt->ev_cjoin[t][k], side[k]));
        TFG_HIP_CHECK(hipStreamWaitEvent(ts, pt->ev_cjoin[t][k], 0));
    }
    mark(pe ? pe->fin[t][0] : nullptr, ts);
    if (nrc) hipLaunchKernelGGL(k_finalize<true>, pg, dim3(256), 0, ts, T, image, accum);
    else hipLaunchKernelGGL(k_finalize<false>, pg, dim3(256), 0, ts, T, image, accum);
    mark(pe ? pe->fin[t][1] : nullptr, ts);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

int ensure_train_buf(tfg_pt* pt) {
    const size_t need = nrc_gather_f4(pt->nrc);
    if (need <= pt->train_cap && pt->train_nrec) return TFG_OK;
    TFG_HIP_CHECK(hipDeviceSynchronize());   // steps still pending on the old buffer
    if (pt->train_buf) (void)hipFree(pt->train_buf);
    pt->train_buf = nullptr;
    pt->train_cap = 0;
    TFG_HIP_CHECK(hipMalloc((void**)&pt->train_buf, (need ? need : 1) * sizeof(float4)));
    if (!pt->train_nrec) TFG_HIP_CHECK(hipMalloc((void**)&pt->train_nrec, sizeof(uint32_t)));
    pt->train_cap = need;
    return TFG_OK;
}

// Enqueues the pending optimiser steps on q, after their prepared samples are
// complete; the next cache inference waits for them.
int run_pending_training(tfg_pt* pt, hipStream_t q) {
    if (!pt->train_pending || !pt->nrc) return TFG_OK;
    TFG_HIP_CHECK(hipStreamWaitEvent(q, pt->ev_gathered, 0));
    const int rc = nrc_train_gathered_enqueue(pt->nrc, pt->train_buf, pt->train_nrec, q);
    if (rc != TFG_OK) return rc;
    TFG_HIP_CHECK(hipEventRecord(pt->ev_trained, q));
    pt->train_pending = false;
    pt->wait_trained = true;
    return TFG_OK;
}

// After every tile: cache inference (composite) and training, then the
// accumulation of the composited image (cache modes only).  With deferred
// training (direct launches) the training is replaced by the copy of the
// records its steps will draw.
int enqueue_post(tfg_pt* pt, hipStream_t s, tfg_color* image, double* accum, bool direct) {
    if (pt->P.mode == 0) return TFG_OK;
    if (pt->nrc) {
        if (pt->wait_trained) {
            TFG_HIP_CHECK(hipStreamWaitEvent(s, pt->ev_trained, 0));
            pt->wait_trained = false;
        }
        const bool defer = direct && pt->defer_train && (pt->P.mode & MODE_TRAIN) && nrc_gather_f4(pt->nrc) > 0;
        tfg_nrc_frame f = pt->nrc_frame;
        if (defer) f.mode &= ~MODE_TRAIN;
        int rc = nrc_frame_enqueue(pt->nrc, f, image, s);
        if (rc == TFG_OK && defer) rc = ensure_train_buf(pt);
        if (rc == TFG_OK && defer) rc = nrc_gather_enqueue(pt->nrc, pt->nrc_frame, pt->train_buf, pt->train_nrec, s);
        if (rc != TFG_OK) return rc;
        if (defer) {
            TFG_HIP_CHECK(hipEventRecord(pt->ev_gathered, s));
            pt->train_pending = true;
        }
    }
    if (accum)
        hipLaunchKernelGGL(k_accumulate, dim3(path_grid(pt, pt->P.n_paths)), dim3(256), 0, s, image, accum,
                           pt->P.n_paths);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

// The stream of tile t's chain for a frame on s.  on_s (graph replay of a
// single tile): s itself, no fork and join, two cross-stream waits fewer per
// frame (1080p: 4.79 instead of 4.81 ms).  Direct launches keep the tile's
// own stream: issued on s, the chain shares a hardware queue with a side
// stream (1080p: 5.82 instead of 4.79 ms).
hipStream_t tile_stream(const tfg_pt* pt, hipStream_t s, int t, bool on_s) { return on_s ? s : pt->tstream[t]; }

// Frame prologue on s (reset of the frame counters, the frame index) and the
// fork of the tile streams: each waits for the prologue.
int fork_tiles(tfg_pt* pt, hipStream_t s, int32_t frame, bool on_s) {
    hipLaunchKernelGGL(k_frame_reset, dim3(1), dim3(128), 0, s, pt->P.counters, const_cast<int32_t*>(pt->P.frame),
                       frame);
    if (on_s) return TFG_OK;
    TFG_HIP_CHECK(hipEventRecord(pt->ev_start, s));
    for (int t = 0; t < pt->tiles; ++t) TFG_HIP_CHECK(hipStreamWaitEvent(pt->tstream[t], pt->ev_start, 0));
    return TFG_OK;
}

int join_tiles(tfg_pt* pt, hipStream_t s, bool on_s) {
    if (on_s) return TFG_OK;
    for (int t = 0; t < pt->tiles; ++t) {
        TFG_HIP_CHECK(hipEventRecord(pt->ev_tdone[t], pt->tstream[t]));
        TFG_HIP_CHECK(hipStreamWaitEvent(s, pt->ev_tdone[t], 0));
    }
    return TFG_OK;
}

// The whole frame, issued eagerly (no graph): prologue on s, the tiles on
// their streams, the cache work on s after every tile.  Tiles are
// independent, so the latency tails of their small late-depth launches
// overlap.
int enqueue_frame(tfg_pt* pt, hipStream_t s, int32_t frame, tfg_color* image, double* accum,
                  StageEvents* pe = nullptr, bool stats = false) {
    make_tiles(pt);
    mark(pe ? pe->frame[0] : nullptr, s);
    int rc = run_pending_training(pt, pt->cstream);   // the previous frame's steps, beside this frame's tracing
    if (rc == TFG_OK) rc = fork_tiles(pt, s, frame, false);
    for (int t = 0; t < pt->tiles && rc == TFG_OK; ++t)
        rc = enqueue_tile(pt, t, tile_stream(pt, s, t, false), pt->tside[t], image, accum, pe, stats);
    if (rc == TFG_OK) rc = join_tiles(pt, s, false);
    if (rc == TFG_OK) rc = enqueue_post(pt, s, image, accum, true);
    mark(pe ? pe->frame[1] : nullptr, s);
    return rc;
}

// Captures one graph per tile (origin: the tile's capture stream, plus its
// side capture streams) and one for the cache work (single stream), all on
// the capture-only streams.
int capture_graphs(tfg_pt* pt, tfg_color* image, double* accum) {
    make_tiles(pt);
    for (int t = 0; t < pt->tiles; ++t) {
        hipStream_t ts = pt->cap_tile[t];
        TFG_HIP_CHECK(hipStreamBeginCapture(ts, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_tile(pt, t, ts, pt->cap_side[t], image, accum, nullptr, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ts, &g);
        if (rc != TFG_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc != TFG_OK ? rc : hip_fail(e, "hipStreamEndCapture(tile)");
        }
        pt->graph_tile[t] = g;
        TFG_HIP_CHECK(hipGraphInstantiate(&pt->exec_tile[t], g, nullptr, nullptr, 0));
    }
    if (pt->P.mode != 0) {
        hipStream_t cs = pt->cap_post;
        TFG_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_post(pt, cs, image, accum, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc != TFG_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc != TFG_OK ? rc : hip_fail(e, "hipStreamEndCapture(cache)");
        }
        pt->graph_post = g;
        TFG_HIP_CHECK(hipGraphInstantiate(&pt->exec_post, g, nullptr, nullptr, 0));
    }
    pt->graphs_ready = true;
    return TFG_OK;
}

// The frame from the captured graphs: prologue on s, each tile graph on its
// tile stream (s itself for a single tile), the cache graph on s after every
// tile.
int launch_graphs(tfg_pt* pt, hipStream_t s, int32_t frame) {
    const bool on_s = pt->tiles == 1;
    int rc = fork_tiles(pt, s, frame, on_s);
    for (int t = 0; t < pt->tiles && rc == TFG_OK; ++t)
        TFG_HIP_CHECK(hipGraphLaunch(pt->exec_tile[t], tile_stream(pt, s, t, on_s)));
    if (rc == TFG_OK) rc = join_tiles(pt, s, on_s);
    if (rc == TFG_OK && pt->exec_post) TFG_HIP_CHECK(hipGraphLaunch(pt->exec_post, s));
    return rc;
}

// The HIP runtime maps streams onto a small pool of hardware queues (4 per
// process by default) that it creates as streams appear.  Measured on the
// 1080p workload, a tracer whose streams are the first ones the process
// creates runs the overlapped schedules 15-35% slower (2 side streams: 3.60
// ms per frame instead of 2.70) than any tracer created after some streams
// already existed (torch streams, or a tracer created and destroyed without
// rendering).  So, once per device and process, the pool is primed by
// creating and destroying as many streams as a tracer uses before the first
// tracer creates its own.
void prime_hw_queues() {
    static std::atomic<bool> primed[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || primed[dev].exchange(true)) return;
    constexpr int kPrime = 2 + kMaxTiles * (1 + kMaxSide);
    hipStream_t q[kPrime] = {};
    for (int i = 0; i < kPrime; ++i)
        if (hipStreamCreateWithFlags(&q[i], hipStreamNonBlocking) != hipSuccess) q[i] = nullptr;
    for (int i = 0; i < kPrime; ++i)
        if (q[i]) (void)hipStreamDestroy(q[i]);
}

}  // namespace
}  // namespace tfg

namespace tfg {
// Frame without cache inference even in INFER mode: the caller exports the
// captured queries (drop-in path).  Camera already in device memory.
int pt_render_export(tfg_pt* pt, int32_t frame, tfg_color* d_image, hipStream_t stream) {
    pt->camera_set = true;
    tfg_nrc* saved = pt->nrc;
    pt->nrc = nullptr;
    pt->allow_no_cache = true;
    const int rc = tfg_pt_render(pt, frame, d_image, nullptr, stream);
    pt->allow_no_cache = false;
    pt->nrc = saved;
    return rc;
}
const PtParams& pt_params(const tfg_pt* pt) { return pt->P; }
hipStream_t pt_stream(const tfg_pt* pt) { return pt->stream; }
}  // namespace tfg

using namespace tfg;

extern "C" {

int tfg_pt_create(const tfg_pt_config* cfg, tfg_pt** out) {
    if (!cfg || !out) { set_error("tfg_pt_create: NULL argument"); return TFG_E_INVALID; }
    *out = nullptr;
    if (cfg->width == 0 || cfg->height == 0 || (uint64_t)cfg->width * cfg->height > (1ull << 26)) {
        set_error("tfg_pt_create: bad resolution %ux%u", cfg->width, cfg->height);
        return TFG_E_INVALID;
    }
    if (cfg->max_depth == 0 || cfg->max_depth > kMaxDepth) {
        set_error("tfg_pt_create: max_depth must be 1..%u", kMaxDepth);
        return TFG_E_INVALID;
    }
    if (cfg->mode > 3 || !(cfg->train_fraction >= 0.0f && cfg->train_fraction <= 1.0f)) {
        set_error("tfg_pt_create: bad mode or train_fraction");
        return TFG_E_INVALID;
    }
    auto* pt = new (std::nothrow) tfg_pt();
    if (!pt) { set_error("tfg_pt_create: out of memory"); return TFG_E_NOMEM; }
    pt->cfg = *cfg;
    const uint32_t n = cfg->width * cfg->height;
    PtParams& P = pt->P;
    P.width = cfg->width;
    P.height = cfg->height;
    P.n_paths = n;
    P.pix_base = 0;
    P.depth_stride = n;
    P.xcd_group = 4096;
    P.octet_below = 1u << 18;
    P.sort_from = 1;   // shadow entries of depths >= 1 in spatial order (no result changes)
    P.refill_from = 1;   // lane refill in the closest-hit stage from depth 1 (no result changes)
    P.refill_idle = 32;
    P.max_depth = cfg->max_depth;
    P.rr_depth = cfg->rr_depth;
    P.mode = cfg->mode;
    P.train_frac = cfg->train_fraction;
    int rc = TFG_OK;
    float4* f4p = nullptr;
    uint32_t* u32p = nullptr;
#define ALLOC(field, T, count) \
    if (rc == TFG_OK) rc = dalloc<T>(pt, &P.field, (count));
    ALLOC(rays[0], float4, 2 * (size_t)n)
    ALLOC(rays[1], float4, 2 * (size_t)n)
    ALLOC(qpath[0], uint32_t, n)
    ALLOC(qpath[1], uint32_t, n)
    ALLOC(hits, float4, n)
    ALLOC(shadow, float4, 4 * (size_t)n * cfg->max_depth)
    ALLOC(skey, uint16_t, (size_t)n * cfg->max_depth)
    ALLOC(sperm, uint32_t, (size_t)n * cfg->max_depth)
    ALLOC(sort_scr, uint32_t, kSortScr * cfg->max_depth * kMaxTiles)
    ALLOC(rkey, uint16_t, n)
    ALLOC(rperm, uint32_t, n)
    ALLOC(ray_scr, uint32_t, kSortScr * kMaxTiles)
    ALLOC(rcur, uint32_t, kRefillCurWords * kMaxTiles)
    ALLOC(nee, float4, (size_t)n * cfg->max_depth)
    ALLOC(neemask, uint32_t, n)
    ALLOC(counters, uint32_t, (size_t)CNT_ALLOC * (1 + kMaxTiles))
    ALLOC(thr, float4, n)
    ALLOC(rad, float4, n)
    ALLOC(rng, uint32_t, n)
    ALLOC(flags, uint32_t, n)
    if (cfg->mode) {
        ALLOC(foot, float4, n)
        ALLOC(pre, float4, n)
        ALLOC(tthr, float4, n)
        ALLOC(head, float4, (size_t)kVertF4 * n)
        ALLOC(tail, float4, (size_t)kVertF4 * n)
        ALLOC(tsuf, float4, n)
        ALLOC(records, float4, (size_t)kRecF4 * n)
        ALLOC(queries, float4, (size_t)kQueryF4 * n)
    }
#undef ALLOC
    (void)f4p;
    (void)u32p;
    int32_t* dframe = nullptr;
    tfg_camera* dcam = nullptr;
    PtView* dviews = nullptr;
    if (rc == TFG_OK) rc = dalloc<int32_t>(pt, &dframe, 1);
    if (rc == TFG_OK) rc = dalloc<tfg_camera>(pt, &dcam, 1);
    if (rc == TFG_OK) rc = dalloc<PtView>(pt, &dviews, kMaxViews);
    P.frame = dframe;
    P.cam = dcam;
    P.views = dviews;
    P.n_views = 0;
    P.gcounters = P.counters;
    if (rc == TFG_OK) {
        int dev = 0;
        hipDeviceProp_t prop;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) prime_hw_queues();
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&pt->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_in, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_out, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_start, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&pt->cstream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_gathered, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_trained, hipEventDisableTiming);
        if (e == hipSuccess && cfg->use_graph) e = hipStreamCreateWithFlags(&pt->cap_post, hipStreamNonBlocking);
        for (uint32_t t = 0; t < kMaxTiles && e == hipSuccess && cfg->use_graph; ++t) {
            e = hipStreamCreateWithFlags(&pt->cap_tile[t], hipStreamNonBlocking);
            for (uint32_t k = 0; k < kMaxSide && e == hipSuccess; ++k)
                e = hipStreamCreateWithFlags(&pt->cap_side[t][k], hipStreamNonBlocking);
        }
        for (uint32_t t = 0; t < kMaxTiles && e == hipSuccess; ++t) {
            e = hipStreamCreateWithFlags(&pt->tstream[t], hipStreamNonBlocking);
            for (uint32_t k = 0; k < kMaxSide && e == hipSuccess; ++k) {
                e = hipStreamCreateWithFlags(&pt->tside[t][k], hipStreamNonBlocking);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_cjoin[t][k], hipEventDisableTiming);
            }
            if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_tdone[t], hipEventDisableTiming);
            for (uint32_t d = 0; d < cfg->max_depth && e == hipSuccess; ++d)
                e = hipEventCreateWithFlags(&pt->ev_shaded[t][d], hipEventDisableTiming);
        }
        if (e != hipSuccess) rc = hip_fail(e, "tfg_pt_create");
        else pt->n_cu = prop.multiProcessorCount;
    }
    if (rc == TFG_OK) {
        hipError_t e = hipMemset(P.counters, 0, (size_t)CNT_ALLOC * (1 + kMaxTiles) * sizeof(uint32_t));
        // the sorts' done counters start at zero (each sort leaves its own at zero)
        if (e == hipSuccess) e = hipMemset(P.sort_scr, 0, kSortScr * cfg->max_depth * kMaxTiles * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(P.ray_scr, 0, kSortScr * kMaxTiles * sizeof(uint32_t));
        if (e != hipSuccess) rc = hip_fail(e, "hipMemset(counters)");
    }
    if (rc != TFG_OK) {
        tfg_pt_destroy(pt);
        return rc;
    }
    set_trace_shape(pt, 64, 16);   // measured default (top_nodes 9): many short-lived workgroups
    pt->run_graph = cfg->use_graph != 0;
    *out = pt;
    return TFG_OK;
}

int tfg_pt_set_scene(tfg_pt* pt, const tfg_pt_scene* sc) {
    if (!pt || !sc) { set_error("tfg_pt_set_scene: NULL argument"); return TFG_E_INVALID; }
    if (!sc->bvh.d_nodos || !sc->bvh.d_tris || !sc->d_shading || !sc->d_materials ||
        (sc->n_lights && !sc->d_lights)) {
        set_error("tfg_pt_set_scene: missing scene buffer");
        return TFG_E_INVALID;
    }
    if (sc->n_lights > kMaxLights) {
        set_error("tfg_pt_set_scene: at most %u point lights", kMaxLights);
        return TFG_E_LIMIT;
    }
    if (sc->n_analytic > kMaxAnalytic || (sc->n_analytic && !sc->d_analytic)) {
        set_error("tfg_pt_set_scene: at most %u analytic primitives", kMaxAnalytic);
        return TFG_E_LIMIT;
    }
    if (sc->tri_format != TFG_TRI_WATERTIGHT && sc->tri_format != TFG_TRI_AFFINE) {
        set_error("tfg_pt_set_scene: unknown triangle format %u", sc->tri_format);
        return TFG_E_INVALID;
    }
    const uint32_t known_fmt = known_tree_format(sc->bvh.d_nodos);
    if (known_fmt != 0xFFFFFFFFu && known_fmt != sc->tri_format) {
        set_error("tfg_pt_set_scene: the tree's triangle records are in format %u, the scene says %u", known_fmt,
                  sc->tri_format);
        return TFG_E_INVALID;
    }
    if (sc->tri_format == TFG_TRI_AFFINE && (!sc->d_verts || sc->n_verts == 0)) {
        set_error("tfg_pt_set_scene: an affine tree needs its vertices (d_verts, n_verts)");
        return TFG_E_INVALID;
    }
    uint32_t depth = known_tree_depth(sc->bvh.d_nodos);
    if (depth == 0) {   // a tree this library did not build: walk it once
        int rc = TFG_OK;
        depth = walk_tree_depth(sc->bvh.d_nodos, sc->bvh.num_nodos, &rc);
        if (rc != TFG_OK) return rc;
        if (depth == 0) {
            set_error("tfg_pt_set_scene: malformed CWBVH (child index outside the %u nodes, or a node reached twice)",
                      sc->bvh.num_nodos);
            return TFG_E_INVALID;
        }
    }
    if (depth > TFG_MAX_TREE_DEPTH) {
        set_error("tfg_pt_set_scene: tree depth %u exceeds %d", depth, TFG_MAX_TREE_DEPTH);
        return TFG_E_LIMIT;
    }
    PtParams& P = pt->P;
    P.nodes = static_cast<const float4*>(sc->bvh.d_nodos);
    P.tris = static_cast<const float4*>(sc->bvh.d_tris);
    if (sc->tri_format == TFG_TRI_AFFINE) {
        const size_t n_rec = sc->bvh.num_tris;
        if (n_rec > pt->wt_cap) {
            if (pt->wt_tris) (void)hipFree(pt->wt_tris);
            pt->wt_tris = nullptr;
            pt->wt_cap = 0;
            TFG_HIP_CHECK(hipMalloc((void**)&pt->wt_tris, 4 * n_rec * sizeof(float4)));
            pt->wt_cap = n_rec;
        }
        TFG_HIP_CHECK(hipDeviceSynchronize());   // the caller's writes of the vertices
        const uint32_t grid = (uint32_t)std::min<size_t>((n_rec + 255) / 256, (size_t)pt->n_cu * 16);
        hipLaunchKernelGGL(k_watertight_records, dim3(grid ? grid : 1), dim3(256), 0, pt->stream,
                           static_cast<const float4*>(sc->bvh.d_tris), (uint32_t)n_rec, sc->d_verts, sc->n_verts,
                           pt->wt_tris);
        TFG_HIP_CHECK(hipGetLastError());
        TFG_HIP_CHECK(hipStreamSynchronize(pt->stream));
        P.tris = pt->wt_tris;
    }
    P.stack_depth = depth ? (int)(depth > 1 ? depth - 1 : 1) : TFG_MAX_TREE_DEPTH - 1;
    pt->scene_nodes = sc->bvh.num_nodos;
    P.n_top = pt->scene_nodes < pt->top_nodes ? pt->scene_nodes : pt->top_nodes;
    P.shading = sc->d_shading;
    P.mats = sc->d_materials;
    P.lights = sc->d_lights;
    P.n_lights = sc->n_lights;
    P.light_shift = 0;
    while ((1u << P.light_shift) < sc->n_lights) ++P.light_shift;
    P.analytic = sc->d_analytic;
    P.n_analytic = sc->n_analytic;
    pt->scene_set = true;
    drop_graph(pt);
    set_trace_shape(pt, pt->trace_block, pt->trace_mult);
    return TFG_OK;
}

int tfg_pt_set_camera(tfg_pt* pt, const tfg_camera* cam) {
    if (!pt || !cam) { set_error("tfg_pt_set_camera: NULL argument"); return TFG_E_INVALID; }
    TFG_HIP_CHECK(hipMemcpy(const_cast<tfg_camera*>(pt->P.cam), cam, sizeof(*cam), hipMemcpyHostToDevice));
    if (pt->P.n_views) drop_graph(pt);   // captured launches carry n_views
    pt->P.n_views = 0;
    pt->camera_set = true;
    return TFG_OK;
}

int tfg_pt_set_views(tfg_pt* pt, const tfg_pt_view* views, uint32_t n_views) {
    if (!pt || (!views && n_views)) { set_error("tfg_pt_set_views: NULL argument"); return TFG_E_INVALID; }
    if (n_views == 0 || n_views > kMaxViews) {
        set_error("tfg_pt_set_views: 1..%u views", kMaxViews);
        return TFG_E_LIMIT;
    }
    PtView hv[kMaxViews] = {};
    uint64_t total = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        const uint32_t w = views[v].width, h = views[v].height;
        if (w == 0 || h == 0 || w > 65535 || h > 65535) {
            set_error("tfg_pt_set_views: view %u is %ux%u (1..65535 each way)", v, w, h);
            return TFG_E_INVALID;
        }
        hv[v].base = (uint32_t)total;
        hv[v].width = w;
        hv[v].height = h;
        hv[v].cam = views[v].camera;
        total += (uint64_t)w * h;
    }
    if (total != pt->P.n_paths) {
        set_error("tfg_pt_set_views: the views hold %llu pixels, the canvas %u (tfg_pt_config width x height)",
                  (unsigned long long)total, pt->P.n_paths);
        return TFG_E_INVALID;
    }
    TFG_HIP_CHECK(hipMemcpy(const_cast<PtView*>(pt->P.views), hv, sizeof(PtView) * n_views, hipMemcpyHostToDevice));
    drop_graph(pt);   // captured launches carry n_views
    pt->P.n_views = n_views;
    pt->camera_set = true;
    return TFG_OK;
}


int tfg_pt_render(tfg_pt* pt, int32_t frame, tfg_color* d_image, double* d_accum, tfg_stream stream) {
    if (!pt || !d_image) { set_error("tfg_pt_render: NULL argument"); return TFG_E_INVALID; }
    if (!pt->scene_set || !pt->camera_set) { set_error("tfg_pt_render: scene or camera not set"); return TFG_E_STATE; }
    if ((pt->P.mode & MODE_INFER) && pt->nrc == nullptr && !pt->allow_no_cache) {
        set_error("tfg_pt_render: cache inference requested without an attached cache");
        return TFG_E_STATE;
    }
    hipStream_t s = pt->stream;
    TFG_HIP_CHECK(hipEventRecord(pt->ev_in, stream));
    TFG_HIP_CHECK(hipStreamWaitEvent(s, pt->ev_in, 0));
    int rc = TFG_OK;
    if (pt->run_graph) {
        if (pt->graphs_ready && (pt->key_image != d_image || pt->key_accum != d_accum || pt->key_nrc != pt->nrc ||
                                 pt->key_mode != pt->P.mode)) {
            drop_graph(pt);
        }
        if (!pt->graphs_ready) {
            rc = capture_graphs(pt, d_image, d_accum);
            if (rc != TFG_OK) {
                drop_graph(pt);
                return rc;
            }
            pt->key_image = d_image;
            pt->key_accum = d_accum;
            pt->key_nrc = pt->nrc;
            pt->key_mode = pt->P.mode;
        }
        rc = launch_graphs(pt, s, frame);
        if (rc != TFG_OK) return rc;
    } else {
        rc = enqueue_frame(pt, s, frame, d_image, d_accum);
        if (rc != TFG_OK) return rc;
    }
    TFG_HIP_CHECK(hipEventRecord(pt->ev_out, s));
    TFG_HIP_CHECK(hipStreamWaitEvent(stream, pt->ev_out, 0));
    return TFG_OK;
}

int tfg_pt_counters(tfg_pt* pt, uint32_t* out) {
    if (!pt || !out) { set_error("tfg_pt_counters: NULL argument"); return TFG_E_INVALID; }
    TFG_HIP_CHECK(hipStreamSynchronize(pt->stream));
    TFG_HIP_CHECK(hipMemcpy(out, pt->P.counters, CNT_TOTAL * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return TFG_OK;
}

int tfg_pt_ray_totals(tfg_pt* pt, uint64_t* out) {
    if (!pt || !out) { set_error("tfg_pt_ray_totals: NULL argument"); return TFG_E_INVALID; }
    uint32_t c[CNT_ALLOC];
    TFG_HIP_CHECK(hipStreamSynchronize(pt->stream));
    TFG_HIP_CHECK(hipMemcpy(c, pt->P.counters, sizeof(c), hipMemcpyDeviceToHost));
    uint64_t cum[2];
    std::memcpy(cum, c + CNT_CUM, sizeof(cum));
    uint64_t e = 0, sh = 0;
    for (uint32_t d = 0; d < kMaxDepth; ++d) {
        e += c[CNT_EXT + d];
        sh += c[CNT_SHD + d];
    }
    out[0] = cum[0] + e;
    out[1] = cum[1] + sh;
    return TFG_OK;
}

int tfg_pt_reset_totals(tfg_pt* pt) {
    if (!pt) { set_error("tfg_pt_reset_totals: NULL"); return TFG_E_INVALID; }
    TFG_HIP_CHECK(hipStreamSynchronize(pt->stream));
    TFG_HIP_CHECK(hipMemset(pt->P.counters, 0, (size_t)CNT_ALLOC * (1 + kMaxTiles) * sizeof(uint32_t)));
    return TFG_OK;
}

int tfg_pt_profile_frame(tfg_pt* pt, int32_t frame, tfg_color* d_image, tfg_pt_profile* out) {
    if (!pt || !d_image || !out) { set_error("tfg_pt_profile_frame: NULL argument"); return TFG_E_INVALID; }
    if (!pt->scene_set || !pt->camera_set) { set_error("tfg_pt_profile_frame: scene or camera not set"); return TFG_E_STATE; }
    std::memset(out, 0, sizeof(*out));
    hipStream_t s = pt->stream;
    TFG_HIP_CHECK(hipStreamSynchronize(s));
    TFG_HIP_CHECK(hipMemset(pt->P.counters + CNT_STAT, 0, 8 * sizeof(uint32_t)));
    StageEvents pe;
    int rc = pe.create((uint32_t)pt->tiles, pt->cfg.max_depth);
    if (rc != TFG_OK) return rc;
    // the frame twice: once with the counting kernels (node / triangle
    // visits), then with the plain kernels, timed
    rc = enqueue_frame(pt, s, frame, d_image, nullptr, nullptr, true);
    uint64_t st[4] = {0, 0, 0, 0};
    if (rc == TFG_OK) {
        TFG_HIP_CHECK(hipStreamSynchronize(s));
        TFG_HIP_CHECK(hipMemcpy(st, pt->P.counters + CNT_STAT, sizeof(st), hipMemcpyDeviceToHost));
        rc = enqueue_frame(pt, s, frame, d_image, nullptr, &pe, false);
    }
    if (rc == TFG_OK) {
        TFG_HIP_CHECK(hipStreamSynchronize(s));
        auto span = [](hipEvent_t a, hipEvent_t b) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            return ms;
        };
        // per stage: the sum of its launch durations over the tiles
        for (int t = 0; t < pt->tiles; ++t) {
            out->ms_generate += span(pe.gen[t][0], pe.gen[t][1]);
            for (uint32_t d = 0; d < staged_depths(pt); ++d) {
                out->ms_extend[d] += span(pe.ext[t][d][0], pe.ext[t][d][1]);
                out->ms_shade[d] += span(pe.shd[t][d][0], pe.shd[t][d][1]);
                if (connect_launched(pt, d)) out->ms_connect[d] += span(pe.con[t][d][0], pe.con[t][d][1]);
            }
            out->ms_finalize += span(pe.fin[t][0], pe.fin[t][1]);
            if (staged_depths(pt) < pt->cfg.max_depth) out->ms_tail += span(pe.tail[t][0], pe.tail[t][1]);
        }
        for (uint32_t d = 0; d < staged_depths(pt); ++d) {
            out->at_extend[d] = span(pe.frame[0], pe.ext[0][d][0]);
            out->at_shade[d] = span(pe.frame[0], pe.shd[0][d][0]);
            if (connect_launched(pt, d)) out->at_connect[d] = span(pe.frame[0], pe.con[0][d][0]);
        }
        out->at_finalize = span(pe.frame[0], pe.fin[0][0]);
        if (staged_depths(pt) < pt->cfg.max_depth) out->at_tail = span(pe.frame[0], pe.tail[0][0]);
        out->ms_total = span(pe.frame[0], pe.frame[1]);
        uint32_t c[CNT_ALLOC];
        TFG_HIP_CHECK(hipMemcpy(c, pt->P.counters, sizeof(c), hipMemcpyDeviceToHost));
        for (uint32_t d = 0; d < kMaxDepth; ++d) {
            out->ext_rays[d] = c[CNT_EXT + d];
            out->shadow_rays[d] = c[CNT_SHD + d];
        }
        out->ext_nodes = st[0];
        out->ext_tris = st[1];
        out->shadow_nodes = st[2];
        out->shadow_tris = st[3];
    }
    return rc;
}

int tfg_pt_autotune(tfg_pt* pt, tfg_stream stream) {
    if (!pt) { set_error("tfg_pt_autotune: NULL"); return TFG_E_INVALID; }
    if (!pt->scene_set || !pt->camera_set) { set_error("tfg_pt_autotune: scene or camera not set"); return TFG_E_STATE; }
    (void)stream;
    drop_graph(pt);
    tfg_color* img = nullptr;
    TFG_HIP_CHECK(hipMalloc(&img, sizeof(tfg_color) * pt->P.n_paths));
    hipEvent_t a, b;
    TFG_HIP_CHECK(hipEventCreate(&a));
    TFG_HIP_CHECK(hipEventCreate(&b));
    const uint32_t saved_mode = pt->P.mode;
    pt->P.mode = 0;   // time the plain frame: the traversal stages dominate it
    int rc = TFG_OK;
    // Coordinate search: each step sweeps one setting with the others held;
    // a candidate's time is the faster of two interleaved passes of 2 untimed
    // and 8 timed frames (direct launches; the tuner captures no graph).
    auto sweep = [&](int n, auto&& apply) {
        float best_ms[64];
        for (int i = 0; i < n; ++i) best_ms[i] = 1e30f;
        for (int pass = 0; pass < 2 && rc == TFG_OK; ++pass) {
            for (int i = 0; i < n && rc == TFG_OK; ++i) {
                apply(i);
                // the same 8 timed frame seeds for every candidate
                for (int w = 0; w < 2 && rc == TFG_OK; ++w) rc = enqueue_frame(pt, pt->stream, 1000 + w, img, nullptr);
                (void)hipEventRecord(a, pt->stream);
                for (int f = 0; f < 8 && rc == TFG_OK; ++f) rc = enqueue_frame(pt, pt->stream, f, img, nullptr);
                (void)hipEventRecord(b, pt->stream);
                (void)hipEventSynchronize(b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                if (ms < best_ms[i]) best_ms[i] = ms;
            }
        }
        int bi = 0;
        for (int i = 1; i < n; ++i)
            if (best_ms[i] < best_ms[bi]) bi = i;
        apply(bi);
    };
    pt->tiles = 1;
    pt->sides = 2;
    pt->overlap = true;
    pt->connect_batch = 0;
    pt->tail_from = 0;
    pt->P.sort_from = 0;
    pt->P.ray_sort_from = 0;
    pt->P.refill_from = 0;
    pt->refill_cap = 0;
    // 1. traversal launch shape and LDS-staged top nodes together (large
    //    multipliers give more workgroups than chunks: each then traces one
    //    chunk and retires, which pays only with cheap workgroup start-up)
    const int blocks[3] = {64, 128, 256};
    const int mults[4] = {1, 4, 16, 64};
    const uint32_t tops[4] = {0, 1, 9, 73};
    sweep(48, [&](int i) {
        pt->top_nodes = tops[i / 12];
        pt->P.n_top = pt->scene_nodes < pt->top_nodes ? pt->scene_nodes : pt->top_nodes;
        set_trace_shape(pt, blocks[(i % 12) / 4], mults[i % 4]);
    });
    // 2. tiles, side streams and overlap; 3. XCD super-chunks, octet
    // threshold, shadow-stage grid and batching, tail kernel; then a second
    // round over every setting but the launch shape, from the first round's
    // result (the first round sweeps the top nodes and some settings before
    // the ones that interact with them are tuned)
    struct Cand {
        int tiles, sides;
        bool overlap;
    };
    const Cand cands[] = {{1, 1, false}, {1, 1, true}, {1, 2, true}, {1, 4, true}, {2, 2, true}, {2, 4, true}};
    const uint32_t groups[4] = {0, 1024, 4096, 16384};
    const uint32_t octs[6] = {0, 1u << 16, 1u << 17, 1u << 18, 1u << 19, 1u << 20};
    const int cpcs[3] = {0, 2, 3};
    const uint32_t batches[6] = {0, 2, 3, 4, 5, 6};
    const uint32_t tails[6] = {0, 2, 3, 4, 5, 6};
    const uint32_t sorts[3] = {0, 1, 2};
    const uint32_t idles[3] = {16, 32, 48};
    const int refill_caps[4] = {0, 16, 14, 12};
    for (int round = 0; round < 2 && rc == TFG_OK; ++round) {
        if (round > 0)
            sweep(4, [&](int i) {
                pt->top_nodes = tops[i];
                pt->P.n_top = pt->scene_nodes < pt->top_nodes ? pt->scene_nodes : pt->top_nodes;
            });
        sweep(6, [&](int i) {
            pt->tiles = cands[i].tiles;
            pt->sides = cands[i].sides;
            pt->overlap = cands[i].overlap;
        });
        sweep(4, [&](int i) { pt->P.xcd_group = groups[i]; });
        sweep(6, [&](int i) { pt->P.octet_below = octs[i]; });
        sweep(3, [&](int i) { pt->connect_per_cu = cpcs[i]; });
        sweep(6, [&](int i) { pt->connect_batch = batches[i]; });
        sweep(6, [&](int i) { pt->tail_from = tails[i]; });
        sweep(3, [&](int i) { pt->P.sort_from = sorts[i]; });
        sweep(3, [&](int i) { pt->P.ray_sort_from = sorts[i]; });
        sweep(3, [&](int i) { pt->P.refill_from = sorts[i]; });
        if (pt->P.refill_from) sweep(3, [&](int i) { pt->P.refill_idle = idles[i]; });
        if (pt->P.refill_from) sweep(4, [&](int i) {
            pt->refill_cap = refill_caps[i];
            set_trace_shape(pt, pt->trace_block, pt->trace_mult);
        });
    }
    drop_graph(pt);
    pt->P.mode = saved_mode;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(img);
    return rc;
}

int tfg_pt_get_launch(const tfg_pt* pt, uint32_t* trace_block, uint32_t* trace_grid, uint32_t* shade_grid) {
    if (!pt) { set_error("tfg_pt_get_launch: NULL"); return TFG_E_INVALID; }
    if (trace_block) *trace_block = (uint32_t)pt->trace_block;
    if (trace_grid) *trace_grid = (uint32_t)pt->trace_grid;
    if (shade_grid) *shade_grid = path_grid(pt, (pt->P.n_paths + pt->tiles - 1) / pt->tiles);   // tile 0's
    return TFG_OK;
}

int tfg_pt_get_schedule(const tfg_pt* pt, tfg_pt_schedule* out) {
    if (!pt || !out) { set_error("tfg_pt_get_schedule: NULL argument"); return TFG_E_INVALID; }
    out->trace_block = (uint32_t)pt->trace_block;
    out->grid_mult = (uint32_t)pt->trace_mult;
    out->tiles = (uint32_t)pt->tiles;
    out->side_streams = (uint32_t)pt->sides;
    out->overlap = pt->overlap ? 1 : 0;
    out->use_graph = pt->run_graph ? 1 : 0;
    out->xcd_group = pt->P.xcd_group;
    out->octet_below = pt->P.octet_below;
    out->connect_per_cu = (uint32_t)pt->connect_per_cu;
    out->connect_batch = pt->connect_batch;
    out->top_nodes = pt->top_nodes;
    out->tail_from = pt->tail_from;
    out->defer_training = pt->defer_train ? 1 : 0;
    out->shadow_sort_from = pt->P.sort_from;
    out->ray_sort_from = pt->P.ray_sort_from;
    out->refill_from = pt->P.refill_from;
    out->refill_idle = pt->P.refill_idle;
    out->refill_per_cu = (uint32_t)pt->refill_cap;
    return TFG_OK;
}

int tfg_pt_set_schedule(tfg_pt* pt, const tfg_pt_schedule* sc) {
    if (!pt || !sc) { set_error("tfg_pt_set_schedule: NULL argument"); return TFG_E_INVALID; }
    const uint32_t b = sc->trace_block;
    if ((b != 64 && b != 128 && b != 256) || sc->grid_mult == 0 || sc->grid_mult > 64 || sc->tiles == 0 ||
        sc->tiles > kMaxTiles || sc->side_streams == 0 || sc->side_streams > kMaxSide ||
        (sc->xcd_group != 0 && (sc->xcd_group % 256 != 0 || sc->xcd_group > (1u << 20))) ||
        sc->connect_per_cu > 64 || sc->connect_batch > kMaxDepth || sc->top_nodes > kTopNodes ||
        sc->tail_from > kMaxDepth || sc->defer_training > 1 || sc->shadow_sort_from > kMaxDepth ||
        sc->ray_sort_from > kMaxDepth || sc->refill_from > kMaxDepth || sc->refill_idle == 0 ||
        sc->refill_idle > 64 || sc->refill_per_cu > 64) {
        set_error("tfg_pt_set_schedule: block 64/128/256, grid_mult 1..64, tiles 1..%u, side streams 1..%u, "
                  "xcd_group 0 or a multiple of 256 up to 2^20, connect_per_cu 0..64, connect_batch, tail_from, "
                  "shadow_sort_from, ray_sort_from and refill_from 0..%u, refill_idle 1..64, refill_per_cu 0..64, top_nodes 0..%u, "
                  "defer_training 0/1", kMaxTiles, kMaxSide, kMaxDepth, kTopNodes);
        return TFG_E_INVALID;
    }
    if ((sc->use_graph != 0) != pt->run_graph) {
        set_error("tfg_pt_set_schedule: graph replay is fixed when the tracer is created");
        return TFG_E_STATE;
    }
    if (sc->defer_training && pt->run_graph) {
        set_error("tfg_pt_set_schedule: defer_training needs direct launches (tfg_pt_config.use_graph = 0)");
        return TFG_E_STATE;
    }
    if (!sc->defer_training) {
        const int rc = run_pending_training(pt, pt->stream);
        if (rc != TFG_OK) return rc;
    }
    pt->defer_train = sc->defer_training != 0;
    drop_graph(pt);
    pt->top_nodes = sc->top_nodes;
    pt->P.n_top = pt->scene_nodes < pt->top_nodes ? pt->scene_nodes : pt->top_nodes;
    pt->refill_cap = (int)sc->refill_per_cu;
    set_trace_shape(pt, (int)b, (int)sc->grid_mult);
    pt->tiles = (int)sc->tiles;
    pt->sides = (int)sc->side_streams;
    pt->overlap = sc->overlap != 0;
    pt->P.xcd_group = sc->xcd_group;
    pt->P.octet_below = sc->octet_below;
    pt->connect_per_cu = (int)sc->connect_per_cu;
    pt->connect_batch = sc->connect_batch;
    pt->tail_from = sc->tail_from;
    pt->P.sort_from = sc->shadow_sort_from;
    pt->P.ray_sort_from = sc->ray_sort_from;
    pt->P.refill_from = sc->refill_from;
    pt->P.refill_idle = sc->refill_idle;
    return TFG_OK;
}

int tfg_pt_set_nrc(tfg_pt* pt, tfg_nrc* nrc) {
    if (!pt) { set_error("tfg_pt_set_nrc: NULL"); return TFG_E_INVALID; }
    if (nrc && pt->cfg.mode == 0) { set_error("tfg_pt_set_nrc: tracer was created in plain mode"); return TFG_E_STATE; }
    if (nrc != pt->nrc) {   // the previous cache's pending steps first
        const int rc = run_pending_training(pt, pt->stream);
        if (rc != TFG_OK) return rc;
    }
    pt->nrc = nrc;
    pt->nrc_frame.records = pt->P.records;
    pt->nrc_frame.queries = pt->P.queries;
    pt->nrc_frame.counters = pt->P.counters;
    pt->nrc_frame.n_paths = pt->P.n_paths;
    pt->nrc_frame.mode = pt->P.mode;
    drop_graph(pt);
    return TFG_OK;
}

int tfg_pt_flush_training(tfg_pt* pt, tfg_stream stream) {
    if (!pt) { set_error("tfg_pt_flush_training: NULL"); return TFG_E_INVALID; }
    if (!pt->train_pending) return TFG_OK;
    TFG_HIP_CHECK(hipEventRecord(pt->ev_in, stream));
    TFG_HIP_CHECK(hipStreamWaitEvent(pt->stream, pt->ev_in, 0));
    const int rc = run_pending_training(pt, pt->stream);
    if (rc != TFG_OK) return rc;
    TFG_HIP_CHECK(hipEventRecord(pt->ev_out, pt->stream));
    TFG_HIP_CHECK(hipStreamWaitEvent(stream, pt->ev_out, 0));
    return TFG_OK;
}

int tfg_pt_set_mode(tfg_pt* pt, uint32_t mode) {
    if (!pt || mode > 3) { set_error("tfg_pt_set_mode: bad argument"); return TFG_E_INVALID; }
    if (mode && pt->cfg.mode == 0) { set_error("tfg_pt_set_mode: tracer was created in plain mode"); return TFG_E_STATE; }
    if (pt->P.mode != mode) drop_graph(pt);
    pt->P.mode = mode;
    pt->nrc_frame.mode = mode;
    return TFG_OK;
}

int tfg_pt_nrc_buffers(tfg_pt* pt, float** d_records, float** d_queries, uint32_t** d_counters) {
    if (!pt) { set_error("tfg_pt_nrc_buffers: NULL"); return TFG_E_INVALID; }
    if (d_records) *d_records = reinterpret_cast<float*>(pt->P.records);
    if (d_queries) *d_queries = reinterpret_cast<float*>(pt->P.queries);
    if (d_counters) *d_counters = pt->P.counters;
    return TFG_OK;
}

int tfg_pt_read_cache_data(tfg_pt* pt, float* records, uint32_t max_records, float* queries, uint32_t max_queries,
                           uint32_t* n_records, uint32_t* n_queries) {
    if (!pt) { set_error("tfg_pt_read_cache_data: NULL"); return TFG_E_INVALID; }
    if (pt->cfg.mode == 0) { set_error("tfg_pt_read_cache_data: tracer has no cache buffers"); return TFG_E_STATE; }
    TFG_HIP_CHECK(hipStreamSynchronize(pt->stream));
    uint32_t c[CNT_TOTAL];
    TFG_HIP_CHECK(hipMemcpy(c, pt->P.counters, sizeof(c), hipMemcpyDeviceToHost));
    const uint32_t nr = c[CNT_RECORDS] < max_records ? c[CNT_RECORDS] : max_records;
    const uint32_t nq = c[CNT_QUERIES] < max_queries ? c[CNT_QUERIES] : max_queries;
    if (records && nr)
        TFG_HIP_CHECK(hipMemcpy(records, pt->P.records, (size_t)nr * kRecF4 * sizeof(float4), hipMemcpyDeviceToHost));
    if (queries && nq)
        TFG_HIP_CHECK(hipMemcpy(queries, pt->P.queries, (size_t)nq * kQueryF4 * sizeof(float4), hipMemcpyDeviceToHost));
    if (n_records) *n_records = c[CNT_RECORDS];
    if (n_queries) *n_queries = c[CNT_QUERIES];
    return TFG_OK;
}

int tfg_pt_destroy(tfg_pt* pt) {
    if (!pt) return TFG_OK;
    if (pt->stream) (void)hipStreamSynchronize(pt->stream);
    if (pt->cstream) (void)hipStreamSynchronize(pt->cstream);   // pending steps not yet launched are dropped
    for (uint32_t t = 0; t < kMaxTiles; ++t) {
        if (pt->tstream[t]) (void)hipStreamSynchronize(pt->tstream[t]);
        for (hipStream_t q : pt->tside[t])
            if (q) (void)hipStreamSynchronize(q);
    }
    drop_graph(pt);
    for (void* p : pt->allocs) (void)hipFree(p);
    if (pt->ev_in) (void)hipEventDestroy(pt->ev_in);
    if (pt->ev_out) (void)hipEventDestroy(pt->ev_out);
    if (pt->ev_start) (void)hipEventDestroy(pt->ev_start);
    if (pt->ev_gathered) (void)hipEventDestroy(pt->ev_gathered);
    if (pt->ev_trained) (void)hipEventDestroy(pt->ev_trained);
    if (pt->train_buf) (void)hipFree(pt->train_buf);
    if (pt->train_nrec) (void)hipFree(pt->train_nrec);
    if (pt->wt_tris) (void)hipFree(pt->wt_tris);
    for (uint32_t t = 0; t < kMaxTiles; ++t) {
        for (hipEvent_t e : pt->ev_cjoin[t])
            if (e) (void)hipEventDestroy(e);
        if (pt->ev_tdone[t]) (void)hipEventDestroy(pt->ev_tdone[t]);
        for (hipEvent_t e : pt->ev_shaded[t])
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t q : pt->tside[t])
            if (q) (void)hipStreamDestroy(q);
        for (hipStream_t q : pt->cap_side[t])
            if (q) (void)hipStreamDestroy(q);
        if (pt->cap_tile[t]) (void)hipStreamDestroy(pt->cap_tile[t]);
        if (pt->tstream[t]) (void)hipStreamDestroy(pt->tstream[t]);
    }
    if (pt->cap_post) (void)hipStreamDestroy(pt->cap_post);
    if (pt->stream) (void)hipStreamDestroy(pt->stream);
    if (pt->cstream) (void)hipStreamDestroy(pt->cstream);
    delete pt;
    return TFG_OK;
}

}  // extern "C"
