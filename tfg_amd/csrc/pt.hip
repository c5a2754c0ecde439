// pt.hip - host side of the wavefront path tracer (tfg_pt_* C ABI).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "host_common.h"
#include "nrc_iface.h"
#include "pt_kernels.h"

namespace tfg {
uint32_t known_tree_depth(const void* d_nodes);
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
    int shade_grid = 0;
    int gen_grid = 0;
    // graph
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;    // shadow-ray stage (connect), overlapped with the next depth
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipEvent_t ev_shaded[tfg::kMaxDepth] = {};   // shade of depth d done: its shadow queue is complete
    hipEvent_t ev_joined = nullptr;              // every connect of the frame done
    bool overlap = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key_image = nullptr;
    const void* key_accum = nullptr;
    const void* key_nrc = nullptr;
    uint32_t key_mode = 0;
    // radiance cache attached to the frame (nrc.hip)
    tfg_nrc* nrc = nullptr;
    tfg_nrc_frame nrc_frame{};
    bool allow_no_cache = false;   // drop-in export frames: queries leave the frame uncomposited
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

void drop_graph(tfg_pt* pt) {
    if (pt->exec) (void)hipGraphExecDestroy(pt->exec);
    if (pt->graph) (void)hipGraphDestroy(pt->graph);
    pt->exec = nullptr;
    pt->graph = nullptr;
}

constexpr uint32_t kTopNodes = 73;   // root + 8 + 64 nodes staged in LDS

size_t trace_lds(const tfg_pt* pt, int block) {
    return (size_t)pt->P.n_top * 80 + (size_t)pt->P.stack_depth * block * 8;
}

template <int B>
int occupancy_extend(const tfg_pt* pt, int* blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_extend<B>, B, trace_lds(pt, B)) == hipSuccess ? 0 : -1;
}

int set_trace_shape(tfg_pt* pt, int block, int mult) {
    int per_cu = 0;
    if (block == 64) occupancy_extend<64>(pt, &per_cu);
    else if (block == 128) occupancy_extend<128>(pt, &per_cu);
    else occupancy_extend<256>(pt, &per_cu);
    if (per_cu <= 0) per_cu = 1;
    pt->trace_block = block;
    pt->trace_mult = mult;
    pt->trace_grid = pt->n_cu * per_cu * mult;
    return TFG_OK;
}

template <bool kStats>
void launch_extend(tfg_pt* pt, hipStream_t s, uint32_t q, uint32_t d) {
    const dim3 g(pt->trace_grid);
    const size_t lds = trace_lds(pt, pt->trace_block);
    switch (pt->trace_block) {
        case 64: hipLaunchKernelGGL((k_extend<64, kStats>), g, dim3(64), lds, s, pt->P, q, d); break;
        case 256: hipLaunchKernelGGL((k_extend<256, kStats>), g, dim3(256), lds, s, pt->P, q, d); break;
        default: hipLaunchKernelGGL((k_extend<128, kStats>), g, dim3(128), lds, s, pt->P, q, d); break;
    }
}

template <bool kStats>
void launch_connect(tfg_pt* pt, hipStream_t s, uint32_t d) {
    const dim3 g(pt->trace_grid);
    const size_t lds = trace_lds(pt, pt->trace_block);
    switch (pt->trace_block) {
        case 64: hipLaunchKernelGGL((k_connect<64, kStats>), g, dim3(64), lds, s, pt->P, d); break;
        case 256: hipLaunchKernelGGL((k_connect<256, kStats>), g, dim3(256), lds, s, pt->P, d); break;
        default: hipLaunchKernelGGL((k_connect<128, kStats>), g, dim3(128), lds, s, pt->P, d); break;
    }
}

// Per-stage events of a profiled frame (nullptr: plain frame): a start and
// an end event around each stage, recorded on the stream the stage runs on.
struct StageEvents {
    hipEvent_t gen[2] = {}, fin[2] = {};
    hipEvent_t ext[kMaxDepth][2] = {}, shd[kMaxDepth][2] = {}, con[kMaxDepth][2] = {};
    int create(uint32_t depth) {
        hipError_t e = hipSuccess;
        auto mk = [&](hipEvent_t* p) {
            if (e == hipSuccess) e = hipEventCreate(p);
        };
        for (int k = 0; k < 2; ++k) {
            mk(&gen[k]);
            mk(&fin[k]);
            for (uint32_t d = 0; d < depth; ++d) {
                mk(&ext[d][k]);
                mk(&shd[d][k]);
                mk(&con[d][k]);
            }
        }
        return e == hipSuccess ? TFG_OK : hip_fail(e, "hipEventCreate(profile)");
    }
    ~StageEvents() {
        for (int k = 0; k < 2; ++k) {
            if (gen[k]) (void)hipEventDestroy(gen[k]);
            if (fin[k]) (void)hipEventDestroy(fin[k]);
            for (uint32_t d = 0; d < kMaxDepth; ++d) {
                if (ext[d][k]) (void)hipEventDestroy(ext[d][k]);
                if (shd[d][k]) (void)hipEventDestroy(shd[d][k]);
                if (con[d][k]) (void)hipEventDestroy(con[d][k]);
            }
        }
    }
};

inline void mark(StageEvents* pe, hipEvent_t e, hipStream_t s) {
    if (pe) (void)hipEventRecord(e, s);
}

// The frame body: every launch reads its work size from device counters.
// The main stream runs generate, then extend(d) -> shade(d) per depth, then
// finalize; connect(d) runs on the side stream after shade(d), overlapping
// extend/shade of the following depths, and finalize waits for the last
// connect.  Without overlap every stage runs on the main stream.
int enqueue_frame(tfg_pt* pt, hipStream_t s, tfg_color* image, double* accum, StageEvents* pe = nullptr) {
    const bool nrc = pt->P.mode != 0;
    hipStream_t sc = pt->overlap ? pt->side : s;
    mark(pe, pe ? pe->gen[0] : nullptr, s);
    hipLaunchKernelGGL(k_frame_reset, dim3(1), dim3(128), 0, s, pt->P);
    hipLaunchKernelGGL(k_generate, dim3(pt->gen_grid), dim3(256), 0, s, pt->P);
    mark(pe, pe ? pe->gen[1] : nullptr, s);
    for (uint32_t d = 0; d < pt->cfg.max_depth; ++d) {
        const uint32_t q = d & 1u;
        mark(pe, pe ? pe->ext[d][0] : nullptr, s);
        if (pe) launch_extend<true>(pt, s, q, d);
        else launch_extend<false>(pt, s, q, d);
        mark(pe, pe ? pe->ext[d][1] : nullptr, s);
        mark(pe, pe ? pe->shd[d][0] : nullptr, s);
        if (nrc) hipLaunchKernelGGL(k_shade<true>, dim3(pt->shade_grid), dim3(256), 0, s, pt->P, q, d);
        else hipLaunchKernelGGL(k_shade<false>, dim3(pt->shade_grid), dim3(256), 0, s, pt->P, q, d);
        mark(pe, pe ? pe->shd[d][1] : nullptr, s);
        if (pt->overlap) {
            TFG_HIP_CHECK(hipEventRecord(pt->ev_shaded[d], s));
            TFG_HIP_CHECK(hipStreamWaitEvent(sc, pt->ev_shaded[d], 0));
        }
        mark(pe, pe ? pe->con[d][0] : nullptr, sc);
        if (pe) launch_connect<true>(pt, sc, d);
        else launch_connect<false>(pt, sc, d);
        mark(pe, pe ? pe->con[d][1] : nullptr, sc);
    }
    if (pt->overlap) {
        TFG_HIP_CHECK(hipEventRecord(pt->ev_joined, sc));
        TFG_HIP_CHECK(hipStreamWaitEvent(s, pt->ev_joined, 0));
    }
    mark(pe, pe ? pe->fin[0] : nullptr, s);
    if (nrc) {
        hipLaunchKernelGGL(k_finalize<true>, dim3(pt->gen_grid), dim3(256), 0, s, pt->P, image, accum);
        if (pt->nrc) {
            const int rc = nrc_frame_enqueue(pt->nrc, pt->nrc_frame, image, s);
            if (rc != TFG_OK) return rc;
        }
        if (accum) hipLaunchKernelGGL(k_accumulate, dim3(pt->gen_grid), dim3(256), 0, s, image, accum, pt->P.n_paths);
    } else {
        hipLaunchKernelGGL(k_finalize<false>, dim3(pt->gen_grid), dim3(256), 0, s, pt->P, image, accum);
    }
    mark(pe, pe ? pe->fin[1] : nullptr, s);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
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
    ALLOC(nee, float4, (size_t)n * cfg->max_depth)
    ALLOC(neemask, uint32_t, n)
    ALLOC(counters, uint32_t, CNT_ALLOC)
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
    if (rc == TFG_OK) rc = dalloc<int32_t>(pt, &dframe, 1);
    if (rc == TFG_OK) rc = dalloc<tfg_camera>(pt, &dcam, 1);
    P.frame = dframe;
    P.cam = dcam;
    if (rc == TFG_OK) {
        int dev = 0;
        hipDeviceProp_t prop;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&pt->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&pt->side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_in, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_out, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pt->ev_joined, hipEventDisableTiming);
        for (uint32_t d = 0; d < cfg->max_depth && e == hipSuccess; ++d)
            e = hipEventCreateWithFlags(&pt->ev_shaded[d], hipEventDisableTiming);
        if (e != hipSuccess) rc = hip_fail(e, "tfg_pt_create");
        else pt->n_cu = prop.multiProcessorCount;
    }
    if (rc == TFG_OK) {
        const hipError_t e = hipMemset(P.counters, 0, CNT_ALLOC * sizeof(uint32_t));
        if (e != hipSuccess) rc = hip_fail(e, "hipMemset(counters)");
    }
    if (rc != TFG_OK) {
        tfg_pt_destroy(pt);
        return rc;
    }
    set_trace_shape(pt, 128, 1);
    pt->shade_grid = pt->n_cu * 8;
    pt->gen_grid = pt->n_cu * 8;
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
    if (sc->n_analytic > kMaxAnalytic || (sc->n_analytic && !sc->d_analytic)) {
        set_error("tfg_pt_set_scene: at most %u analytic primitives", kMaxAnalytic);
        return TFG_E_LIMIT;
    }
    if (known_tree_depth(sc->bvh.d_nodos) > TFG_MAX_TREE_DEPTH) {
        set_error("tfg_pt_set_scene: tree too deep");
        return TFG_E_LIMIT;
    }
    PtParams& P = pt->P;
    P.nodes = static_cast<const float4*>(sc->bvh.d_nodos);
    P.tris = static_cast<const float4*>(sc->bvh.d_tris);
    const uint32_t depth = known_tree_depth(sc->bvh.d_nodos);
    P.stack_depth = depth ? (int)(depth > 1 ? depth - 1 : 1) : TFG_MAX_TREE_DEPTH - 1;
    P.n_top = sc->bvh.num_nodos < kTopNodes ? sc->bvh.num_nodos : kTopNodes;
    P.shading = sc->d_shading;
    P.mats = sc->d_materials;
    P.lights = sc->d_lights;
    P.n_lights = sc->n_lights;
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
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, s, const_cast<int32_t*>(pt->P.frame), frame);
    int rc = TFG_OK;
    if (pt->cfg.use_graph) {
        if (pt->exec && (pt->key_image != d_image || pt->key_accum != d_accum || pt->key_nrc != pt->nrc ||
                         pt->key_mode != pt->P.mode)) {
            drop_graph(pt);
        }
        if (!pt->exec) {
            TFG_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rc = enqueue_frame(pt, s, d_image, d_accum);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != TFG_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            pt->graph = g;
            TFG_HIP_CHECK(hipGraphInstantiate(&pt->exec, g, nullptr, nullptr, 0));
            pt->key_image = d_image;
            pt->key_accum = d_accum;
            pt->key_nrc = pt->nrc;
            pt->key_mode = pt->P.mode;
        }
        TFG_HIP_CHECK(hipGraphLaunch(pt->exec, s));
    } else {
        rc = enqueue_frame(pt, s, d_image, d_accum);
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
    TFG_HIP_CHECK(hipMemset(pt->P.counters, 0, CNT_ALLOC * sizeof(uint32_t)));
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
    int rc = pe.create(pt->cfg.max_depth);
    if (rc != TFG_OK) return rc;
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, s, const_cast<int32_t*>(pt->P.frame), frame);
    rc = enqueue_frame(pt, s, d_image, nullptr, &pe);
    if (rc == TFG_OK) {
        TFG_HIP_CHECK(hipStreamSynchronize(s));
        auto span = [](hipEvent_t a, hipEvent_t b) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            return ms;
        };
        out->ms_generate = span(pe.gen[0], pe.gen[1]);
        for (uint32_t d = 0; d < pt->cfg.max_depth; ++d) {
            out->ms_extend[d] = span(pe.ext[d][0], pe.ext[d][1]);
            out->ms_shade[d] = span(pe.shd[d][0], pe.shd[d][1]);
            out->ms_connect[d] = span(pe.con[d][0], pe.con[d][1]);
        }
        out->ms_finalize = span(pe.fin[0], pe.fin[1]);
        out->ms_total = span(pe.gen[0], pe.fin[1]);
        uint32_t c[CNT_ALLOC];
        TFG_HIP_CHECK(hipMemcpy(c, pt->P.counters, sizeof(c), hipMemcpyDeviceToHost));
        for (uint32_t d = 0; d < kMaxDepth; ++d) {
            out->ext_rays[d] = c[CNT_EXT + d];
            out->shadow_rays[d] = c[CNT_SHD + d];
        }
        uint64_t st[4];
        std::memcpy(st, c + CNT_STAT, sizeof(st));
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
    drop_graph(pt);
    tfg_color* img = nullptr;
    TFG_HIP_CHECK(hipMalloc(&img, sizeof(tfg_color) * pt->P.n_paths));
    hipEvent_t a, b;
    TFG_HIP_CHECK(hipEventCreate(&a));
    TFG_HIP_CHECK(hipEventCreate(&b));
    const int blocks[3] = {64, 128, 256};
    const int mults[3] = {1, 2, 4};
    float best = 1e30f;
    int bb = 128, bm = 1;
    bool bo = true;
    const uint32_t saved_mode = pt->P.mode;
    pt->P.mode = 0;   // time the plain frame: the traversal stages dominate it
    int rc = TFG_OK;
    for (int oi = 0; oi < 2 && rc == TFG_OK; ++oi)
    for (int bi = 0; bi < 3 && rc == TFG_OK; ++bi) {
        for (int mi = 0; mi < 3 && rc == TFG_OK; ++mi) {
            set_trace_shape(pt, blocks[bi], mults[mi]);
            pt->overlap = oi == 0;
            float ms_tot = 0.f;
            for (int rep = 0; rep < 3 && rc == TFG_OK; ++rep) {
                (void)hipEventRecord(a, pt->stream);
                rc = enqueue_frame(pt, pt->stream, img, nullptr);
                (void)hipEventRecord(b, pt->stream);
                (void)hipEventSynchronize(b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                if (rep > 0) ms_tot += ms;
            }
            if (ms_tot < best) {
                best = ms_tot;
                bb = blocks[bi];
                bm = mults[mi];
                bo = pt->overlap;
            }
        }
    }
    pt->P.mode = saved_mode;
    set_trace_shape(pt, bb, bm);
    pt->overlap = bo;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(img);
    (void)stream;
    return rc;
}

int tfg_pt_get_launch(const tfg_pt* pt, uint32_t* trace_block, uint32_t* trace_grid, uint32_t* shade_grid) {
    if (!pt) { set_error("tfg_pt_get_launch: NULL"); return TFG_E_INVALID; }
    if (trace_block) *trace_block = (uint32_t)pt->trace_block;
    if (trace_grid) *trace_grid = (uint32_t)pt->trace_grid;
    if (shade_grid) *shade_grid = (uint32_t)pt->shade_grid;
    return TFG_OK;
}

int tfg_pt_set_overlap(tfg_pt* pt, int enable) {
    if (!pt) { set_error("tfg_pt_set_overlap: NULL"); return TFG_E_INVALID; }
    if (pt->overlap != (enable != 0)) drop_graph(pt);
    pt->overlap = enable != 0;
    return TFG_OK;
}

int tfg_pt_get_overlap(const tfg_pt* pt, int* enabled) {
    if (!pt || !enabled) { set_error("tfg_pt_get_overlap: NULL"); return TFG_E_INVALID; }
    *enabled = pt->overlap ? 1 : 0;
    return TFG_OK;
}

int tfg_pt_set_nrc(tfg_pt* pt, tfg_nrc* nrc) {
    if (!pt) { set_error("tfg_pt_set_nrc: NULL"); return TFG_E_INVALID; }
    if (nrc && pt->cfg.mode == 0) { set_error("tfg_pt_set_nrc: tracer was created in plain mode"); return TFG_E_STATE; }
    pt->nrc = nrc;
    pt->nrc_frame.records = pt->P.records;
    pt->nrc_frame.queries = pt->P.queries;
    pt->nrc_frame.counters = pt->P.counters;
    pt->nrc_frame.n_paths = pt->P.n_paths;
    pt->nrc_frame.mode = pt->P.mode;
    drop_graph(pt);
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
    if (pt->side) (void)hipStreamSynchronize(pt->side);
    drop_graph(pt);
    for (void* p : pt->allocs) (void)hipFree(p);
    if (pt->ev_in) (void)hipEventDestroy(pt->ev_in);
    if (pt->ev_out) (void)hipEventDestroy(pt->ev_out);
    if (pt->ev_joined) (void)hipEventDestroy(pt->ev_joined);
    for (hipEvent_t e : pt->ev_shaded)
        if (e) (void)hipEventDestroy(e);
    if (pt->side) (void)hipStreamDestroy(pt->side);
    if (pt->stream) (void)hipStreamDestroy(pt->stream);
    delete pt;
    return TFG_OK;
}

}  // extern "C"
