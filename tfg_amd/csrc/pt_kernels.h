// pt_kernels.h - the wavefront stages of one frame.
//
//   generate : per pixel seed, jitter and primary ray     (render.h:416-443)
//   extend   : closest hit of every queued ray             (escenario.h:316-366)
//   shade    : emission, NEE request, cache capture, lobe
//              sampling and Russian roulette               (render.h:67-241)
//   connect  : shadow rays of the NEE requests             (escenario.h:144-184)
//   finalize : pixel colour, training records, queries     (render.h:243-260,449-465)
//
// The path loop of lanzarRayoIterativo becomes max_depth rounds of
// extend/shade/connect over compacted queues.  Queue appends are
// wave-aggregated: one ballot, one atomic per wave, lane offsets by popcount.
// Every kernel is a grid-stride (persistent) loop whose trip count comes from
// device counters, so the whole frame can be captured in one hipGraph.
#pragma once
#include "pt_internal.h"
#include "shade_device.h"
#include "trace_device.h"
#include "wave_ops.h"

namespace tfg {

__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// ------------------------------------------------------------------ reset
// One block of 128 threads on the frame counter block: folds the previous
// frame's ray counts into the 64-bit running totals, then clears the
// per-frame counts, and sets the frame index (the seed generate reads).
// (Each tile's queue counters are set by its generate.)
__global__ void k_frame_reset(uint32_t* gc, int32_t* frame, int32_t value) {
    if (threadIdx.x == 0) *frame = value;
    const uint32_t i = threadIdx.x;
    if (i < 64) {
        const uint32_t e = wave_sum(i < kMaxDepth ? gc[CNT_EXT + i] : 0u);
        const uint32_t s = wave_sum(i < kMaxDepth ? gc[CNT_SHD + i] : 0u);
        if (i == 0) {
            unsigned long long* cum = reinterpret_cast<unsigned long long*>(gc + CNT_CUM);
            cum[0] += e;
            cum[1] += s;
        }
    }
    __syncthreads();
    if (i < CNT_TOTAL) gc[i] = 0u;
}

// --------------------------------------------------------------- generate
// Depth-0 queue slot k -> path.  Paths are pixels; the primary rays of one
// wave cover an 8 x 8 pixel block rather than 64 pixels of one row (a
// smaller angular footprint, fewer divergent node visits).  Applies to each
// band of 8 full rows inside this launch's pixel range when one camera
// fills the frame and its width is a multiple of 8; elsewhere slot k holds
// path k.  A bijection on the range, so every path is generated once.
__device__ __forceinline__ uint32_t primary_order(const PtParams& P, uint32_t k) {
    const uint32_t w = P.width;
    if (P.n_views != 0 || (w & 7u) != 0u) return k;
    const uint32_t p = P.pix_base + k;
    const uint32_t band = p / (8u * w), b0 = band * 8u * w;
    if (b0 < P.pix_base || b0 + 8u * w > P.pix_base + P.n_paths) return k;
    const uint32_t o = p - b0, j = o & 63u;
    return (band * 8u + (j >> 3)) * w + (o >> 6) * 8u + (j & 7u) - P.pix_base;
}

// Primary rays of the tile's pixels; block 0 also sets the tile's queue
// counters and zeroes its refill cursors (nothing else reads them while
// generate runs).
__global__ __launch_bounds__(256) void k_generate(PtParams P) {
    if (blockIdx.x == 0) {
        for (uint32_t k = threadIdx.x; k < kMaxDepth; k += blockDim.x) P.counters[CNT_SHQ + k] = 0u;
        for (uint32_t k = threadIdx.x; k <= kMaxDepth; k += blockDim.x) P.counters[CNT_QD + k] = k ? 0u : P.n_paths;
        for (uint32_t k = threadIdx.x; k < kMaxDepth * kRefillCursors; k += blockDim.x) P.rcur[k * kRefillCurStride] = 0u;
    }
    const int32_t frame = *P.frame;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < P.n_paths; k += gridDim.x * blockDim.x) {
        const uint32_t i = primary_order(P, k);
        const uint32_t px = P.pix_base + i;
        // the pixel's camera and its pixel within that camera's image (a
        // view of a multi-view canvas is traced as a tracer of its own size
        // would trace it: same seed, same ray)
        tfg_camera cam;
        uint32_t lp = px, w = P.width, h = P.height;
        if (P.n_views == 0) {
            cam = *P.cam;
        } else {
            uint32_t lo = 0, hi = P.n_views - 1;   // last view with base <= px
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (P.views[mid].base <= px) lo = mid;
                else hi = mid - 1;
            }
            cam = P.views[lo].cam;
            lp = px - P.views[lo].base;
            w = P.views[lo].width;
            h = P.views[lo].height;
        }
        const int x = (int)(lp % w), y = (int)(lp / w);
        uint32_t s = pixel_seed(x, y, frame);
        const float ox = pcg32_float(s);
        const float oy = pcg32_float(s);
        const V3 d = camera_dir(cam, x, y, (int)w, (int)h, ox, oy);
        uint32_t fl = 0;
        // render.h:443: training pixel draw only when both flags are set
        if ((P.mode & MODE_TRAIN) && (P.mode & MODE_INFER)) {
            if (pcg32_float(s) < P.train_frac) fl |= PF_TRAIN;
        } else if (P.mode & MODE_TRAIN) {
            fl |= PF_TRAIN;
        }
        // render.h:65: unbiased suffix draw only for training paths
        if ((fl & PF_TRAIN) && pcg32_float(s) < 0.0625f) fl |= PF_UNBIASED;
        P.rays[0][2 * (size_t)k] = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], 0.0f);
        P.rays[0][2 * (size_t)k + 1] = f4(d, __int_as_float(0x7f800000));
        P.qpath[0][k] = i;
        P.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        P.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        P.rng[i] = s;
        P.flags[i] = fl;
        P.neemask[i] = 0u;
        if (P.mode) P.foot[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ----------------------------------------------------------------- extend
__device__ __forceinline__ void add_stats(PtParams& P, int slot, const TraceStats& st) {
    const uint32_t n = wave_sum(st.nodes), t = wave_sum(st.tris);
    if (__lane_id() == 0 && (n | t)) {
        unsigned long long* c = reinterpret_cast<unsigned long long*>(P.gcounters + CNT_STAT);
        atomicAdd(&c[slot], (unsigned long long)n);
        atomicAdd(&c[slot + 1], (unsigned long long)t);
    }
}

// kStats: also count decoded nodes and tested triangles (profiling frames).
// Dynamic LDS: [n_top nodes x 5 float4][stack_depth x kBlock uint2].
__device__ __forceinline__ void stage_top_nodes(const PtParams& P, float4* top) {
    for (uint32_t k = threadIdx.x; k < 5 * P.n_top; k += blockDim.x) top[k] = P.nodes[k];
    __syncthreads();
}

// Whether depth d's extension rays are traced in key order (count: the
// depth's queue length).  The sort passes and the closest-hit stage decide
// alike.
__device__ __forceinline__ bool rays_sorted(const PtParams& P, uint32_t d, uint32_t count) {
    return P.ray_sort_from != 0 && d >= P.ray_sort_from && count > 0 && count >= (P.n_paths >> 4);
}

// The closest hit of the analytic primitives, after the mesh's
// (escenario.h:323-334): a nearer analytic hit replaces (t, u, v, prim).
__device__ __forceinline__ void analytic_hits(const PtParams& P, float4 a, float4 b, float& t, float& u, float& v,
                                              uint32_t& prim) {
    for (uint32_t k = 0; k < P.n_analytic; ++k) {
        const float ta = analytic_t(P.analytic[k], v3(a.x, a.y, a.z), v3(b.x, b.y, b.z));
        if (ta > a.w && ta < t) {
            t = ta;
            u = 0.0f;
            v = 0.0f;
            prim = kAnalyticBit | k;
        }
    }
}

// Closest hit of a launch with few rays: 8 lanes per ray (traverse_octet),
// kBlock / 8 rays per chunk.  Workgroups without a chunk leave before
// staging the top nodes.
template <int kBlock, bool kStats>
__device__ __forceinline__ void extend_octet(PtParams& P, const float4* __restrict__ rays,
                                             const uint32_t* __restrict__ order, uint32_t count, float4* top,
                                             uint2* stack) {
    constexpr uint32_t kRays = kBlock / 8;
    ChunkIter it((count + kRays - 1) / kRays, P.xcd_group / kRays, P.xcd_group != 0);
    if (!it.valid()) return;
    stage_top_nodes(P, top);
    TraceStats st;   // accumulated over the chunks, flushed once per wave
    const uint32_t j = threadIdx.x & 7u;
    for (; it.valid(); it.next()) {
        const uint32_t iq = it.chunk * kRays + (threadIdx.x >> 3);
        if (iq < count) {
            const uint32_t i = order ? order[iq] : iq;
            const float4 a = rays[2 * (size_t)i];
            const float4 b = rays[2 * (size_t)i + 1];
            const RayPre r = make_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
            float t, u, v;
            uint32_t prim;
            traverse_octet<false>(P.nodes, P.tris, top, P.n_top, r, stack + (threadIdx.x >> 3), kRays,
                                  P.stack_depth, t, u, v, prim, kStats ? &st : nullptr);
            analytic_hits(P, a, b, t, u, v, prim);
            if (j == 0) P.hits[i] = make_float4(t, u, v, __uint_as_float(prim));
        }
    }
    if (kStats) add_stats(P, 0, st);
}

// Closest hit in lockstep: one ray per lane, kBlock rays per chunk, chunks
// in XCD super-chunk order.  Workgroups without a chunk leave before staging
// the top nodes.
template <int kBlock, bool kStats>
__device__ __forceinline__ void extend_lockstep(PtParams& P, const float4* __restrict__ rays,
                                                const uint32_t* __restrict__ order, uint32_t count, float4* top,
                                                uint2* stack) {
    ChunkIter it((count + kBlock - 1) / kBlock, P.xcd_group / kBlock, P.xcd_group != 0);
    if (!it.valid()) return;
    stage_top_nodes(P, top);
    TraceStats st;   // accumulated over the chunks, flushed once per wave
    for (; it.valid(); it.next()) {
        const uint32_t iq = it.chunk * kBlock + threadIdx.x;
        if (iq < count) {
            const uint32_t i = order ? order[iq] : iq;
            const float4 a = rays[2 * (size_t)i];
            const float4 b = rays[2 * (size_t)i + 1];
            const RayPre r = make_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
            float t, u, v;
            uint32_t prim;
            traverse<0, false>(P.nodes, P.tris, top, P.n_top, r, stack + threadIdx.x, kBlock, P.stack_depth, t, u, v,
                               prim, kStats ? &st : nullptr);
            analytic_hits(P, a, b, t, u, v, prim);
            P.hits[i] = make_float4(t, u, v, __uint_as_float(prim));
        }
    }
    if (kStats) add_stats(P, 0, st);
}

template <int kBlock, bool kStats = false>
__global__ __launch_bounds__(kBlock) void k_extend(PtParams P, uint32_t q, uint32_t depth) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* top = reinterpret_cast<float4*>(lds);
    uint2* stack = reinterpret_cast<uint2*>(top + 5 * P.n_top);
    const uint32_t count = P.counters[CNT_QD + depth];
    if (blockIdx.x == 0 && threadIdx.x == 0 && count) atomicAdd(&P.gcounters[CNT_EXT + depth], count);
    const float4* rays = P.rays[q];
    // with the ray order on, queue position i traces the ray at rperm[i] and
    // writes its hit at that ray's own slot (shade reads hits in queue order)
    const uint32_t* order = rays_sorted(P, depth, count) ? P.rperm : nullptr;
    if (count <= P.octet_below) extend_octet<kBlock, kStats>(P, rays, order, count, top, stack);
    else extend_lockstep<kBlock, kStats>(P, rays, order, count, top, stack);
}

// Closest hit with lane refill (dynamic ray fetch, Aila & Laine 2009,
// "Understanding the efficiency of ray traversal on GPUs").  Each lane runs
// one ray's traversal loop (the steps of traverse(), so each ray decodes the
// same nodes, tests the same triangles and gets the same hit); a wave no
// longer waits for its slowest ray: whenever at least refill_idle of its
// lanes have finished, they take the next rays of the queue.
//
// Hand-out: n = min(kRefillCursors, grid) cursors, workgroup b on cursor
// b % n (with n a multiple of 8, each cursor's waves share one XCD:
// workgroup b runs on XCD b % 8; every cursor has a workgroup).  Cursor c
// deals out the queue's super-chunks c, c + n, ...
// (kRefillGroup batches of kRefillBatch consecutive queue positions each),
// one batch per atomic: every cursor gets an even share of the queue, and a
// batch is a run of neighbouring rays.  Batch j of cursor c starts at a
// position that grows with j, so once a wave draws a batch past the queue's
// end every later draw on that cursor does too: the wave then stops drawing
// and leaves when its lanes are done.  Each refill either starts a ray or
// marks the wave drained, so every wave leaves.  A wave keeps the rest of its
// batch (res_lo..res_hi) for its next refills.
//
// Launches with fewer than two rays per resident lane trace in lockstep (too
// few batches per wave for refill to pay), launches of at most octet_below
// rays with 8 lanes per ray.  The cursors (counters of this depth) are zeroed
// by generate.
template <int kBlock, bool kStats = false>
__global__ __launch_bounds__(kBlock) void k_extend_refill(PtParams P, uint32_t q, uint32_t depth) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* top = reinterpret_cast<float4*>(lds);
    uint2* stack_base = reinterpret_cast<uint2*>(top + 5 * P.n_top);
    const uint32_t count = P.counters[CNT_QD + depth];
    if (blockIdx.x == 0 && threadIdx.x == 0 && count) atomicAdd(&P.gcounters[CNT_EXT + depth], count);
    if (count == 0) return;
    const float4* rays = P.rays[q];
    const uint32_t* order = rays_sorted(P, depth, count) ? P.rperm : nullptr;
    if (count <= P.octet_below) {
        extend_octet<kBlock, kStats>(P, rays, order, count, top, stack_base);
        return;
    }
    if (count < 2u * gridDim.x * kBlock) {
        extend_lockstep<kBlock, kStats>(P, rays, order, count, top, stack_base);
        return;
    }
    stage_top_nodes(P, top);
    uint2* stack = stack_base + threadIdx.x;
    const uint32_t n_cur = min(kRefillCursors, gridDim.x);
    const uint32_t c = blockIdx.x % n_cur;
    uint32_t* cursor = P.rcur + ((size_t)depth * kRefillCursors + c) * kRefillCurStride;
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;   // lanes below this one
    uint32_t res_lo = 0, res_hi = 0;                // wave-uniform
    bool drained = false;
    // lane state: the ray and its traversal
    bool busy = false;
    uint32_t slot = 0;
    RayPre r = make_ray(0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f);
    int sp = 0;
    uint32_t ng_x = 0, ng_y = 0, tg_x = 0, tg_y = 0, best_prim = kMiss;
    float tcur = 0.f, bu = 0.f, bv = 0.f;
    bool found = false;
    TraceStats st;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!drained && n_idle >= P.refill_idle) {
            const uint32_t have = res_hi - res_lo;
            uint32_t nb_lo = 0, nb_hi = 0;   // a new batch, drawn when the reservoir is short
            if (have < n_idle) {
                uint32_t j = 0;
                if (lane == 0) j = atomicAdd(cursor, 1u);
                j = __builtin_amdgcn_readfirstlane(j);   // every lane is active here
                const uint32_t pos = ((j / kRefillGroup) * n_cur + c) * (kRefillGroup * kRefillBatch) +
                                     (j % kRefillGroup) * kRefillBatch;
                if (pos < count) {
                    nb_lo = pos;
                    nb_hi = min(pos + kRefillBatch, count);
                } else {
                    drained = true;
                }
            }
            if (!busy) {   // the k-th idle lane takes the reservoir's k-th ray, then the batch's
                const uint32_t k = (uint32_t)__popcll(idle & below);
                uint32_t iq = kMiss;
                if (k < have) iq = res_lo + k;
                else if (k - have < nb_hi - nb_lo) iq = nb_lo + (k - have);
                if (iq != kMiss) {
                    slot = order ? order[iq] : iq;
                    const float4 ra = rays[2 * (size_t)slot];
                    const float4 rb = rays[2 * (size_t)slot + 1];
                    r = make_ray(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, ra.w, rb.w);
                    sp = 0;
                    ng_x = 0;
                    ng_y = 0x80000000u;
                    tg_x = 0;
                    tg_y = 0;
                    tcur = r.tmax;
                    best_prim = kMiss;
                    bu = 0.f;
                    bv = 0.f;
                    found = false;
                    busy = true;
                }
            }
            if (n_idle <= have) {
                res_lo += n_idle;
            } else {
                res_lo = nb_lo + min(n_idle - have, nb_hi - nb_lo);
                res_hi = nb_hi;
            }
        }
        if (__ballot(busy) == 0) {
            if (drained) break;
            continue;
        }
        if (busy) {   // one step of traverse()'s loop
            if (ng_y <= 0x00FFFFFFu && sp > 0) {
                --sp;
                const uint2 e = stack[sp * kBlock];
                ng_x = e.x;
                ng_y = e.y;
            }
            const bool have_node = ng_y > 0x00FFFFFFu;
            float4 n0, n1, n2, n3, n4;
            if (have_node) {
                const uint32_t imask = ng_y;
                const int bit = highest_bit(ng_y & 0xFF000000u);
                ng_y &= ~(1u << bit);
                if (ng_y > 0x00FFFFFFu) {
                    if (sp < P.stack_depth) {
                        stack[sp * kBlock] = make_uint2(ng_x, ng_y);
                        ++sp;
                    } else if (kStats) {
                        st.overflow = 1;   // host refuses trees this deep
                    }
                }
                const uint32_t cslot = ((uint32_t)bit - 24u) ^ r.octinv;
                const uint32_t rel = (uint32_t)__popc(imask & ~(0xFFFFFFFFu << cslot));
                load_node(P.nodes, top, P.n_top, ng_x + rel, n0, n1, n2, n3, n4);
            }
            test_triangles<0, false>(P.tris, tg_x, tg_y, r, tcur, bu, bv, best_prim, found,
                                     kStats ? &st : nullptr);
            if (have_node) {
                if (kStats) st.nodes++;
                const uint32_t hm = node_hitmask(n0, n1, n2, n3, n4, r, tcur);
                ng_x = __float_as_uint(n1.x);
                ng_y = (hm & 0xFF000000u) | (__float_as_uint(n0.w) >> 24);
                tg_x = __float_as_uint(n1.y);
                tg_y = hm & 0x00FFFFFFu;
            } else {   // the ray is done: traverse()'s result, then the analytic primitives
                float t = tcur, u = found ? bu : 0.f, v = found ? bv : 0.f;
                uint32_t prim = found ? best_prim : kMiss;
                analytic_hits(P, make_float4(r.O[0], r.O[1], r.O[2], r.tmin),
                              make_float4(r.D[0], r.D[1], r.D[2], r.tmax), t, u, v, prim);
                P.hits[slot] = make_float4(t, u, v, __uint_as_float(prim));
                busy = false;
            }
        }
    }
    if (kStats) add_stats(P, 0, st);
}

// ------------------------------------------------------------------ shade
__device__ __forceinline__ void pack_vertex(float4* dst, V3 p, V3 dir_mapped, V3 n_mapped, const Mat& m) {
    dst[0] = f4(p, m.rough);
    dst[1] = f4(dir_mapped, m.kd.x);
    dst[2] = f4(n_mapped, m.kd.y);
    dst[3] = make_float4(m.kd.z, m.ks.x, m.ks.y, m.ks.z);
}

// One queue entry of the shading stage (render.h:62-224): the hit (h) of ray
// (ro, rd) of path pid at this depth.  Updates the path's state; returns the
// NEE request (want_shadow, shadow entry e0..e3) and the next ray (want_ray,
// n0, n1).
template <bool kNrc>
__device__ __forceinline__ void shade_entry(const PtParams& P, uint32_t depth, bool last, uint32_t pid, float4 ro,
                                            float4 rd, float4 h, bool& want_shadow, float4& e0, float4& e1,
                                            float4& e2, float4& e3, bool& want_ray, float4& n0, float4& n1) {
    const uint32_t prim = __float_as_uint(h.w);
    const float4 thr4 = P.thr[pid];
    V3 thr = xyz(thr4);
    float pdf = thr4.w;
    const float4 rad4 = P.rad[pid];
    V3 rad = xyz(rad4);
    float ior = rad4.w;
    uint32_t s = P.rng[pid];
    uint32_t fl = P.flags[pid];
    const uint32_t fl_in = fl;
    float4 foot = kNrc ? P.foot[pid] : make_float4(0.f, 0.f, 0.f, 0.f);
    bool rad_dirty = false;
    const V3 d = xyz(rd);
    if (prim != kMiss) {
        const float t = h.x;
        const V3 p = xyz(ro) + d * t;
        uint32_t mat_id;
        V3 n;
        if (prim & kAnalyticBit) {
            const tfg_analytic q = P.analytic[prim & ~kAnalyticBit];
            mat_id = q.material;
            n = analytic_normal(q, p, d);
        } else {
            const tfg_tri_shading sh = P.shading[prim];
            mat_id = sh.material;
            n = shading_normal(sh, ld3(sh.ng), h.y, h.z, d);
        }
        const Mat m = load_mat(&P.mats[mat_id]);
        const float cos_theta = fabsf(dot(n, -d));
        const bool diffuse = mat_diffuse(m), glossy = mat_glossy(m), delta = mat_delta(m);
        bool foot_ok = false;
        if (kNrc) {   // render.h:86-102 path-footprint spread
            const float dist_sq = t * t;
            if (depth == 0) {
                foot.x = fdiv(dist_sq, 4.0f * kPi * fmaxf(cos_theta, 1e-4f));
                foot.y = 0.0f;
            } else {
                const float spread = fdiv(dist_sq, fmaxf(pdf, 1e-4f) * fmaxf(cos_theta, 1e-4f));
                if (!(fl & PF_HEAD)) foot.y += spread;
                else foot.z += spread;
            }
            foot_ok = foot.y >= foot.x * 0.01f;
        }
        bool alive = true;
        if (vmax(m.le) > 0.0f) {   // render.h:104-123
            rad = rad + thr * m.le;
            rad_dirty = true;
            if (kNrc && (fl & PF_HEAD)) {
                float4* tl = P.tail + (size_t)kVertF4 * pid;
                Mat z = m;
                z.kd = v3(0.f, 0.f, 0.f);
                z.ks = v3(0.f, 0.f, 0.f);
                pack_vertex(tl, p, -d, n, z);
                P.tsuf[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
                fl |= PF_TAIL;
            }
            alive = false;
        } else {
            const V3 wo = -d;
            if (diffuse || glossy) {   // NEE request (render.h:126-131)
                want_shadow = true;
                e0 = f4(p + n * kEps, __uint_as_float(pid));
                e1 = f4(n, __uint_as_float(mat_id));
                e2 = f4(wo, 0.0f);
                e3 = f4(thr, 0.0f);
            }
            bool done = false;
            const V3 dm = v3(wo.x * 0.5f + 0.5f, wo.y * 0.5f + 0.5f, wo.z * 0.5f + 0.5f);
            const V3 nm = v3(n.x * 0.5f + 0.5f, n.y * 0.5f + 0.5f, n.z * 0.5f + 0.5f);
            if (kNrc && (fl & PF_HEAD)) {   // render.h:134-161 suffix end
                const bool stop = !(fl & PF_UNBIASED) && foot.z >= foot.w;
                if (stop) {
                    pack_vertex(P.tail + (size_t)kVertF4 * pid, p, dm, nm, m);
                    P.tsuf[pid] = f4(thr * (m.kd + m.ks + m.kt), 0.0f);
                    fl |= PF_TAIL;
                    alive = false;
                    done = true;
                }
            }
            if (kNrc && !done && diffuse && !delta && foot_ok) {   // render.h:166-202
                V3 refl = m.kd + m.ks + m.kt;
                refl = v3(fmaxf(refl.x, 1e-3f), fmaxf(refl.y, 1e-3f), fmaxf(refl.z, 1e-3f));
                if (!(fl & PF_TRAIN) && (P.mode & MODE_INFER)) {
                    pack_vertex(P.head + (size_t)kVertF4 * pid, p, dm, nm, m);
                    P.tthr[pid] = f4(thr * refl, 0.0f);
                    fl |= PF_QUERY;
                    alive = false;
                    done = true;
                } else if ((fl & PF_TRAIN) && !(fl & PF_HEAD)) {
                    pack_vertex(P.head + (size_t)kVertF4 * pid, p, dm, nm, m);
                    P.tthr[pid] = f4(thr * refl, 0.0f);
                    P.pre[pid] = f4(rad, 0.0f);
                    foot.w = foot.y;
                    fl |= PF_HEAD | (depth << PF_HEAD_DEPTH_SHIFT);
                }
            }
            if (!done) {   // calcularLuzIndirecta, render.h:264-335
                const float mkd = vmax(m.kd), mks = vmax(m.ks), mkt = vmax(m.kt);
                const float sum = mkd + mks + mkt;
                V3 wi = v3(0.f, 0.f, 1.f);
                if (!(sum > 0.0f)) {
                    alive = false;
                } else {
                    const float isum = frcp(sum);
                    const float pd = mkd * isum, ps = mks * isum, pt = 1.0f - pd - ps;
                    const float r = pcg32_float(s);
                    if (r < pd) {
                        thr = cdiv(thr * m.kd, pd);
                        const float r1 = pcg32_float(s);
                        const float r2 = pcg32_float(s);
                        const V3 wl = cosine_sample(r1, r2);
                        if (wl.z <= kEps) {
                            alive = false;
                        } else {
                            wi = to_world(make_frame(n), wl);
                            pdf = fmaxf(fabsf(dot(wi, n)), 1e-4f) * (1.0f / kPi);
                        }
                    } else if (r < pd + ps) {
                        if (glossy) {
                            const float r1 = pcg32_float(s);
                            const float r2 = pcg32_float(s);
                            const Frame fr = make_frame(n);
                            const V3 wol = to_local(fr, wo);
                            const float alpha = ggx_alpha(m.rough);
                            if (wol.z <= 0.0f) {
                                alive = false;
                            } else {
                                const V3 hm = ggx_sample_vndf(wol, alpha, r1, r2);
                                const float oh = dot(wol, hm);
                                const V3 wil = hm * (2.0f * oh) - wol;
                                if (wil.z <= 0.0f) {
                                    alive = false;
                                } else {
                                    const float a2 = alpha * alpha;
                                    thr = thr * schlick(m.ks, oh) * fdiv(ggx_G1(a2, wil.z), ps);
                                    wi = normalize(to_world(fr, wil));
                                    pdf = fdiv(ggx_D(a2, hm.z) * ggx_G1(a2, wol.z), 4.0f * wol.z);
                                }
                            }
                        } else {   // render.h:337-340
                            wi = normalize(d - n * (2.0f * dot(d, n)));
                            thr = cdiv(thr * m.ks, ps);
                            pdf = 1e6f;
                        }
                    } else {   // render.h:342-376
                        V3 N = n;
                        float cosi = fminf(1.0f, fmaxf(-1.0f, dot(d, N)));
                        float etat;
                        if (cosi > 0.0f) {
                            N = -N;
                            etat = 1.0f;
                        } else {
                            cosi = -cosi;
                            etat = m.ior;
                        }
                        const float eta = fdiv(ior, etat);
                        const float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
                        if (k < 0.0f) {
                            wi = normalize(d - N * (2.0f * dot(d, N)));
                        } else {
                            ior = etat;
                            wi = normalize(d * eta + N * (eta * cosi - fsqrt(k)));
                        }
                        thr = cdiv(thr * m.kt, pt);
                        pdf = 1e6f;
                    }
                }
                // Russian roulette, render.h:209-224 (after depth rr_depth)
                if (alive && depth > P.rr_depth) {
                    const float prob = fminf(vmax(thr), 0.95f);
                    if (pcg32_float(s) > prob) alive = false;
                    else thr = cdiv(thr, prob);
                }
                if (alive && !last) {
                    want_ray = true;
                    n0 = f4(p + wi * kEps, 0.0f);
                    n1 = f4(wi, __int_as_float(0x7f800000));
                }
                P.thr[pid] = f4(thr, pdf);
                P.rng[pid] = s;
                rad_dirty = rad_dirty || ior != rad4.w;
            }
        }
    }
    if (rad_dirty) P.rad[pid] = f4(rad, ior);
    if (kNrc) {
        P.foot[pid] = foot;
        if (fl != fl_in) P.flags[pid] = fl;
    }
}

// Cells of p in the root node's quantisation frame, where every scene point
// lies in [origin, origin + 256 * 2^e) per axis: `bits` bits per axis,
// interleaved as a Morton code (bits 3b, 3b + 1, 3b + 2: bit b of the x, y, z
// cell); points outside clamp to the border cells.
__device__ __forceinline__ uint32_t root_morton(const PtParams& P, V3 p, int bits) {
    const float4 r = P.nodes[0];
    const uint32_t w = __float_as_uint(r.w);
    const float rel[3] = {p.x - r.x, p.y - r.y, p.z - r.z};
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int e = (int)(int8_t)((w >> (8 * a)) & 0xFFu);
        const float q = fminf(fmaxf(ldexpf(rel[a], -e), 0.0f), 255.0f);
        c[a] = (uint32_t)q >> (8 - bits);
    }
    uint32_t key = 0;
    for (int b = 0; b < bits; ++b)
        key |= (((c[0] >> b) & 1u) << (3 * b)) | (((c[1] >> b) & 1u) << (3 * b + 1)) | (((c[2] >> b) & 1u) << (3 * b + 2));
    return key;
}

// Shadow-entry key: the Morton code of the shadow point, kSortBits / 3 bits
// per axis.
__device__ __forceinline__ uint32_t shadow_sort_key(const PtParams& P, V3 p) {
    return root_morton(P, p, (int)kSortBits / 3);
}

// Ray key: the direction's cube-map cell (6 faces x 4 x 4) times 32, plus the
// origin's 2-bit-per-axis Morton code without its lowest bit (< 96 x 32 =
// 3,072 <= kSortBins): rays of a wave then share direction first, origin
// second.
__device__ __forceinline__ uint32_t ray_sort_key(const PtParams& P, V3 o, V3 w) {
    const float ax = fabsf(w.x), ay = fabsf(w.y), az = fabsf(w.z);
    uint32_t face;
    float m, u, v;
    if (ax >= ay && ax >= az) {
        face = w.x < 0.0f ? 1u : 0u;
        m = ax;
        u = w.y;
        v = w.z;
    } else if (ay >= az) {
        face = w.y < 0.0f ? 3u : 2u;
        m = ay;
        u = w.x;
        v = w.z;
    } else {
        face = w.z < 0.0f ? 5u : 4u;
        m = az;
        u = w.x;
        v = w.y;
    }
    const float inv = m > 0.0f ? 1.0f / m : 0.0f;
    const uint32_t iu = (uint32_t)fminf(fmaxf((u * inv + 1.0f) * 2.0f, 0.0f), 3.0f);
    const uint32_t iv = (uint32_t)fminf(fmaxf((v * inv + 1.0f) * 2.0f, 0.0f), 3.0f);
    return ((face * 4u + iu) * 4u + iv) * 32u + (root_morton(P, o, 2) >> 1);
}

template <bool kNrc>
__global__ __launch_bounds__(256) void k_shade(PtParams P, uint32_t q, uint32_t depth) {
    const uint32_t count = P.counters[CNT_QD + depth];
    const uint32_t qo = q ^ 1u;
    const bool last = depth + 1 >= P.max_depth;
    uint32_t facing = 0;   // shadow rays this thread's entries will trace (lights above the surface)
    for (uint32_t b = blockIdx.x * blockDim.x; b < count; b += gridDim.x * blockDim.x) {
        const uint32_t i = b + threadIdx.x;
        bool want_shadow = false, want_ray = false;
        uint32_t pid = 0;
        float4 e0, e1, e2, e3;           // shadow entry
        float4 n0, n1;                   // next ray
        if (i < count) {
            pid = P.qpath[q][i];
            shade_entry<kNrc>(P, depth, last, pid, P.rays[q][2 * (size_t)i], P.rays[q][2 * (size_t)i + 1], P.hits[i],
                              want_shadow, e0, e1, e2, e3, want_ray, n0, n1);
        }
        uint32_t ss, rs;
        block_append2<256>(want_shadow, &P.counters[CNT_SHQ + depth], ss, want_ray, &P.counters[CNT_QD + depth + 1],
                           rs);
        if (want_shadow) {
            for (uint32_t l = 0; l < P.n_lights; ++l) {
                V3 dir;
                float dist, c;
                facing += light_facing(P.lights[l], xyz(e0), xyz(e1), dir, dist, c) ? 1u : 0u;
            }
            float4* e = P.shadow + 4 * ((size_t)depth * P.depth_stride + ss);
            e[0] = e0;
            e[1] = e1;
            e[2] = e2;
            e[3] = e3;
            if (P.sort_from && depth >= P.sort_from)
                P.skey[(size_t)depth * P.depth_stride + ss] = (uint16_t)shadow_sort_key(P, xyz(e0));
            P.neemask[pid] |= 1u << depth;
        }
        if (want_ray) {
            P.rays[qo][2 * (size_t)rs] = n0;
            P.rays[qo][2 * (size_t)rs + 1] = n1;
            P.qpath[qo][rs] = pid;
            if (P.ray_sort_from && depth + 1 >= P.ray_sort_from) P.rkey[rs] = (uint16_t)ray_sort_key(P, xyz(n0), xyz(n1));
        }
    }
    block_add<256>(facing, &P.gcounters[CNT_SHD + depth]);
}

// ------------------------------------------------------------ key sorts
// Counting sort of up to 2^24 keys (< kSortBins) in three launches: count
// (per workgroup share), exclusive scan over the workgroups per bin (whose
// last workgroup also scans the bin totals), scatter of the positions.  A
// stage that reads its queue through the permutation traces the same
// entries in another order: every entry's result lands where it did
// before.  Scratch (kSortScr words): counts [kSortBlocks][kSortBins]
// (workgroup-major), bin totals [kSortBins], bin bases [kSortBins], a done
// counter that is zero between sorts (zeroed at creation, reset by the last
// workgroup).  Used for the shadow entries of a depth (keys from shade) and
// for the extension rays of a depth (keys from the previous depth's shade).
struct SortJob {
    const uint16_t* key;
    const uint32_t* count;   // device count of keys
    uint32_t* perm;          // positions in key order
    uint32_t* scr;
    uint32_t min_count;      // fewer keys: nothing is sorted (the consumer decides alike)
};
constexpr uint32_t kSortThreads = 512;   // threads of the count and scatter workgroups
constexpr uint32_t kSortTileBins = 16;   // bins per workgroup of the per-bin scan

// Whether depth d's shadow entries are traced in key order (count: the
// depth's shadow queue length).
__device__ __forceinline__ bool shadow_sorted(const PtParams& P, uint32_t d, uint32_t count) {
    return P.sort_from != 0 && d >= P.sort_from && count > 0 && count >= (P.n_paths >> 4);
}

__device__ __forceinline__ bool sort_share(const SortJob& J, uint32_t& lo, uint32_t& hi) {
    const uint32_t count = *J.count;
    if (count == 0 || count < J.min_count) return false;
    const uint32_t per = (count + kSortBlocks - 1) / kSortBlocks;
    lo = blockIdx.x * per;
    hi = lo + per < count ? lo + per : count;   // empty when lo >= count
    return true;
}

// Loads (up to) kSortBatch keys of a thread's stride in flight at once.
constexpr uint32_t kSortBatch = 8;

__global__ __launch_bounds__(kSortThreads) void k_sort_count(SortJob J) {
    __shared__ uint32_t h[kSortBins];
    uint32_t lo, hi;
    if (!sort_share(J, lo, hi)) return;
#pragma unroll
    for (uint32_t m = 0; m < kSortBins / kSortThreads; ++m) h[threadIdx.x + m * kSortThreads] = 0u;
    __syncthreads();
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kSortBatch * kSortThreads) {
        uint32_t k[kSortBatch];
#pragma unroll
        for (uint32_t m = 0; m < kSortBatch; ++m) {
            const uint32_t i = i0 + m * kSortThreads;
            k[m] = i < hi ? (uint32_t)J.key[i] : kSortBins;
        }
#pragma unroll
        for (uint32_t m = 0; m < kSortBatch; ++m)
            if (k[m] < kSortBins) atomicAdd(&h[k[m]], 1u);
    }
    __syncthreads();
    uint32_t* cnt = J.scr + (size_t)blockIdx.x * kSortBins;
#pragma unroll
    for (uint32_t m = 0; m < kSortBins / kSortThreads; ++m) cnt[threadIdx.x + m * kSortThreads] = h[threadIdx.x + m * kSortThreads];
}

// kSortTileBins bins per workgroup: their kSortBlocks x kSortTileBins counts
// through LDS, each bin's column summed by 256 / kSortTileBins threads in
// row ranges, then rewritten as exclusive prefixes over the workgroups; the
// bin totals go to the totals array.  The last workgroup to finish (done
// counter) turns the totals into the bins' bases by an exclusive scan.
__global__ __launch_bounds__(256) void k_sort_scan_bins(SortJob J) {
    constexpr uint32_t kParts = 256 / kSortTileBins, kRows = kSortBlocks / kParts;
    constexpr uint32_t kPer = kSortBins / 256;   // bins per thread of the totals scan
    __shared__ uint32_t tile[kSortBlocks][kSortTileBins + 1];
    __shared__ uint32_t psum[kParts][kSortTileBins];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t is_last;
    const uint32_t count = *J.count;
    if (count == 0 || count < J.min_count) return;
    uint32_t* cnt = J.scr;
    uint32_t* tot = cnt + (size_t)kSortBlocks * kSortBins;
    uint32_t* base = tot + kSortBins;
    uint32_t* done = base + kSortBins;
    const uint32_t bin0 = blockIdx.x * kSortTileBins;
    constexpr uint32_t kLoads = kSortBlocks * kSortTileBins / 256;
    uint32_t tv[kLoads];
#pragma unroll
    for (uint32_t m = 0; m < kLoads; ++m) {   // every load in flight before the LDS stores
        const uint32_t k = threadIdx.x + 256u * m, r = k / kSortTileBins, c = k % kSortTileBins;
        tv[m] = cnt[(size_t)r * kSortBins + bin0 + c];
    }
#pragma unroll
    for (uint32_t m = 0; m < kLoads; ++m) {
        const uint32_t k = threadIdx.x + 256u * m;
        tile[k / kSortTileBins][k % kSortTileBins] = tv[m];
    }
    __syncthreads();
    const uint32_t c = threadIdx.x % kSortTileBins, part = threadIdx.x / kSortTileBins;
    uint32_t s = 0;
    for (uint32_t r = part * kRows; r < part * kRows + kRows; ++r) s += tile[r][c];
    psum[part][c] = s;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t k = 0; k < part; ++k) run += psum[k][c];
    for (uint32_t r = part * kRows; r < part * kRows + kRows; ++r) {
        const uint32_t v = tile[r][c];
        tile[r][c] = run;
        run += v;
    }
    if (part == kParts - 1) tot[bin0 + c] = run;
    __syncthreads();
#pragma unroll
    for (uint32_t m = 0; m < kLoads; ++m) {
        const uint32_t k = threadIdx.x + 256u * m, r = k / kSortTileBins, cc = k % kSortTileBins;
        cnt[(size_t)r * kSortBins + bin0 + cc] = tile[r][cc];
    }
    if (threadIdx.x == 0) {
        __threadfence();   // this workgroup's totals before its count
        is_last = atomicAdd(done, 1u) == gridDim.x - 1u ? 1u : 0u;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t v[kPer], ts = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {   // other workgroups' totals: coherent loads
        v[k] = __hip_atomic_load(&tot[kPer * t + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ts += v[k];
    }
    uint32_t x = ts;   // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if ((int)lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t b = x - ts;
    for (uint32_t w = 0; w < wave; ++w) b += wsum[w];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        base[kPer * t + k] = b;
        b += v[k];
    }
    if (t == 0) *done = 0u;   // for the next sort with this scratch
}

// Key i of workgroup b's share goes to base[key] + (keys of that bin in the
// shares before b) + its rank among its share's keys of that bin (LDS
// atomics: the order within a bin is arbitrary, which changes no result).
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(SortJob J) {
    __shared__ uint32_t cur[kSortBins];
    uint32_t lo, hi;
    if (!sort_share(J, lo, hi)) return;
    const uint32_t* cnt = J.scr;
    const uint32_t* base = J.scr + (size_t)kSortBlocks * kSortBins + kSortBins;
    {
        constexpr uint32_t kPer = kSortBins / kSortThreads;
        uint32_t a[kPer], b[kPer];
#pragma unroll
        for (uint32_t m = 0; m < kPer; ++m) {
            a[m] = base[threadIdx.x + m * kSortThreads];
            b[m] = cnt[(size_t)blockIdx.x * kSortBins + threadIdx.x + m * kSortThreads];
        }
#pragma unroll
        for (uint32_t m = 0; m < kPer; ++m) cur[threadIdx.x + m * kSortThreads] = a[m] + b[m];
    }
    __syncthreads();
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kSortBatch * kSortThreads) {
        uint32_t k[kSortBatch];
#pragma unroll
        for (uint32_t m = 0; m < kSortBatch; ++m) {
            const uint32_t i = i0 + m * kSortThreads;
            k[m] = i < hi ? (uint32_t)J.key[i] : kSortBins;
        }
#pragma unroll
        for (uint32_t m = 0; m < kSortBatch; ++m)
            if (k[m] < kSortBins) J.perm[atomicAdd(&cur[k[m]], 1u)] = i0 + m * kSortThreads;
    }
}

// ---------------------------------------------------------------- connect
// Shadow rays of the NEE requests made at depths d_lo..d_hi
// (escenario.h:144-184), one launch for the range (several depths when the
// late, small depths are batched).  Each entry gets a group of
// 2^light_shift adjacent lanes, one per light (or 8 lanes per light when the
// launch has few rays, see octet_below), so the rays of one shading point are
// traced side by side; the group's first lane then sums the unoccluded
// contributions in light order (the order of the reference's loop).  The
// stage reads only the shadow entries of its depths and writes only
// nee[depth][pid], so it can run on a side stream while later depths are
// extended and shaded; finalize adds the contributions in depth order.  The
// shadow rays are counted by the shading stage (light_facing).  At most 96
// registers (5 waves per SIMD): left to itself the compiler takes 98 and 4
// waves, and the launch runs about 5% longer.
template <int kBlock, bool kStats = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_connect(PtParams P, uint32_t d_lo,
                                                                                            uint32_t d_hi) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* top = reinterpret_cast<float4*>(lds);
    uint2* stack = reinterpret_cast<uint2*>(top + 5 * P.n_top);
    uint32_t count = 0;   // entries of all depths in the range
    uint32_t sorted_mask = 0;   // bit d - d_lo: depth d's entries are traced in key order
    for (uint32_t d = d_lo; d <= d_hi; ++d) {
        const uint32_t c = P.counters[CNT_SHQ + d];
        count += c;
        if (shadow_sorted(P, d, c)) sorted_mask |= 1u << (d - d_lo);
    }
    const uint32_t shift = P.light_shift;
    const uint32_t group = 1u << shift;
    // few shadow rays: 8 lanes per ray (traverse_octet); needs <= 8 lights so
    // that an entry's lanes stay within one wave
    const uint32_t os = ((uint64_t)count << shift) <= P.octet_below && shift <= 3 ? 3u : 0u;
    const uint32_t sh = shift + os;
    const uint64_t work = (uint64_t)count << sh;
    const uint32_t lane = __lane_id();
    const int lead = (int)(lane & ~((1u << sh) - 1u));
    const uint32_t rays_per_chunk = kBlock >> os;
    ChunkIter it((uint32_t)((work + kBlock - 1) / kBlock), P.xcd_group / rays_per_chunk, P.xcd_group != 0);
    if (!it.valid()) return;   // workgroups without a chunk leave before staging
    stage_top_nodes(P, top);
    TraceStats st;            // accumulated over the chunks, flushed once per wave
    for (; it.valid(); it.next()) {
        const uint64_t b = (uint64_t)it.chunk * kBlock;
        const uint64_t w = b + threadIdx.x;
        uint32_t i = (uint32_t)(w >> sh);   // entry over the range, then within its depth
        uint32_t d = d_lo;
        for (; d < d_hi; ++d) {
            const uint32_t c = P.counters[CNT_SHQ + d];
            if (i < c) break;
            i -= c;
        }
        const float4* entries = P.shadow + 4 * (size_t)d * P.depth_stride;
        const uint32_t l = (uint32_t)(w >> os) & (group - 1u);
        const uint32_t j = (uint32_t)w & ((1u << os) - 1u);
        V3 c = v3(0.f, 0.f, 0.f);
        float4 e0 = make_float4(0.f, 0.f, 0.f, 0.f), e3 = e0;
        if (w < work) {
            // the entry at position i of the key order (or of the queue order)
            const uint32_t ei = (sorted_mask >> (d - d_lo)) & 1u ? P.sperm[(size_t)d * P.depth_stride + i] : i;
            const float4* e = entries + 4 * (size_t)ei;
            e0 = e[0];
            e3 = e[3];
            if (l < P.n_lights) {
                const float4 e1 = e[1], e2 = e[2];
                const V3 ps = xyz(e0);
                const V3 n = xyz(e1);
                const Mat m = load_mat(&P.mats[__float_as_uint(e1.w) & 0x00FFFFFFu]);
                V3 contrib, dir;
                float dist;
                if (light_term(P.lights[l], ps, n, xyz(e2), m, contrib, dir, dist)) {
                    const V3 o = ps + dir * kEps;
                    bool occ = false;
                    for (uint32_t k = 0; k < P.n_analytic && !occ; ++k) {   // escenario.h:255-266
                        const float ta = analytic_t(P.analytic[k], o, dir);
                        occ = ta > 0.0f && ta < dist;
                    }
                    if (!occ) {
                        const RayPre r = make_ray(o.x, o.y, o.z, dir.x, dir.y, dir.z, 0.0f, dist);
                        float t, u, v;
                        uint32_t prim;
                        if (os)
                            occ = traverse_octet<true>(P.nodes, P.tris, top, P.n_top, r, stack + (threadIdx.x >> 3),
                                                       kBlock / 8, P.stack_depth, t, u, v, prim,
                                                       kStats ? &st : nullptr);
                        else
                            occ = traverse<0, true>(P.nodes, P.tris, top, P.n_top, r, stack + threadIdx.x, kBlock,
                                                    P.stack_depth, t, u, v, prim, kStats ? &st : nullptr);
                    }
                    if (!occ) c = contrib;
                }
            }
        }
        // ordered sum over the entry's lights (wave-uniform trip count)
        V3 sum = v3(0.f, 0.f, 0.f);
        for (uint32_t k = 0; k < group; ++k) {
            const int src = lead + (int)(k << os);
            sum = sum + v3(__shfl(c.x, src), __shfl(c.y, src), __shfl(c.z, src));
        }
        if (w < work && l == 0 && j == 0)
            P.nee[(size_t)d * P.depth_stride + __float_as_uint(e0.w)] = f4(xyz(e3) * sum, 0.0f);
    }
    if (kStats) add_stats(P, 2, st);
}

// ------------------------------------------------------------------- tail
// The frame's last depths in one launch: from depth d0 on, each path still
// in queue q is traced to its end by one aligned octet of lanes.  Per depth:
// the closest hit (traverse_octet), the shading step (shade_entry, the code
// of the shading stage), and the NEE of the request it makes, one lane per
// light (8 lights per round) with the unoccluded terms summed in light order
// by shuffles, as the shadow stage does; nee[depth][pid] and the mask are
// written as the shadow and shading stages write them.  Results are those
// of the per-depth stages bit for bit.  Late depths hold few paths (training
// paths past their head vertex, paths that survive Russian roulette), and
// one launch replaces a chain of 3 small launches per depth, each of which
// lasts as long as its slowest ray.
// Dynamic LDS: [n_top nodes][stack_depth x kBlock uint2 (per-lane shadow
// rays)][stack_depth x kBlock/8 uint2 (per-octet closest hit)].
template <int kBlock, bool kNrc, bool kStats = false>
__global__ __launch_bounds__(kBlock) void k_tail(PtParams P, uint32_t q, uint32_t d0) {
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr uint32_t kOct = kBlock / 8;
    __shared__ uint32_t n_ext[kMaxDepth], n_shd[kMaxDepth];
    float4* top = reinterpret_cast<float4*>(lds);
    uint2* lane_stack = reinterpret_cast<uint2*>(top + 5 * P.n_top);
    uint2* oct_stack = lane_stack + (size_t)P.stack_depth * kBlock;
    const uint32_t count = P.counters[CNT_QD + d0];
    const uint32_t chunks = (count + kOct - 1) / kOct;
    if (blockIdx.x >= chunks) return;   // workgroups without paths leave before staging
    for (uint32_t k = threadIdx.x; k < kMaxDepth; k += kBlock) n_ext[k] = n_shd[k] = 0u;
    stage_top_nodes(P, top);   // synchronises the block
    const uint32_t j = threadIdx.x & 7u;
    const int base = (int)(__lane_id() & ~7u);
    TraceStats st_e, st_s;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint32_t i = c * kOct + (threadIdx.x >> 3);   // uniform over the octet
        if (i >= count) continue;
        const uint32_t pid = P.qpath[q][i];
        float4 ro = P.rays[q][2 * (size_t)i];
        float4 rd = P.rays[q][2 * (size_t)i + 1];
        for (uint32_t d = d0; d < P.max_depth; ++d) {
            if (j == 0) atomicAdd(&n_ext[d], 1u);
            const RayPre r = make_ray(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, ro.w, rd.w);
            float t, u, v;
            uint32_t prim;
            traverse_octet<false>(P.nodes, P.tris, top, P.n_top, r, oct_stack + (threadIdx.x >> 3), kOct,
                                  P.stack_depth, t, u, v, prim, kStats ? &st_e : nullptr);
            for (uint32_t k = 0; k < P.n_analytic; ++k) {   // escenario.h:323-334
                const float ta = analytic_t(P.analytic[k], v3(ro.x, ro.y, ro.z), v3(rd.x, rd.y, rd.z));
                if (ta > ro.w && ta < t) {
                    t = ta;
                    u = 0.0f;
                    v = 0.0f;
                    prim = kAnalyticBit | k;
                }
            }
            bool want_shadow = false, want_ray = false;
            float4 e0, e1, e2, e3, n0, n1;
            shade_entry<kNrc>(P, d, d + 1 >= P.max_depth, pid, ro, rd, make_float4(t, u, v, __uint_as_float(prim)),
                              want_shadow, e0, e1, e2, e3, want_ray, n0, n1);
            if (want_shadow) {   // the shadow stage for this entry (escenario.h:144-184)
                const V3 ps = xyz(e0);
                const V3 n = xyz(e1);
                const Mat m = load_mat(&P.mats[__float_as_uint(e1.w) & 0x00FFFFFFu]);
                V3 sum = v3(0.f, 0.f, 0.f);
                uint32_t facing = 0;
                for (uint32_t l0 = 0; l0 < P.n_lights; l0 += 8) {
                    const uint32_t l = l0 + j;
                    V3 ch = v3(0.f, 0.f, 0.f);
                    if (l < P.n_lights) {
                        V3 contrib, dir;
                        float dist;
                        if (light_term(P.lights[l], ps, n, xyz(e2), m, contrib, dir, dist)) {
                            ++facing;
                            const V3 o = ps + dir * kEps;
                            bool occ = false;
                            for (uint32_t k = 0; k < P.n_analytic && !occ; ++k) {   // escenario.h:255-266
                                const float ta = analytic_t(P.analytic[k], o, dir);
                                occ = ta > 0.0f && ta < dist;
                            }
                            if (!occ) {
                                const RayPre rs = make_ray(o.x, o.y, o.z, dir.x, dir.y, dir.z, 0.0f, dist);
                                float ts, us, vs;
                                uint32_t prs;
                                occ = traverse<0, true>(P.nodes, P.tris, top, P.n_top, rs, lane_stack + threadIdx.x,
                                                        kBlock, P.stack_depth, ts, us, vs, prs,
                                                        kStats ? &st_s : nullptr);
                            }
                            if (!occ) ch = contrib;
                        }
                    }
                    const uint32_t nk = P.n_lights - l0 < 8u ? P.n_lights - l0 : 8u;
                    for (uint32_t k = 0; k < nk; ++k) {
                        const int src = base + (int)k;
                        sum = sum + v3(__shfl(ch.x, src), __shfl(ch.y, src), __shfl(ch.z, src));
                    }
                }
                if (facing) atomicAdd(&n_shd[d], facing);
                if (j == 0) {
                    P.nee[(size_t)d * P.depth_stride + pid] = f4(xyz(e3) * sum, 0.0f);
                    P.neemask[pid] |= 1u << d;
                }
            }
            if (!want_ray) break;
            ro = n0;
            rd = n1;
        }
    }
    if (kStats) {
        add_stats(P, 0, st_e);
        add_stats(P, 2, st_s);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kMaxDepth; k += kBlock) {
        if (n_ext[k]) atomicAdd(&P.gcounters[CNT_EXT + k], n_ext[k]);
        if (n_shd[k]) atomicAdd(&P.gcounters[CNT_SHD + k], n_shd[k]);
    }
}

// --------------------------------------------------------------- finalize
// Path radiance = emitted radiance + the NEE contributions of every depth,
// added in depth order.  Plain mode: pixel = path radiance.  Cache modes:
// training pixels keep the colour before their head (render.h:251-260):
// the emission collected before the head plus the NEE of depths up to and
// including the head's; every captured head/query vertex becomes a cache
// query; every training path with a head becomes a record with
// L_suffix = radiance - prefix colour.  Cache modes store the totals back
// into rad / pre for the drop-in export.
template <bool kNrc>
__global__ __launch_bounds__(256) void k_finalize(PtParams P, tfg_color* image, double* accum) {
    for (uint32_t b = blockIdx.x * blockDim.x; b < P.n_paths; b += gridDim.x * blockDim.x) {
        const uint32_t pid = b + threadIdx.x;
        bool want_rec = false, want_q = false;
        uint32_t fl = 0;
        V3 pix = v3(0.f, 0.f, 0.f);
        V3 tot = v3(0.f, 0.f, 0.f), pre = v3(0.f, 0.f, 0.f);
        if (pid < P.n_paths) {
            const float4 r = P.rad[pid];
            tot = xyz(r);
            bool head = false;
            uint32_t head_depth = 0;
            if (kNrc) {
                fl = P.flags[pid];
                head = (fl & PF_TRAIN) && (fl & PF_HEAD);
                head_depth = (fl >> PF_HEAD_DEPTH_SHIFT) & (kMaxDepth - 1);
                if (head) pre = xyz(P.pre[pid]);
            }
            uint32_t mask = P.neemask[pid];
            while (mask) {
                const uint32_t d = (uint32_t)__ffs(mask) - 1u;
                mask &= mask - 1u;
                const V3 c = xyz(P.nee[(size_t)d * P.depth_stride + pid]);
                tot = tot + c;
                if (head && d <= head_depth) pre = pre + c;
            }
            pix = head ? pre : tot;
            if (kNrc) {
                P.rad[pid] = f4(tot, r.w);
                if (head) P.pre[pid] = f4(pre, 0.0f);
                want_rec = head;
                want_q = head || (fl & PF_QUERY);
            }
            const size_t px = (size_t)P.pix_base + pid;
            image[px] = tfg_color{pix.x, pix.y, pix.z};
            if (!kNrc && accum) {
                accum[3 * px] += (double)pix.x;
                accum[3 * px + 1] += (double)pix.y;
                accum[3 * px + 2] += (double)pix.z;
            }
        }
        if (!kNrc) continue;
        // The block's records and queries take contiguous slots (ranked in
        // lane order), so they are written through LDS: each thread leaves
        // its path id and the fields it computed at its rank, then the block
        // copies the ranges float4 by float4, consecutive lanes writing
        // consecutive float4 (and reading the head / tail vertices of
        // consecutive paths).  The values and slots are those of writing
        // each record and query from its own thread.
        __shared__ uint32_t s_rpid[256], s_qpid[256], s_rtail[256];
        __shared__ float4 s_rx[256][3], s_qx[256];
        uint32_t rs, rb, nr, qs, qb, nq;
        block_append2_range<256>(want_rec, &P.gcounters[CNT_RECORDS], rs, rb, nr, want_q, &P.gcounters[CNT_QUERIES],
                                 qs, qb, nq);
        if (want_rec) {
            const uint32_t r = rs - rb;
            const bool tl = fl & PF_TAIL;
            s_rpid[r] = pid;
            s_rtail[r] = tl ? 1u : 0u;
            s_rx[r][0] = make_float4(tot.x - pre.x, tot.y - pre.y, tot.z - pre.z, 1.0f);
            float4 ts = tl ? P.tsuf[pid] : make_float4(0.f, 0.f, 0.f, 0.f);
            ts.w = __uint_as_float(P.pix_base + pid);   // pixel index, for record-level checks
            s_rx[r][1] = ts;
            s_rx[r][2] = P.tthr[pid];
        }
        if (want_q) {
            const uint32_t r = qs - qb;
            const float4 tt = P.tthr[pid];
            s_qpid[r] = pid;
            s_qx[r] = make_float4(tt.x, tt.y, tt.z, __uint_as_float(P.pix_base + pid));
        }
        __syncthreads();
        float4* R = P.records + (size_t)kRecF4 * rb;
        for (uint32_t j = threadIdx.x; j < nr * (uint32_t)kRecF4; j += 256) {
            const uint32_t r = j / (uint32_t)kRecF4, k = j - r * (uint32_t)kRecF4;
            const uint32_t p = s_rpid[r];
            float4 v;
            if (k < (uint32_t)kVertF4) v = P.head[(size_t)kVertF4 * p + k];
            else if (k < 2u * kVertF4)
                v = s_rtail[r] ? P.tail[(size_t)kVertF4 * p + (k - kVertF4)] : make_float4(0.f, 0.f, 0.f, 0.f);
            else v = s_rx[r][k - 2u * kVertF4];
            R[j] = v;
        }
        float4* Q = P.queries + (size_t)kQueryF4 * qb;
        for (uint32_t j = threadIdx.x; j < nq * (uint32_t)kQueryF4; j += 256) {
            const uint32_t r = j / (uint32_t)kQueryF4, k = j - r * (uint32_t)kQueryF4;
            Q[j] = k < (uint32_t)kVertF4 ? P.head[(size_t)kVertF4 * s_qpid[r] + k] : s_qx[r];
        }
        __syncthreads();   // the staging arrays are reused by the next iteration
    }
}

// image -> accum (after the cache composite)
__global__ __launch_bounds__(256) void k_accumulate(const tfg_color* image, double* accum, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_color c = image[i];
        accum[3 * (size_t)i] += (double)c.r;
        accum[3 * (size_t)i + 1] += (double)c.g;
        accum[3 * (size_t)i + 2] += (double)c.b;
    }
}

}  // namespace tfg
