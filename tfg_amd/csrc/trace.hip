// trace.hip - batched CWBVH closest-hit and any-hit kernels and their C ABI.
#include <hip/hip_runtime.h>

#include "host_common.h"
#include "trace_device.h"
#include "wave_ops.h"

namespace tfg {
uint32_t known_tree_depth(const void* d_nodes);

namespace {

// Measured on MI355X (1M-triangle microbenchmark): one short-lived
// workgroup of 64 rays per chunk staging the root and its 8 children beats
// 128-ray workgroups staging 73 nodes by ~18%: every workgroup pays its
// staging, and the upper levels are served from L2 anyway.
constexpr int kBlock = 64;
constexpr uint32_t kTopNodes = 9;

// Dynamic LDS: [n_top nodes x 5 float4][stack_depth x kBlock uint2].
template <int kFormat, bool kAnyHit, bool kStats>
__global__ __launch_bounds__(kBlock) void k_trace(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                                  const float4* __restrict__ rays, float4* __restrict__ hits,
                                                  uint32_t* __restrict__ occluded, uint32_t* __restrict__ counts,
                                                  uint32_t n, uint32_t n_top, int stack_depth) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* top = reinterpret_cast<float4*>(lds);
    uint2* stack = reinterpret_cast<uint2*>(top + 5 * n_top);
    for (uint32_t k = threadIdx.x; k < 5 * n_top; k += kBlock) top[k] = nodes[k];
    __syncthreads();
    // one chunk of kBlock rays per workgroup, placed XCD-aware
    for (ChunkIter it(gridDim.x, 4096 / kBlock); it.valid(); it.next()) {
        const uint32_t i = it.chunk * kBlock + threadIdx.x;
        if (i >= n) continue;
        const float4 a = rays[2 * (size_t)i];
        const float4 b = rays[2 * (size_t)i + 1];
        const RayPre r = make_ray(a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w);
        TraceStats st;
        float t, u, v;
        uint32_t prim;
        // byte-converted bounds: at 80 registers (6 waves per SIMD) this kernel runs faster than with the
        // FP16-subnormal form, which needs 86 (5 waves)
        const bool hit = traverse<kFormat, kAnyHit, false>(nodes, tris, top, n_top, r, stack + threadIdx.x, kBlock,
                                                   stack_depth, t, u, v, prim, kStats ? &st : nullptr);
        if (kAnyHit) {
            occluded[i] = hit ? 1u : 0u;
        } else {
            hits[i] = make_float4(t, u, v, __uint_as_float(prim));
        }
        if (kStats) {
            counts[2 * (size_t)i] = st.nodes | (st.overflow << 31);
            counts[2 * (size_t)i + 1] = st.tris;
        }
    }
}

template <bool kAnyHit>
int launch(const tfg_cwbvh_gpu* bvh, uint32_t format, const tfg_ray* rays, void* out, uint32_t* counts, uint32_t n,
           tfg_stream stream, const char* name) {
    if (!bvh || !rays || !out) { set_error("%s: NULL argument", name); return TFG_E_INVALID; }
    if (format > 1) { set_error("%s: unknown triangle format %u", name, format); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    if (!bvh->d_nodos || !bvh->d_tris || bvh->num_nodos == 0) {
        set_error("%s: empty CWBVH", name);
        return TFG_E_INVALID;
    }
    const uint32_t depth = known_tree_depth(bvh->d_nodos);
    if (depth > TFG_MAX_TREE_DEPTH) {
        set_error("%s: tree depth %u exceeds %d", name, depth, TFG_MAX_TREE_DEPTH);
        return TFG_E_LIMIT;
    }
    const int stack_depth = depth ? (int)(depth > 1 ? depth - 1 : 1) : TFG_MAX_TREE_DEPTH - 1;
    const uint32_t n_top = bvh->num_nodos < kTopNodes ? bvh->num_nodos : kTopNodes;
    const size_t lds = (size_t)n_top * 80 + (size_t)stack_depth * kBlock * 8;
    const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    const float4* nd = static_cast<const float4*>(bvh->d_nodos);
    const float4* tr = static_cast<const float4*>(bvh->d_tris);
    const float4* ry = reinterpret_cast<const float4*>(rays);
    float4* h = kAnyHit ? nullptr : static_cast<float4*>(out);
    uint32_t* o = kAnyHit ? static_cast<uint32_t*>(out) : nullptr;
#define TFG_LAUNCH(F, S) \
    hipLaunchKernelGGL((k_trace<F, kAnyHit, S>), grid, block, lds, stream, nd, tr, ry, h, o, counts, n, n_top, stack_depth)
    if (format == 0) {
        if (counts) TFG_LAUNCH(0, true); else TFG_LAUNCH(0, false);
    } else {
        if (counts) TFG_LAUNCH(1, true); else TFG_LAUNCH(1, false);
    }
#undef TFG_LAUNCH
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

}  // namespace
}  // namespace tfg

extern "C" {

int tfg_trace_closest(const tfg_cwbvh_gpu* bvh, uint32_t format, const tfg_ray* d_rays, tfg_hit* d_hits,
                      uint32_t* d_counts, uint32_t n, tfg_stream stream) {
    return tfg::launch<false>(bvh, format, d_rays, d_hits, d_counts, n, stream, "tfg_trace_closest");
}

int tfg_trace_any(const tfg_cwbvh_gpu* bvh, uint32_t format, const tfg_ray* d_rays, uint32_t* d_occluded,
                  uint32_t* d_counts, uint32_t n, tfg_stream stream) {
    return tfg::launch<true>(bvh, format, d_rays, d_occluded, d_counts, n, stream, "tfg_trace_any");
}

}  // extern "C"
