// refit.hip - GPU refit of an uploaded CWBVH to moved vertices.
//
// The tree is kept (nodes, slots, triangle records and their order); the
// triangle records are rewritten from the new vertices and every node's
// quantisation frame and child bounds are recomputed bottom-up, one launch
// per tree level (nodes are stored breadth-first, so each level is a
// contiguous index range and a node's children lie in the next level).
// The box and quantisation arithmetic is the host builder's
// (cwbvh_build.cpp), done in the same order in FP64 with contraction off,
// so refitting with unchanged vertices reproduces the built nodes and
// records bit for bit.  The reference rebuilds its tree instead
// (tinybvh_wrapper.cpp:53); this is the animated-geometry path next to it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

#include "../../include/tfg.h"
#include "cwbvh_build.h"
#include "host_common.h"

namespace tfg {
namespace {

constexpr int kBlock = 256;

struct FBox {
    float mn[3], mx[3];
};

// std::min / std::max as the host builder applies them (first argument kept
// unless the second compares strictly smaller / larger)
__device__ __forceinline__ float hmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float hmax(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ void grow(FBox& b, const float* p) {
    for (int a = 0; a < 3; ++a) {
        b.mn[a] = hmin(b.mn[a], p[a]);
        b.mx[a] = hmax(b.mx[a], p[a]);
    }
}

__device__ __forceinline__ void grow(FBox& b, const FBox& c) {
    for (int a = 0; a < 3; ++a) {
        b.mn[a] = hmin(b.mn[a], c.mn[a]);
        b.mx[a] = hmax(b.mx[a], c.mx[a]);
    }
}

__device__ __forceinline__ FBox empty_box() {
    FBox b;
    for (int a = 0; a < 3; ++a) {
        b.mn[a] = __int_as_float(0x7f800000);
        b.mx[a] = __int_as_float(0xff800000);
    }
    return b;
}

// Smallest e (clamped to -100..127) with 255 * 2^e >= extent; the host's
// quant_exponent gives the same value (its ceil(log2) start and this
// frexp start are both <= the smallest such e, and both then step up).
__device__ __forceinline__ int quant_exponent(double extent) {
    if (!(extent > 0.0)) return -100;
    int x;
    const double m = frexp(extent / 255.0, &x);   // extent / 255 = m * 2^x, m in [0.5, 1)
    int e = m == 0.5 ? x - 1 : x;
    while (ldexp(255.0, e) < extent) ++e;
    return e < -100 ? -100 : (e > 127 ? 127 : e);
}

// The float below f (std::nextafter(f, -inf) for finite f)
__device__ __forceinline__ float float_below(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) == 0u) return __uint_as_float(0x80000001u);
    return __uint_as_float(f > 0.0f ? u - 1u : u + 1u);
}

// Triangle records from the new vertices (write_tri of the builder).
__global__ __launch_bounds__(kBlock) void k_refit_tris(float4* __restrict__ tris, const float* __restrict__ verts,
                                                       uint32_t n, uint32_t format) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float4* T = tris + 4 * (size_t)i;
        const uint32_t prim = __float_as_uint(T[3].w);
        const float* a = verts + 9 * (size_t)prim;
        const float* b = a + 3;
        const float* c = a + 6;
        if (format == TRI_WATERTIGHT) {
            T[0] = make_float4(a[0], a[1], a[2], 0.f);
            T[1] = make_float4(b[0], b[1], b[2], 0.f);
            T[2] = make_float4(c[0], c[1], c[2], 0.f);
            T[3] = make_float4(0.f, 0.f, 0.f, __uint_as_float(prim));
            continue;
        }
        // affine world -> unit-triangle transform: rows of inverse([e1 e2 n])
        double e1[3], e2[3], nn[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = (double)b[k] - a[k];
            e2[k] = (double)c[k] - a[k];
        }
        nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double r0[3], r1[3];
        r0[0] = e2[1] * nn[2] - e2[2] * nn[1];
        r0[1] = e2[2] * nn[0] - e2[0] * nn[2];
        r0[2] = e2[0] * nn[1] - e2[1] * nn[0];
        r1[0] = nn[1] * e1[2] - nn[2] * e1[1];
        r1[1] = nn[2] * e1[0] - nn[0] * e1[2];
        r1[2] = nn[0] * e1[1] - nn[1] * e1[0];
        const double det = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
        float o[12];
        if (!(det > 0.0)) {
            for (int k = 0; k < 12; ++k) o[k] = 0.f;
            o[11] = 1.f;
        } else {
            const double* rows[3] = {r0, r1, nn};
            for (int r = 0; r < 3; ++r) {
                double w = 0.0;
                for (int k = 0; k < 3; ++k) {
                    o[4 * r + k] = (float)(rows[r][k] / det);
                    w -= (rows[r][k] / det) * a[k];
                }
                o[4 * r + 3] = (float)w;
            }
        }
        T[0] = make_float4(o[0], o[1], o[2], o[3]);
        T[1] = make_float4(o[4], o[5], o[6], o[7]);
        T[2] = make_float4(o[8], o[9], o[10], o[11]);
        T[3] = make_float4(a[0], a[1], a[2], __uint_as_float(prim));
    }
}

// One level of nodes [lo, hi): child boxes (internal children from `boxes`,
// written by the previous launch; leaf slots from their triangles), the
// node's box, its quantisation frame and quantised child bounds.
__global__ __launch_bounds__(kBlock) void k_refit_level(float4* __restrict__ nodes, const float4* __restrict__ tris,
                                                        const float* __restrict__ verts, FBox* __restrict__ boxes,
                                                        uint32_t lo, uint32_t hi) {
    const uint32_t i = lo + blockIdx.x * kBlock + threadIdx.x;
    if (i >= hi) return;
    float4* nd = nodes + 5 * (size_t)i;
    const float4 n0 = nd[0], n1 = nd[1];
    const uint32_t imask = __float_as_uint(n0.w) >> 24;
    const uint32_t child_base = __float_as_uint(n1.x);
    const uint32_t tri_base = __float_as_uint(n1.y) / 4u;
    const uint32_t m0 = __float_as_uint(n1.z), m1 = __float_as_uint(n1.w);
    FBox cb[8];
    bool has[8];
    FBox nb = empty_box();
    for (int s = 0; s < 8; ++s) {
        const uint32_t meta = ((s < 4 ? m0 : m1) >> (8 * (s & 3))) & 0xFFu;
        has[s] = meta != 0u;
        if (!has[s]) continue;
        if (imask & (1u << s)) {
            cb[s] = boxes[child_base + __popc(imask & ((1u << s) - 1u))];
        } else {
            cb[s] = empty_box();
            const uint32_t off = meta & 31u, cnt = (uint32_t)__popc(meta >> 5);
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t prim = __float_as_uint(tris[4 * (size_t)(tri_base + off + k) + 3].w);
                const float* v = verts + 9 * (size_t)prim;
                FBox tb = empty_box();
                grow(tb, v);
                grow(tb, v + 3);
                grow(tb, v + 6);
                grow(cb[s], tb);
            }
        }
        grow(nb, cb[s]);
    }
    // quantisation frame: every child box widened by 2^-18 of the node's
    // magnitude, then rounded outward
    float p[3];
    int e[3];
    double scale[3], margin[3];
    for (int a = 0; a < 3; ++a) {
        const double l = nb.mn[a], h = nb.mx[a];
        const double mag = fmax(fmax(fabs(l), fabs(h)), h - l);
        margin[a] = ldexp(fmax(mag, 1e-20), -18);
        p[a] = (float)(l - margin[a]);
        if ((double)p[a] > l - margin[a]) p[a] = float_below(p[a]);
        e[a] = quant_exponent(h + margin[a] - (double)p[a]);
        scale[a] = ldexp(1.0, e[a]);
    }
    uint32_t q[6][2] = {{0u, 0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}};   // qlo x,y,z, qhi x,y,z
    for (int s = 0; s < 8; ++s) {
        if (!has[s]) continue;
        for (int a = 0; a < 3; ++a) {
            double ql = floor(((double)cb[s].mn[a] - margin[a] - p[a]) / scale[a]);
            double qh = ceil(((double)cb[s].mx[a] + margin[a] - p[a]) / scale[a]);
            ql = fmin(fmax(ql, 0.0), 255.0);
            qh = fmin(fmax(qh, 0.0), 255.0);
            q[a][s >> 2] |= (uint32_t)ql << (8 * (s & 3));
            q[3 + a][s >> 2] |= (uint32_t)qh << (8 * (s & 3));
        }
    }
    const uint32_t ex = (uint32_t)(uint8_t)(int8_t)e[0] | ((uint32_t)(uint8_t)(int8_t)e[1] << 8) |
                        ((uint32_t)(uint8_t)(int8_t)e[2] << 16) | (imask << 24);
    nd[0] = make_float4(p[0], p[1], p[2], __uint_as_float(ex));
    nd[2] = make_float4(__uint_as_float(q[0][0]), __uint_as_float(q[0][1]), __uint_as_float(q[1][0]),
                        __uint_as_float(q[1][1]));
    nd[3] = make_float4(__uint_as_float(q[2][0]), __uint_as_float(q[2][1]), __uint_as_float(q[3][0]),
                        __uint_as_float(q[3][1]));
    nd[4] = make_float4(__uint_as_float(q[4][0]), __uint_as_float(q[4][1]), __uint_as_float(q[5][0]),
                        __uint_as_float(q[5][1]));
    boxes[i] = nb;
}

}  // namespace
}  // namespace tfg

struct tfg_cwbvh_refitter {
    std::vector<uint32_t> level_start;   // level k: nodes [level_start[k], level_start[k + 1])
    uint32_t num_nodes = 0, num_tris = 0, format = 0;
    tfg::FBox* boxes = nullptr;          // device, one per node
};

extern "C" {

int tfg_cwbvh_refit_create(const tfg_cwbvh_host* h, tfg_cwbvh_refitter** out) {
    if (!h || !out) { tfg::set_error("tfg_cwbvh_refit_create: NULL argument"); return TFG_E_INVALID; }
    *out = nullptr;
    const uint32_t nn = h->b.num_nodes, nt = h->b.num_tris;
    const float* nodes = h->b.nodes.data();
    // levels by a breadth-first walk; the builder stores nodes level by level
    std::vector<uint32_t> level(nn, 0xFFFFFFFFu);
    level[0] = 0;
    uint32_t deepest = 0;
    for (uint32_t i = 0; i < nn; ++i) {
        if (level[i] == 0xFFFFFFFFu || (i > 0 && level[i] < level[i - 1])) {
            tfg::set_error("tfg_cwbvh_refit_create: nodes are not stored level by level");
            return TFG_E_INVALID;
        }
        uint32_t w, base, tbase4, meta[2];
        std::memcpy(&w, nodes + 20 * (size_t)i + 3, 4);
        std::memcpy(&base, nodes + 20 * (size_t)i + 4, 4);
        std::memcpy(&tbase4, nodes + 20 * (size_t)i + 5, 4);
        std::memcpy(meta, nodes + 20 * (size_t)i + 6, 8);
        const uint32_t imask = w >> 24;
        for (int s = 0; s < 8; ++s) {   // leaf slots address records inside the tree
            const uint32_t m = (meta[s >> 2] >> (8 * (s & 3))) & 0xFFu;
            if (m == 0u || (imask & (1u << s))) continue;
            if (tbase4 / 4u + (m & 31u) + (uint32_t)__builtin_popcount(m >> 5) > nt) {
                tfg::set_error("tfg_cwbvh_refit_create: leaf records out of range");
                return TFG_E_INVALID;
            }
        }
        for (uint32_t k = 0; k < (uint32_t)__builtin_popcount(imask); ++k) {
            if (base + k >= nn) {
                tfg::set_error("tfg_cwbvh_refit_create: child index out of range");
                return TFG_E_INVALID;
            }
            level[base + k] = level[i] + 1;
        }
        deepest = level[i] > deepest ? level[i] : deepest;
    }
    auto* r = new (std::nothrow) tfg_cwbvh_refitter();
    if (!r) { tfg::set_error("tfg_cwbvh_refit_create: out of memory"); return TFG_E_LIMIT; }
    r->level_start.assign(deepest + 2, nn);
    for (uint32_t i = nn; i-- > 0;) r->level_start[level[i]] = i;
    r->num_nodes = nn;
    r->num_tris = nt;
    r->format = h->b.format;
    const hipError_t e = hipMalloc(&r->boxes, sizeof(tfg::FBox) * (size_t)(nn ? nn : 1));
    if (e != hipSuccess) {
        delete r;
        return tfg::hip_fail(e, "hipMalloc(refit boxes)");
    }
    *out = r;
    return TFG_OK;
}

int tfg_cwbvh_refit(tfg_cwbvh_refitter* r, tfg_cwbvh_gpu* bvh, const float* d_verts, uint32_t n_tris, tfg_stream stream) {
    if (!r || !bvh || !d_verts || !bvh->d_nodos || !bvh->d_tris) {
        tfg::set_error("tfg_cwbvh_refit: NULL argument");
        return TFG_E_INVALID;
    }
    if (bvh->num_nodos != r->num_nodes || bvh->num_tris != r->num_tris || n_tris != r->num_tris) {
        tfg::set_error("tfg_cwbvh_refit: the tree has %u nodes / %u triangles, got %u / %u (and %u vertices' "
                       "triangles)", r->num_nodes, r->num_tris, bvh->num_nodos, bvh->num_tris, n_tris);
        return TFG_E_INVALID;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    float4* nodes = static_cast<float4*>(bvh->d_nodos);
    float4* tris = static_cast<float4*>(bvh->d_tris);
    const uint32_t tgrid = (r->num_tris + tfg::kBlock - 1) / tfg::kBlock;
    if (tgrid) {
        hipLaunchKernelGGL(tfg::k_refit_tris, dim3(tgrid < 65535u ? tgrid : 65535u), dim3(tfg::kBlock), 0, s, tris,
                           d_verts, r->num_tris, r->format);
        TFG_HIP_CHECK(hipGetLastError());
    }
    for (size_t k = r->level_start.size() - 1; k-- > 0;) {
        const uint32_t lo = r->level_start[k], hi = r->level_start[k + 1];
        if (hi <= lo) continue;
        hipLaunchKernelGGL(tfg::k_refit_level, dim3((hi - lo + tfg::kBlock - 1) / tfg::kBlock), dim3(tfg::kBlock), 0,
                           s, nodes, tris, d_verts, r->boxes, lo, hi);
        TFG_HIP_CHECK(hipGetLastError());
    }
    return TFG_OK;
}

int tfg_cwbvh_refit_destroy(tfg_cwbvh_refitter* r) {
    if (!r) return TFG_OK;
    const hipError_t e = r->boxes ? hipFree(r->boxes) : hipSuccess;
    delete r;
    return e == hipSuccess ? TFG_OK : tfg::hip_fail(e, "hipFree(refit boxes)");
}

}  // extern "C"
