// api_bvh.cpp - C ABI for CWBVH construction and upload, and the
// thread-local error message.
#include <cstring>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>
#include <algorithm>

#include "cwbvh_build.h"
#include "host_common.h"

namespace tfg {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

// Depth of every tree uploaded by this library, keyed by its node buffer,
// so the traversal entry points can refuse trees deeper than their stack.
static std::mutex g_depth_mu;
static std::unordered_map<const void*, std::pair<uint32_t, uint32_t>> g_depth;   // depth, format

uint32_t known_tree_depth(const void* d_nodes) {
    std::lock_guard<std::mutex> lk(g_depth_mu);
    auto it = g_depth.find(d_nodes);
    return it == g_depth.end() ? 0u : it->second.first;
}

// Triangle format of a tree uploaded by this library, or 0xFFFFFFFF.
uint32_t known_tree_format(const void* d_nodes) {
    std::lock_guard<std::mutex> lk(g_depth_mu);
    auto it = g_depth.find(d_nodes);
    return it == g_depth.end() ? 0xFFFFFFFFu : it->second.second;
}

void register_tree_depth(const void* d_nodes, uint32_t depth, uint32_t format = 0) {
    std::lock_guard<std::mutex> lk(g_depth_mu);
    if (depth == 0) g_depth.erase(d_nodes);
    else g_depth[d_nodes] = {depth, format};
}

// Depth of a tree this library did not build (root = 1), from a walk over a
// host copy of its nodes: the children of node i are nodes n1.x + rank, one
// per set bit of imask (byte 3 of n0.w), the node layout the traversal
// decodes (traverse_cwbvh.cuh:73-99).  0 when a child index leaves the node
// array, a node is reached twice or the tree is deeper than 64 levels.
uint32_t walk_tree_depth(const void* d_nodes, uint32_t num_nodes, int* rc) {
    *rc = TFG_OK;
    if (!d_nodes || num_nodes == 0) return 0;
    std::vector<uint32_t> words((size_t)num_nodes * 20);
    const hipError_t e = hipMemcpy(words.data(), d_nodes, words.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        *rc = hip_fail(e, "hipMemcpy(CWBVH nodes)");
        return 0;
    }
    std::vector<uint8_t> seen(num_nodes, 0);
    std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, 1u}};
    uint32_t depth = 0;
    seen[0] = 1;
    while (!todo.empty()) {
        const auto [i, d] = todo.back();
        todo.pop_back();
        if (d > 64) return 0;
        depth = std::max(depth, d);
        const uint32_t* n = &words[(size_t)i * 20];
        const uint32_t imask = n[3] >> 24, base = n[4];
        for (uint32_t slot = 0; slot < 8; ++slot) {
            if (!(imask & (1u << slot))) continue;
            const uint64_t c = (uint64_t)base + (uint32_t)__builtin_popcount(imask & ((1u << slot) - 1u));
            if (c >= num_nodes || seen[c]) return 0;
            seen[c] = 1;
            todo.push_back({(uint32_t)c, d + 1});
        }
    }
    return depth;
}

}  // namespace tfg

extern "C" {

const char* tfg_last_error(void) { return tfg::g_error.c_str(); }
int tfg_version(void) { return 1; }

int tfg_cwbvh_build(const float* verts, uint32_t n_tris, uint32_t format, tfg_cwbvh_host** out) {
    return tfg_cwbvh_build_ex(verts, n_tris, format, nullptr, out);
}

void tfg_cwbvh_default_build_options(tfg_cwbvh_build_options* o) {
    if (!o) return;
    const tfg::BuildOptions d;
    o->node_cost = (float)d.node_cost;
    o->tri_cost = (float)d.tri_cost;
    o->max_leaf = (uint32_t)d.max_leaf;
    o->bins = (uint32_t)d.bins;
}

int tfg_cwbvh_build_ex(const float* verts, uint32_t n_tris, uint32_t format, const tfg_cwbvh_build_options* opts,
                       tfg_cwbvh_host** out) {
    if (out == nullptr) { tfg::set_error("tfg_cwbvh_build: out is NULL"); return TFG_E_INVALID; }
    tfg::BuildOptions bo;
    if (opts) {
        if (opts->max_leaf < 1 || opts->max_leaf > 3 || opts->bins < 4 || opts->bins > 64 ||
            !(opts->node_cost > 0.0f) || !(opts->tri_cost > 0.0f)) {
            tfg::set_error("tfg_cwbvh_build: options out of range (max_leaf 1..3, bins 4..64, costs > 0)");
            return TFG_E_INVALID;
        }
        bo.node_cost = opts->node_cost;
        bo.tri_cost = opts->tri_cost;
        bo.max_leaf = (int)opts->max_leaf;
        bo.bins = (int)opts->bins;
    }
    *out = nullptr;
    if (verts == nullptr || n_tris == 0) { tfg::set_error("tfg_cwbvh_build: no triangles"); return TFG_E_INVALID; }
    if (format != TFG_TRI_WATERTIGHT && format != TFG_TRI_AFFINE) {
        tfg::set_error("tfg_cwbvh_build: unknown triangle format %u", format);
        return TFG_E_INVALID;
    }
    auto* h = new (std::nothrow) tfg_cwbvh_host();
    if (!h) { tfg::set_error("tfg_cwbvh_build: out of host memory"); return TFG_E_NOMEM; }
    if (!tfg::build_cwbvh(verts, n_tris, format, h->b, bo)) {
        delete h;
        tfg::set_error("tfg_cwbvh_build: invalid input (non-finite vertex or too many triangles)");
        return TFG_E_INVALID;
    }
    if (h->b.max_depth > TFG_MAX_TREE_DEPTH) {
        const uint32_t d = h->b.max_depth;
        delete h;
        tfg::set_error("tfg_cwbvh_build: tree depth %u exceeds %d", d, TFG_MAX_TREE_DEPTH);
        return TFG_E_LIMIT;
    }
    *out = h;
    return TFG_OK;
}

int tfg_cwbvh_host_info(const tfg_cwbvh_host* h, uint32_t* num_nodes, uint32_t* num_tris, uint32_t* max_depth,
                        double* sah_cost) {
    if (!h) { tfg::set_error("tfg_cwbvh_host_info: NULL handle"); return TFG_E_INVALID; }
    if (num_nodes) *num_nodes = h->b.num_nodes;
    if (num_tris) *num_tris = h->b.num_tris;
    if (max_depth) *max_depth = h->b.max_depth;
    if (sah_cost) *sah_cost = h->b.sah_cost;
    return TFG_OK;
}

const float* tfg_cwbvh_host_nodes(const tfg_cwbvh_host* h) { return h ? h->b.nodes.data() : nullptr; }
const float* tfg_cwbvh_host_tris(const tfg_cwbvh_host* h) { return h ? h->b.tris.data() : nullptr; }
void tfg_cwbvh_host_free(tfg_cwbvh_host* h) { delete h; }

int tfg_cwbvh_upload(const tfg_cwbvh_host* h, tfg_cwbvh_gpu* out) {
    if (!h || !out) { tfg::set_error("tfg_cwbvh_upload: NULL argument"); return TFG_E_INVALID; }
    std::memset(out, 0, sizeof(*out));
    const size_t nb = h->b.nodes.size() * sizeof(float);
    const size_t tb = h->b.tris.size() * sizeof(float);
    void* dn = nullptr;
    void* dt = nullptr;
    hipError_t e = hipMalloc(&dn, nb);
    if (e != hipSuccess) return tfg::hip_fail(e, "hipMalloc(nodes)");
    e = hipMalloc(&dt, tb);
    if (e != hipSuccess) { (void)hipFree(dn); return tfg::hip_fail(e, "hipMalloc(tris)"); }
    e = hipMemcpy(dn, h->b.nodes.data(), nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, h->b.tris.data(), tb, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dn);
        (void)hipFree(dt);
        return tfg::hip_fail(e, "hipMemcpy(cwbvh)");
    }
    out->d_nodos = dn;
    out->d_tris = dt;
    out->num_nodos = h->b.num_nodes;
    out->num_tris = h->b.num_tris;
    tfg::register_tree_depth(dn, h->b.max_depth, h->b.format);
    return TFG_OK;
}

int tfg_cwbvh_gpu_download(const tfg_cwbvh_gpu* g, float* nodes, float* tris) {
    if (!g || !nodes || !tris || !g->d_nodos || !g->d_tris) {
        tfg::set_error("tfg_cwbvh_gpu_download: NULL argument");
        return TFG_E_INVALID;
    }
    TFG_HIP_CHECK(hipDeviceSynchronize());
    TFG_HIP_CHECK(hipMemcpy(nodes, g->d_nodos, (size_t)g->num_nodos * 80, hipMemcpyDeviceToHost));
    TFG_HIP_CHECK(hipMemcpy(tris, g->d_tris, (size_t)g->num_tris * 64, hipMemcpyDeviceToHost));
    return TFG_OK;
}

int tfg_cwbvh_gpu_free(tfg_cwbvh_gpu* g) {
    if (!g) { tfg::set_error("tfg_cwbvh_gpu_free: NULL"); return TFG_E_INVALID; }
    tfg::register_tree_depth(g->d_nodos, 0);
    hipError_t e1 = g->d_nodos ? hipFree(g->d_nodos) : hipSuccess;
    hipError_t e2 = g->d_tris ? hipFree(g->d_tris) : hipSuccess;
    g->d_nodos = g->d_tris = nullptr;
    g->num_nodos = g->num_tris = 0;
    if (e1 != hipSuccess) return tfg::hip_fail(e1, "hipFree(nodes)");
    if (e2 != hipSuccess) return tfg::hip_fail(e2, "hipFree(tris)");
    return TFG_OK;
}

}  // extern "C"
