// cwbvh_build.h - host builder for the 8-wide compressed BVH (CWBVH).
//
// Replaces tinybvh::BVH8_CWBVH::Build as called from tinybvh_wrapper.cpp:53.
// The output follows the node/triangle contract decoded by
// traverse_cwbvh.cuh:51-257 (80 B nodes: quantisation origin, per-axis
// power-of-two exponents, imask, child/triangle bases, 8 meta bytes and
// 6 x 8 quantised child bounds; 64 B triangle records addressed as
// triBase + 4*k with the primitive index in the last word).
#pragma once
#include <cstdint>
#include <vector>

namespace tfg {

// Triangle record encodings (64 B each).
enum TriFormat : uint32_t {
    // T0..T2 = vertices v0, v1, v2 (w = 0), T3 = (0, 0, 0, primIdx).
    // Intersected with the watertight test (Woop, Benthin, Wald 2013).
    TRI_WATERTIGHT = 0,
    // T0..T2 = rows of the affine world -> unit-triangle transform,
    // T3 = (v0, primIdx): the layout traverse_cwbvh.cuh:189-209 decodes.
    TRI_AFFINE = 1,
};

struct CwbvhHost {
    std::vector<float> nodes;   // 20 floats per node
    std::vector<float> tris;    // 16 floats per triangle record
    uint32_t num_nodes = 0;
    uint32_t num_tris = 0;      // triangle records (== input triangles)
    uint32_t max_depth = 0;     // deepest node level (root = 1)
    uint32_t num_leaf_slots = 0;
    uint32_t format = TRI_WATERTIGHT;
    double sah_cost = 0.0;      // SAH cost of the 8-wide tree (node=1, tri=1)
};

// Defaults measured on MI355X (traversal microbenchmark and the 1080p path
// tracer): equal node and triangle costs and at most 2 triangles per leaf
// slot beat the classic 1 / 0.3 with 3-triangle leaves by 10-20%.
struct BuildOptions {
    double node_cost = 1.0;   // SAH cost of a wide-node visit
    double tri_cost = 1.0;    // SAH cost of one triangle test
    int max_leaf = 2;         // triangles per leaf slot (1..3, the format's limit)
    int bins = 32;            // SAH bins per axis (4..64)
};

// verts: 9 floats per triangle (v0, v1, v2).  Returns false on bad input.
bool build_cwbvh(const float* verts, uint32_t n_tris, uint32_t format, CwbvhHost& out,
                 const BuildOptions& opt = BuildOptions());

}  // namespace tfg

// The C ABI's opaque host tree (tfg_cwbvh_build / tfg_cwbvh_refit_create).
struct tfg_cwbvh_host {
    tfg::CwbvhHost b;
};
