// pt_internal.h - device-side layout of the wavefront path tracer's state,
// shared by pt.hip (generate / extend / shade / connect / finalize) and
// nrc.hip (radiance-cache inference, targets and training).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tfg_types.h"

namespace tfg {

constexpr uint32_t kMaxDepth = 32;
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kAnalyticBit = 0x80000000u;
constexpr uint32_t kMaxAnalytic = 64;
constexpr uint32_t kMaxLights = 64;   // one wave64 lane per light of a shadow entry

// Counter blocks (uint32, device).  The frame block holds the frame-wide
// counts; each pixel tile has its own block for its queue lengths.  Both use
// this layout; a block uses only its own entries.
enum : uint32_t {
    CNT_RESERVED = 2,
    CNT_RECORDS = 3,      // frame: training records written this frame
    CNT_QUERIES = 4,      // frame: cache queries written this frame
    CNT_EXT = 8,          // frame, + depth: extension rays traced at that depth
    CNT_SHD = 8 + kMaxDepth,   // frame, + depth: shadow rays traced at that depth
    CNT_TOTAL = 8 + 2 * kMaxDepth,
    CNT_CUM = 80,         // frame: 2 x uint64, rays of all previous frames
    CNT_STAT = 84,        // frame: 4 x uint64, extend nodes, extend tris, shadow nodes, shadow tris
    CNT_SHQ = 96,         // tile, + depth: shadow queue length of that depth
    // tile, + depth: length of the ray queue traced at that depth (the ray
    // buffers alternate, the counts do not: a launch that appends to the next
    // depth's queue never shares its counter with a launch that reads one)
    CNT_QD = CNT_SHQ + kMaxDepth,
    CNT_ALLOC = 192,      // >= CNT_QD + kMaxDepth + 1
};
constexpr uint32_t kMaxTiles = 4;   // pixel tiles traced as independent stream chains
constexpr uint32_t kMaxViews = 64;  // cameras sharing one canvas (tfg_pt_set_views)
constexpr uint32_t kMaxSide = 4;    // side streams per tile for the shadow-ray stage

// path flags
enum : uint32_t {
    PF_TRAIN = 1u << 0,      // training path (render.h:443)
    PF_UNBIASED = 1u << 1,   // unbiased suffix (render.h:65)
    PF_HEAD = 1u << 2,       // head vertex captured (render.h:191-201)
    PF_QUERY = 1u << 3,      // cache query captured (render.h:182-190)
    PF_TAIL = 1u << 4,       // suffix ended at a surface: tail data set
};
constexpr uint32_t PF_HEAD_DEPTH_SHIFT = 8;   // flags bits 8..12: depth of the head vertex

// frame mode bits
enum : uint32_t {
    MODE_INFER = 1u << 0,   // usar_red_inferencia
    MODE_TRAIN = 1u << 1,   // entrenar_red
};

// One view of a multi-view canvas: its pixels are canvas pixels
// base .. base + width * height - 1, row-major.
struct PtView {
    uint32_t base, width, height, pad;
    tfg_camera cam;
};

struct PtParams {
    // scene
    const float4* nodes;
    const float4* tris;
    uint32_t n_top;          // leading nodes staged in LDS by the traversal stages
    int stack_depth;         // LDS stack entries per lane (tree depth - 1)
    uint32_t xcd_group;      // rays per XCD super-chunk of the traversal stages (0: plain grid stride)
    uint32_t octet_below;    // traversal launches with at most this many rays trace 8 lanes per ray
    const tfg_tri_shading* shading;
    const tfg_material* mats;
    const tfg_point_light* lights;
    uint32_t n_lights;
    uint32_t light_shift;    // lanes per shadow entry = 2^light_shift >= n_lights (<= 64)
    const tfg_analytic* analytic;
    uint32_t n_analytic;
    const tfg_camera* cam;
    const PtView* views;     // n_views > 0: the canvas holds these views, cam unused
    uint32_t n_views;
    const int32_t* frame;   // device frame index (seed), read by generate
    uint32_t width, height;
    uint32_t n_paths;        // paths of this tile (path ids 0..n_paths-1)
    uint32_t pix_base;       // pixel index of the tile's path 0
    uint32_t depth_stride;   // pixels of the frame: stride of the per-depth shadow / nee arrays
    uint32_t max_depth, rr_depth, mode;
    float train_frac;
    // queues
    float4* rays[2];        // 2 float4 per ray (tfg_ray)
    uint32_t* qpath[2];
    float4* hits;           // 1 float4 per queue entry
    float4* shadow;         // per depth: depth_stride entries of 4 float4 (depth-major)
    uint32_t* counters;     // this tile's counter block
    uint32_t* gcounters;    // the frame counter block
    // path state (indexed by path id = pixel)
    float4* thr;            // throughput rgb, pdf of the last bounce
    float4* rad;            // emitted radiance rgb (totals after finalize), current ior
    float4* nee;            // per depth: depth_stride NEE contributions (rgb), written by connect
    uint32_t* neemask;      // bit d: the path has a NEE contribution at depth d
    uint32_t* rng;
    uint32_t* flags;
    float4* foot;           // a0, a_current, a_suffix, a_prefix (render.h:50-54)
    float4* pre;            // colour before the head vertex's suffix (rgb)
    float4* tthr;           // throughput x reflectance at the head / query vertex
    float4* head;           // 4 float4: head (training) or query vertex data
    float4* tail;           // 4 float4: tail vertex data
    float4* tsuf;           // suffix throughput x tail reflectance
    // cache I/O (frame-wide, appended through the frame counter block;
    // the path ids they carry are pixel indices)
    float4* records;        // kRecF4 float4 per record
    float4* queries;        // kQueryF4 float4 per query
    // spatial order of the shadow entries (depths >= sort_from; 0: queue order)
    uint32_t sort_from;
    uint16_t* skey;         // per depth: depth_stride keys (kSortBits-bit Morton code of the shadow point)
    uint32_t* sperm;        // per depth: depth_stride entry indices in key order
    uint32_t* sort_scr;     // per depth: kSortScr words (per-workgroup bin counts, bin totals, bin bases)
    // direction / origin order of the extension rays (depths >= ray_sort_from; 0: queue order)
    uint32_t ray_sort_from;
    uint16_t* rkey;         // key of each ray of the queue being built (written by shade)
    uint32_t* rperm;        // queue slots in key order (for the next closest-hit launch)
    uint32_t* ray_scr;      // kSortScr words
    // lane refill in the closest-hit stage (depths >= refill_from; 0: off)
    uint32_t refill_from;
    uint32_t refill_idle;   // a wave refills once at least this many of its lanes are idle (1..64)
    uint32_t* rcur;         // per depth: kRefillCursors cursors, kRefillCurStride words apart (zeroed by generate)
};

// Lane refill: per depth, kRefillCursors cursors (kRefillCurStride words
// apart, one 64-byte line each) deal out the queue in batches of
// kRefillBatch positions, super-chunks of kRefillGroup batches interleaved
// over the cursors.
constexpr uint32_t kRefillCursors = 32;
constexpr uint32_t kRefillCurStride = 16;
constexpr uint32_t kRefillBatch = 64;
constexpr uint32_t kRefillGroup = 16;
constexpr size_t kRefillCurWords = (size_t)kMaxDepth * kRefillCursors * kRefillCurStride;

// Counting sort of a depth's shadow entries by the Morton code of their point
// in the root node's quantisation frame (kSortBits / 3 bits per axis), so that
// the rays a wave traces start close together.
constexpr uint32_t kSortBits = 12;
constexpr uint32_t kSortBins = 1u << kSortBits;
constexpr uint32_t kSortBlocks = 128;   // workgroups (of 512 threads) of the count and scatter passes
// counts [kSortBlocks][kSortBins], bin totals, bin bases, a done counter (zero between sorts)
constexpr size_t kSortScr = (size_t)kSortBins * kSortBlocks + 2 * kSortBins + 64;
// Depths whose queue holds fewer entries than 1/16 of the tile's paths keep
// the queue order (the sort's passes cost more there than the order saves);
// the sort passes and the stage that reads the queue read the same device
// count to decide (shadow_sorted, rays_sorted in pt_kernels.h).

// Vertex data packing (4 float4): (pos, roughness), (dir mapped to [0,1], kd.r),
// (normal mapped to [0,1], kd.g), (kd.b, ks.rgb).  Mirrors DatosGeometricos
// (mlp_types.h:32-38) plus the roughness input the GGX scenes need.
constexpr int kVertF4 = 4;
// Training record: head (4), tail (4), L_suffix (1, w = 1 valid), T_suffix (1),
// normaliser (1).  Mirrors RegistroEntrenamiento (mlp_types.h:41-53).
constexpr int kRecF4 = 11;
// Query: vertex (4) + (throughput rgb, path id as bits).
constexpr int kQueryF4 = 5;

}  // namespace tfg
