/*
 * tfg.h - C ABI of libtfg_hip.so, the MI355X (gfx950) implementation of the
 * reference path tracer's hot path.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative TFG_E_* code on failure; the message of
 *     the last failure on the calling thread is tfg_last_error();
 *   - device pointers are caller-owned unless an entry point says it
 *     allocates; host pointers are read during the call only;
 *   - `stream` may be NULL (the default stream); launches are asynchronous
 *     unless stated otherwise.
 */
#ifndef TFG_H
#define TFG_H

#include <stddef.h>
#include <stdint.h>
#include "tfg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tfg_stream; /* same type as hipStream_t */

enum {
    TFG_OK = 0,
    TFG_E_INVALID = -1,   /* bad argument */
    TFG_E_HIP = -2,       /* HIP runtime error */
    TFG_E_NOMEM = -3,     /* allocation failed */
    TFG_E_LIMIT = -4,     /* input exceeds a documented limit */
    TFG_E_STATE = -5,     /* object not ready for this call */
};

/* Triangle record formats of the 64 B CWBVH triangle array. */
enum {
    TFG_TRI_WATERTIGHT = 0, /* T0..T2 = v0, v1, v2; T3.w = primIdx            */
    TFG_TRI_AFFINE = 1,     /* T0..T2 = affine unit-triangle rows; T3 = (v0, primIdx) */
};

const char* tfg_last_error(void);
int tfg_version(void);

/* ---- CWBVH construction (host) ------------------------------------------
 * Replaces tinybvh::BVH8_CWBVH::Build (tinybvh_wrapper.cpp:53).
 * verts: 9 floats per triangle (v0, v1, v2); primIdx = input position.
 * Runs on the host only; needs no GPU. */
typedef struct tfg_cwbvh_host tfg_cwbvh_host;
int tfg_cwbvh_build(const float* verts, uint32_t n_tris, uint32_t format, tfg_cwbvh_host** out);
int tfg_cwbvh_host_info(const tfg_cwbvh_host* h, uint32_t* num_nodes, uint32_t* num_tris, uint32_t* max_depth,
                        double* sah_cost);
const float* tfg_cwbvh_host_nodes(const tfg_cwbvh_host* h); /* 20 floats per node */
const float* tfg_cwbvh_host_tris(const tfg_cwbvh_host* h);  /* 16 floats per record */
void tfg_cwbvh_host_free(tfg_cwbvh_host* h);

/* Allocates device buffers and copies nodes and triangle records
 * (tinybvh_wrapper.cpp:59-81).  Free with tfg_cwbvh_gpu_free. */
int tfg_cwbvh_upload(const tfg_cwbvh_host* h, tfg_cwbvh_gpu* out);
int tfg_cwbvh_gpu_free(tfg_cwbvh_gpu* g);

/* ---- traversal ----------------------------------------------------------
 * Device function traverse_cwbvh (traverse_cwbvh.cuh:51-257) as batched
 * kernels.  Closest hit: tmin < t <= tmax, ties to the smaller primIdx;
 * miss -> prim = 0xFFFFFFFF, t = tmax.  Any hit: occluded[i] = 1 when some
 * triangle has tmin < t < tmax.  counts (nullable, device): 2 uint32 per
 * ray = {nodes decoded, triangles tested}.  The tree must be at most
 * TFG_MAX_TREE_DEPTH levels deep (trees from tfg_cwbvh_build are checked). */
#define TFG_MAX_TREE_DEPTH 25
int tfg_trace_closest(const tfg_cwbvh_gpu* bvh, uint32_t format, const tfg_ray* d_rays, tfg_hit* d_hits,
                      uint32_t* d_counts, uint32_t n, tfg_stream stream);
int tfg_trace_any(const tfg_cwbvh_gpu* bvh, uint32_t format, const tfg_ray* d_rays, uint32_t* d_occluded,
                  uint32_t* d_counts, uint32_t n, tfg_stream stream);


/* ---- wavefront path tracer ------------------------------------------------
 * Replaces kernelRender_tiny (RenderKernel.cu:28-47, render.h:416-466): one
 * path per pixel per frame as max_depth rounds of extend / shade / connect
 * over compacted device queues, optionally captured in a hipGraph.
 * Pixel seeds follow inicializarSemilla(x, y, frame) (rng.h:25-35). */
typedef struct tfg_pt tfg_pt;

enum {
    TFG_PT_INFER = 1,   /* usar_red_inferencia: capture cache queries */
    TFG_PT_TRAIN = 2,   /* entrenar_red: training paths and records  */
};

typedef struct tfg_pt_config {
    uint32_t width, height;
    uint32_t max_depth;      /* bounces, <= 32 (render.h:63 uses 20)        */
    uint32_t rr_depth;       /* Russian roulette when depth > rr_depth (render.h:210 uses 10) */
    uint32_t mode;           /* TFG_PT_INFER | TFG_PT_TRAIN; 0 = plain path tracing */
    float train_fraction;    /* training-pixel probability (render.h:443 uses 0.05) */
    uint32_t use_graph;      /* 1: capture the frame into a hipGraph and replay it */
} tfg_pt_config;

typedef struct tfg_pt_scene {
    tfg_cwbvh_gpu bvh;                  /* watertight triangle records */
    const tfg_tri_shading* d_shading;   /* per primIdx */
    const tfg_material* d_materials;
    const tfg_point_light* d_lights;
    uint32_t n_lights;
} tfg_pt_scene;

#define TFG_PT_NCOUNTERS 72
int tfg_pt_create(const tfg_pt_config* cfg, tfg_pt** out);
int tfg_pt_set_scene(tfg_pt* pt, const tfg_pt_scene* scene);
int tfg_pt_set_camera(tfg_pt* pt, const tfg_camera* cam);
/* Renders frame `frame` (1 sample per pixel): writes d_image[W*H] and, if
 * d_accum is not NULL, adds the frame to d_accum[3*W*H] (FP64). */
int tfg_pt_render(tfg_pt* pt, int32_t frame, tfg_color* d_image, double* d_accum, tfg_stream stream);
/* Copies the device counters of the last frame to host (synchronises):
 * [8 + d] extension rays at depth d, [40 + d] shadow rays at depth d,
 * [3] training records, [4] cache queries. */
int tfg_pt_counters(tfg_pt* pt, uint32_t* out);
/* Extension and shadow rays traced since creation / the last reset,
 * including the last frame (synchronises). */
int tfg_pt_ray_totals(tfg_pt* pt, uint64_t* out2);
int tfg_pt_reset_totals(tfg_pt* pt);
/* One frame launched stage by stage (no graph) with HIP events around every
 * launch and node/triangle visit counters in the traversal stages: for the
 * roofline of the path tracer's kernels.  Synchronises. */
typedef struct tfg_pt_profile {
    float ms_generate, ms_finalize, ms_total;
    float ms_extend[32], ms_shade[32], ms_connect[32];
    uint32_t ext_rays[32], shadow_rays[32];
    uint64_t ext_nodes, ext_tris, shadow_nodes, shadow_tris;
} tfg_pt_profile;
int tfg_pt_profile_frame(tfg_pt* pt, int32_t frame, tfg_color* d_image, tfg_pt_profile* out);
/* Times candidate launch shapes of the traversal stages on the current scene
 * and keeps the fastest (drops any captured graph). */
int tfg_pt_autotune(tfg_pt* pt, tfg_stream stream);
int tfg_pt_get_launch(const tfg_pt* pt, uint32_t* trace_block, uint32_t* trace_grid, uint32_t* shade_grid);
int tfg_pt_destroy(tfg_pt* pt);

#ifdef __cplusplus
}
#endif
#endif /* TFG_H */
