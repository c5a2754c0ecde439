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

#ifdef __cplusplus
}
#endif
#endif /* TFG_H */
