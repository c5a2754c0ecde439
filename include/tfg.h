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
/* Builder settings (tfg_cwbvh_build uses the defaults): SAH costs of a
 * wide-node visit and of a triangle test for the 8-wide collapse (1, 1),
 * triangles per leaf slot (1..3, default 2) and SAH bins per axis (4..64,
 * default 32).  Any setting gives a valid tree; they trade node visits
 * against triangle tests. */
typedef struct tfg_cwbvh_build_options {
    float node_cost, tri_cost;
    uint32_t max_leaf, bins;
} tfg_cwbvh_build_options;
void tfg_cwbvh_default_build_options(tfg_cwbvh_build_options* o);
int tfg_cwbvh_build_ex(const float* verts, uint32_t n_tris, uint32_t format, const tfg_cwbvh_build_options* opts,
                       tfg_cwbvh_host** out);
int tfg_cwbvh_host_info(const tfg_cwbvh_host* h, uint32_t* num_nodes, uint32_t* num_tris, uint32_t* max_depth,
                        double* sah_cost);
const float* tfg_cwbvh_host_nodes(const tfg_cwbvh_host* h); /* 20 floats per node */
const float* tfg_cwbvh_host_tris(const tfg_cwbvh_host* h);  /* 16 floats per record */
void tfg_cwbvh_host_free(tfg_cwbvh_host* h);

/* Allocates device buffers and copies nodes and triangle records
 * (tinybvh_wrapper.cpp:59-81).  Free with tfg_cwbvh_gpu_free. */
int tfg_cwbvh_upload(const tfg_cwbvh_host* h, tfg_cwbvh_gpu* out);
int tfg_cwbvh_gpu_free(tfg_cwbvh_gpu* g);
/* Copies an uploaded tree back to host memory (20 floats per node, 16 per
 * triangle record); synchronises the device. */
int tfg_cwbvh_gpu_download(const tfg_cwbvh_gpu* g, float* nodes, float* tris);
/* Refit of an uploaded tree to moved vertices (animated geometry whose
 * triangles keep their identity, e.g. rigidly moved instances): the tree is
 * kept, the triangle records are rewritten from the new vertices and every
 * node's quantisation frame and child bounds are recomputed bottom-up on
 * the GPU, one launch per level.  Same box and quantisation arithmetic as
 * the builder: with unchanged vertices the uploaded tree is reproduced bit
 * for bit.  The refit object is made once from the host tree (its level
 * ranges plus a device box per node); d_verts holds 9 finite floats per
 * triangle in the build's input order.  Asynchronous on `stream`.  No
 * reference counterpart (the reference rebuilds, tinybvh_wrapper.cpp:53). */
typedef struct tfg_cwbvh_refitter tfg_cwbvh_refitter;
int tfg_cwbvh_refit_create(const tfg_cwbvh_host* h, tfg_cwbvh_refitter** out);
int tfg_cwbvh_refit(tfg_cwbvh_refitter* r, tfg_cwbvh_gpu* bvh, const float* d_verts, uint32_t n_tris,
                    tfg_stream stream);
int tfg_cwbvh_refit_destroy(tfg_cwbvh_refitter* r);

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

/* Scene of a tracer.  bvh's triangle records are watertight
 * (tri_format TFG_TRI_WATERTIGHT) or tinybvh's affine unit-triangle records
 * (TFG_TRI_AFFINE, the layout the reference uploads, tinybvh_wrapper.cpp:3,
 * traverse_cwbvh.cuh:187-209, primIdx in T3.w).  For an affine tree d_verts
 * holds the triangles' vertices, 9 floats per primIdx (the build's input
 * order): the tracer keeps a watertight copy of the records, made from them
 * on the device at set_scene (call set_scene again after changing them).  A
 * tree not built by this library is walked once on the host to find its
 * depth (at most TFG_MAX_TREE_DEPTH levels). */
typedef struct tfg_pt_scene {
    tfg_cwbvh_gpu bvh;
    const tfg_tri_shading* d_shading;   /* per primIdx */
    const tfg_material* d_materials;
    const tfg_point_light* d_lights;
    uint32_t n_lights;                  /* <= 64 */
    const tfg_analytic* d_analytic;     /* may be NULL */
    uint32_t n_analytic;                /* <= 64 */
    uint32_t tri_format;                /* TFG_TRI_WATERTIGHT (0) or TFG_TRI_AFFINE */
    const float* d_verts;               /* TFG_TRI_AFFINE: 9 floats per primIdx; else unused */
    uint32_t n_verts;                   /* triangles in d_verts (primIdx < n_verts) */
} tfg_pt_scene;

#define TFG_PT_NCOUNTERS 72
/* Creates a tracer (its device state and streams).  The first call on a
 * device in a process first creates and destroys as many streams as a
 * tracer uses: a tracer whose streams are the process's first ones ran its
 * overlapped schedules 15-35% slower (the runtime's hardware-queue pool). */
int tfg_pt_create(const tfg_pt_config* cfg, tfg_pt** out);
int tfg_pt_set_scene(tfg_pt* pt, const tfg_pt_scene* scene);
int tfg_pt_set_camera(tfg_pt* pt, const tfg_camera* cam);
/* Several cameras traced in one frame (one chain of launches): the canvas of
 * the config (width x height pixels) holds the views' images one after the
 * other, row-major, view v at pixels base_v .. base_v + W_v*H_v - 1 with
 * base_0 = 0 and base_{v+1} = base_v + W_v*H_v; the views' pixels must fill
 * the canvas exactly.  Each view's pixels get the seeds and camera rays a
 * tracer of that view's size with tfg_pt_set_camera would give them, so in a
 * plain frame each view's image equals that tracer's, bit for bit.  Every
 * path of the canvas shares the frame's stages, and in cache mode one
 * radiance cache.  1..64 views of 1..65535 pixels each way.  Replaces the
 * camera; tfg_pt_set_camera returns to one camera.  No reference
 * counterpart (the reference renders one camera per process, main.cpp). */
typedef struct tfg_pt_view {
    uint32_t width, height;
    tfg_camera camera;   /* the view's camera at its own resolution */
} tfg_pt_view;
int tfg_pt_set_views(tfg_pt* pt, const tfg_pt_view* views, uint32_t n_views);
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
/* For the roofline of the path tracer's kernels: the frame is run twice
 * with direct launches, first with node/triangle visit counters in the
 * traversal stages, then with the plain kernels and HIP events around every
 * launch (ms_* = sum over tiles of that stage's launch durations; ms_total =
 * the whole frame).  Synchronises. */
typedef struct tfg_pt_profile {
    float ms_generate, ms_finalize, ms_total;
    float ms_extend[32], ms_shade[32], ms_connect[32];
    uint32_t ext_rays[32], shadow_rays[32];
    uint64_t ext_nodes, ext_tris, shadow_nodes, shadow_tris;
    /* start of each of tile 0's launches, ms after the frame's start */
    float at_extend[32], at_shade[32], at_connect[32], at_finalize;
    /* the tail kernel (schedule tail_from): duration summed over the tiles,
     * start of tile 0's launch; 0 without one */
    float ms_tail, at_tail;
} tfg_pt_profile;
int tfg_pt_profile_frame(tfg_pt* pt, int32_t frame, tfg_color* d_image, tfg_pt_profile* out);
/* Times candidate schedules (tfg_pt_schedule) on the current scene and keeps
 * the fastest: first the traversal launch shape together with top_nodes,
 * then tiles, side streams and overlap, then the XCD super-chunk size,
 * octet_below, connect_per_cu, connect_batch and tail_from, then the two
 * queue orders and lane refill (with its workgroups per CU), then a second round over all of them but the
 * launch shape (top_nodes included), all timed on plain frames with direct
 * launches (the tuner captures no graph).
 * Drops any captured graph. */
int tfg_pt_autotune(tfg_pt* pt, tfg_stream stream);
/* The traversal stages' workgroup size and grid, and the grid of the
 * per-path kernels (generate, shade, finalize: 256-thread workgroups, one
 * path per thread within 8-64 workgroups per CU) for tile 0. */
int tfg_pt_get_launch(const tfg_pt* pt, uint32_t* trace_block, uint32_t* trace_grid, uint32_t* shade_grid);
/* How a frame is launched.  None of these change the results: every
 * schedule gives the same image bit for bit.  No reference counterpart (the
 * reference launches one megakernel per frame, RenderKernel.cu:185-200).
 *   trace_block, grid_mult: traversal stages' block size (64, 128, 256) and
 *       persistent grid (resident blocks per CU x grid_mult x CUs);
 *   tiles (1..4): the pixels split into that many contiguous ranges, each
 *       traced by its own chain of launches on its own streams;
 *   side_streams (1..4), overlap: with overlap on, the shadow-ray stage of
 *       depth d runs on side stream d % side_streams of its tile, after the
 *       shading stage of depth d and concurrently with the following depths
 *       (NEE contributions are kept per depth and summed in depth order);
 *   xcd_group: rays per XCD super-chunk in the traversal stages (0 = plain
 *       grid stride; else a multiple of 256): consecutive queue entries are
 *       grouped so that each group is traced by the workgroups of one XCD
 *       (one L2), groups dealt to the 8 XCDs in turn;
 *   octet_below: traversal launches of at most this many rays trace each
 *       ray with 8 lanes (one per child slot, up to 8 triangles at once),
 *       which shortens each ray's latency when too few rays fill the GPU
 *       (shadow rays: only with at most 8 lights);
 *   connect_per_cu (0..64): the shadow stage's persistent grid, workgroups
 *       per CU (0: the closest-hit stages' grid); a small grid leaves CU
 *       slots to the closest-hit chain, the frame's critical path;
 *   connect_batch (0..32): 0 gives every depth's shadow stage its own
 *       launch; k > 0 traces the shadow rays of depths k..max_depth-1 in one
 *       launch after the last shading stage (late depths have few rays, and
 *       one launch of them replaces a chain of small launches);
 *   top_nodes (0..73): leading tree nodes (breadth-first) each traversal
 *       workgroup copies to LDS before tracing; more saves global fetches of
 *       the upper levels but costs every workgroup its staging, which
 *       matters for grids of many short-lived workgroups (default 9);
 *   tail_from (0..32): 0 traces every depth with the per-depth stages;
 *       k > 0 traces depths k..max_depth-1 with one launch of a tail kernel
 *       in which each remaining path is carried to its end by 8 lanes
 *       (closest hit, shading, shadow rays one lane per light): late depths
 *       hold few paths (in cache frames mostly training paths), and the
 *       per-depth chain of 3 launches per depth is bound by latency there;
 *   defer_training (0/1, direct launches only): a training frame's
 *       optimiser steps run at the start of the next tfg_pt_render, on a
 *       stream of their own beside that frame's tracing, on a copy of the
 *       records they draw (made at the end of their frame); the next cache
 *       inference waits for them, so every image is the one the steps at the
 *       end of their own frame give.  The last frame's steps wait for the next
 *       render or tfg_pt_flush_training.  Only for a cache attached to
 *       this tracer alone: tracers sharing a cache would train it from
 *       several streams at once.  The tuner leaves it unchanged;
 *   use_graph: replay captured hipGraphs (one per tile, plus one for the
 *       cache work) instead of launching the kernels directly; read-only,
 *       fixed by tfg_pt_config.use_graph (set_schedule must pass the same
 *       value). */
typedef struct tfg_pt_schedule {
    uint32_t trace_block, grid_mult, tiles, side_streams;
    int32_t overlap, use_graph;
    uint32_t xcd_group;
    uint32_t octet_below;
    uint32_t connect_per_cu;
    uint32_t connect_batch;
    uint32_t top_nodes;
    uint32_t tail_from;
    uint32_t defer_training;
    uint32_t shadow_sort_from;   /* 0: off; k (default 1): the shadow entries of depths >= k are traced in the order of a
                                    12-bit Morton code of their point (a counting sort before each shadow stage),
                                    at depths whose shadow queue holds at least 1/16 of the frame's (tile's) paths */
    uint32_t ray_sort_from;      /* 0: off; k: the extension rays of depths >= k are traced in the order of their
                                    direction's cube-map cell (6 x 4 x 4) and a coarse cell of their origin, at
                                    depths whose queue holds at least 1/16 of the paths */
    uint32_t refill_from;        /* 0: off; k (default 1): the closest-hit stage of depths >= k keeps its lanes busy
                                    (lane refill): a persistent wave takes new rays from the queue whenever at least
                                    refill_idle of its lanes have finished theirs, instead of waiting for its slowest
                                    ray; at depths with at least two rays per resident lane (fewer: lockstep, and at
                                    most octet_below rays: 8 lanes per ray) */
    uint32_t refill_idle;        /* 1..64 (default 32) */
    uint32_t refill_per_cu;      /* 0 (default): the lane-refill closest-hit kernel keeps as many workgroups resident
                                    per CU as fit; k (1..64): at most k, leaving CU slots and LDS to the stages that
                                    run beside it (shadow-entry sort, shadow stage) */
} tfg_pt_schedule;
int tfg_pt_get_schedule(const tfg_pt* pt, tfg_pt_schedule* out);
int tfg_pt_set_schedule(tfg_pt* pt, const tfg_pt_schedule* schedule);
int tfg_pt_destroy(tfg_pt* pt);


/* ---- frame output ----------------------------------------------------------
 * Replaces the host loop of main.cpp:1070-1071 / 1109-1115 and the byte
 * conversion of guardarPNG (imagen_utils.h:58-62): per pixel, the mean of
 * an FP64 accumulation buffer (3 doubles per pixel, as tfg_pt_render
 * accumulates) over `samples` frames, optionally the ACES filmic curve of
 * Imagen::filmic (imagen.h:194-215), then round(clamp(v, 0, 1) * 255) into
 * d_rgb8 (3 bytes per pixel, nullable) and the FP32 mean into d_mean
 * (nullable).  tfg_amd/image.py writes the bytes as an 8-bit RGB PNG. */
int tfg_image_resolve(const double* d_accum, uint32_t n_pixels, uint32_t samples, int tonemap, uint8_t* d_rgb8,
                      float* d_mean, tfg_stream stream);

/* ---- radiance cache (NRC) -------------------------------------------------
 * Replaces ColorMLP (mlp.h:19-70, mlp.cpp:16-170) and the NRC data kernels
 * (NrcKernel.cu, RenderKernel.cu:92-166).  64 encoded inputs -> 5 x 64 ReLU
 * -> 3 outputs, FP32 master weights, BF16 (or FP16) matrix-core operands
 * with FP32 accumulation, RelativeL2Luminance loss with static loss
 * scaling, Adam + EMA.  Attached to a tfg_pt, its inference and training
 * run inside the path tracer's frame (and its hipGraph). */
typedef struct tfg_nrc tfg_nrc;

enum { TFG_NRC_BF16 = 0, TFG_NRC_FP16 = 1 };
#define TFG_NRC_PARAMS 21504   /* 5 x 64 x 64 + 16 x 64 (output rows 3..15 are padding) */

typedef struct tfg_nrc_config {
    uint32_t batch;            /* samples per optimiser step, multiple of 64 (main.cpp:939: 16384) */
    uint32_t steps_per_frame;  /* optimiser steps per frame */
    uint32_t min_records;      /* no training with <= this many records (main.cpp:1043: 1024) */
    float lr, beta1, beta2, eps, ema_decay, loss_scale;
    float bounds_min[3], bounds_max[3];   /* SceneBounds (main.cpp:870-872) */
    uint32_t seed;             /* weight initialisation */
    uint32_t precision;        /* TFG_NRC_BF16 / TFG_NRC_FP16 */
} tfg_nrc_config;

typedef struct tfg_nrc_stats {
    uint32_t steps;              /* optimiser steps applied */
    uint32_t skipped_nonfinite;  /* steps dropped for a non-finite gradient */
    uint32_t skipped_no_data;    /* steps skipped for too few records */
    uint32_t seed_counter;       /* batch-index seed (mlp.cpp:102 step_counter) */
    float last_loss;             /* loss of the last applied step */
} tfg_nrc_stats;

int tfg_nrc_default_config(tfg_nrc_config* cfg);
int tfg_nrc_create(const tfg_nrc_config* cfg, tfg_nrc** out);
int tfg_nrc_destroy(tfg_nrc* nrc);
/* Checkpoint / restore (replaces mlp.cpp:173-276): TFG_NRC_PARAMS floats
 * each, row-major per layer [out][in]; ema may be NULL (= master).
 * set_weights resets the optimiser state. */
int tfg_nrc_set_weights(tfg_nrc* nrc, const float* master, const float* ema);
int tfg_nrc_get_weights(tfg_nrc* nrc, float* master, float* ema);
int tfg_nrc_get_stats(tfg_nrc* nrc, tfg_nrc_stats* out);
/* Evaluates the network on n vertex records (4 float4 each, the layout of
 * the cache queries): raw outputs d_out[3n]; d_dump (nullable) receives the
 * pre-activations of all 6 layers (n x 6 x 64 floats). */
int tfg_nrc_eval(tfg_nrc* nrc, const float* d_verts, uint32_t n, float* d_out, float* d_dump, int use_train_weights,
                 tfg_stream stream);
/* The frame's cache inference on its own: predictions for cache queries (5
 * float4 each: vertex record, then throughput and pixel index) whose count
 * is read from device memory at *d_count, composited into d_image
 * (image[pixel] += throughput * prediction, RenderKernel.cu:49-66).  Used
 * to time the inference kernel; tfg_pt_render runs the same kernel. */
int tfg_nrc_infer_queries(tfg_nrc* nrc, const float* d_queries, const uint32_t* d_count, tfg_color* d_image,
                          tfg_stream stream);
/* `steps` optimiser steps on training records (11 float4 each) whose count
 * is read from device memory at *d_n_records.  Every step forms its targets
 * (L_suffix + T_suffix * cache(tail)) / normaliser, RenderKernel.cu:92-166,
 * with the inference (EMA) weights as they were when the call was enqueued,
 * as the reference computes a frame's targets once before its training step
 * (main.cpp:1047-1056); the call stands for one frame's steps. */
int tfg_nrc_train(tfg_nrc* nrc, const float* d_records, const uint32_t* d_n_records, uint32_t steps,
                  tfg_stream stream);
/* Gradient (unscaled, FP32) of the last optimiser step, to host. */
int tfg_nrc_last_gradient(tfg_nrc* nrc, float* host_grad);
/* Copies the last frame's training records (11 x 4 floats each; the path
 * id is the bit pattern of float 39) and cache queries (5 x 4 floats; path
 * id in float 19) to host memory, at most max_* of each; synchronises. */
int tfg_pt_read_cache_data(tfg_pt* pt, float* records, uint32_t max_records, float* queries, uint32_t max_queries,
                           uint32_t* n_records, uint32_t* n_queries);
/* ColorMLP::inference (mlp.cpp:127-170) on DatosMLP records: sanitised
 * predictions (NrcKernel.cu:99-126) into d_out[n]; synchronous. */
int tfg_nrc_infer_datos(tfg_nrc* nrc, const tfg_datos_mlp* d_in, uint32_t n, tfg_color* d_out, tfg_stream stream);
/* tfg_nrc_infer_datos without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The converted inputs and raw outputs live in one static device array of
 * the library (4.75 MiB on every GPU it runs on, held until the process
 * ends), so at most TFG_NRC_NOALLOC_MAX_QUERIES queries per call (more:
 * TFG_E_LIMIT).  Same kernels and results as tfg_nrc_infer_datos, but
 * asynchronous on `stream`.  Every call on a GPU shares that GPU's array:
 * calls ordered on one stream are safe, calls that overlap on one GPU (other
 * streams, other threads) race. */
#define TFG_NRC_NOALLOC_MAX_QUERIES 65536
int tfg_nrc_infer_datos_noalloc(tfg_nrc* nrc, const tfg_datos_mlp* d_in, uint32_t n, tfg_color* d_out,
                                tfg_stream stream);
/* ColorMLP::train_step (mlp.cpp:95-125): one optimiser step on the
 * configured batch drawn (NrcKernel.cu:10-16) from d_in[0..n) with target =
 * DatosMLP.color; no-op with *loss = 0 when n < 100; synchronous. */
int tfg_nrc_train_datos(tfg_nrc* nrc, const tfg_datos_mlp* d_in, uint32_t n, float* loss, tfg_stream stream);
/* Attaches the cache to a path tracer created with a nonzero mode (pending
 * deferred steps of the previous cache are enqueued first). */
int tfg_pt_set_nrc(tfg_pt* pt, tfg_nrc* nrc);
/* With schedule defer_training: enqueues, ordered after the work on
 * `stream`, the optimiser steps of the last training frame, which otherwise
 * run at the start of the next tfg_pt_render; no-op when none is pending.
 * Call it before reading or saving the cache's weights or statistics after
 * the last frame.  tfg_pt_destroy drops steps still pending. */
int tfg_pt_flush_training(tfg_pt* pt, tfg_stream stream);
/* Switches the frame mode (0 = plain, TFG_PT_INFER | TFG_PT_TRAIN); a
 * tracer created in plain mode cannot switch to a cache mode. */
int tfg_pt_set_mode(tfg_pt* pt, uint32_t mode);
/* Device buffers of the last frame's cache data: records (11 float4 each),
 * queries (5 float4 each) and the counter array (records at [3], queries at [4]). */
int tfg_pt_nrc_buffers(tfg_pt* pt, float** d_records, float** d_queries, uint32_t** d_counters);


/* ---- reference launchers on the reference's struct layouts ------------------
 * Same argument meaning as the host wrappers of RenderKernel.h:63-78 and
 * NrcKernel.h:24-41; the CUDA grid/block arguments are dropped (the library
 * picks its own launch), cudaStream_t becomes tfg_stream. */
int tfg_launch_inicializar_camara(tfg_camara* d_cam, int width, int height, tfg_stream stream);   /* RenderKernel.h:67 */
int tfg_launch_composite(tfg_color* d_img, const tfg_color* d_pred, const tfg_color* d_thr, int width, int height,
                         tfg_stream stream);                                                      /* RenderKernel.h:63-65 */
int tfg_launch_calcular_targets(const tfg_registro_entrenamiento* d_recs, const tfg_color* d_pred_tail,
                                tfg_datos_mlp* d_out, int n, tfg_stream stream);                  /* RenderKernel.h:69-73 */
int tfg_launch_preparar_inferencia_tail(const tfg_registro_entrenamiento* d_recs, tfg_datos_mlp* d_out, int n,
                                        tfg_stream stream);                                       /* RenderKernel.h:75-78 */
int tfg_launch_generar_indices(int* d_idx, int max_range, int count, uint32_t seed, tfg_stream stream);   /* NrcKernel.h:24-25 */
int tfg_launch_preparar_entrenamiento(const tfg_datos_mlp* d_buf, const int* d_idx, int batch, int n_in, int n_out,
                                      tfg_scene_bounds bounds, float* d_in, float* d_tgt, tfg_stream stream);   /* NrcKernel.h:27-35 */
int tfg_launch_preparar_inferencia(const tfg_datos_mlp* d_buf, uint32_t n, uint32_t n_in, float* d_in,
                                   tfg_scene_bounds bounds, tfg_stream stream);                    /* NrcKernel.h:37-38 */
int tfg_launch_guardar_salida(const float* d_out, tfg_color* d_col, int n, tfg_stream stream);     /* NrcKernel.h:40-41 */


/* ---- launchKernelRenderTiny drop-in ----------------------------------------
 * RenderKernel.h:55-61 with the reference's argument list (grid/block
 * dropped, bool flags as int, plus a stream): renders one frame of the
 * scene given by the reference's buffers and writes the image, the cache
 * queries (buffer_inference / buffer_throughput, throughput 0 where no
 * query, render.h:452-458) and the training records
 * (buffer_training[atomicAdd(counter_train, 1) % max_cap_train],
 * render.h:461-464) in the reference layouts.  The CWBVH may come from this
 * library (tfg_cwbvh_upload / the TinyBVH shim, either triangle format) or
 * be any tree in tinybvh's layout (80 B nodes, 64 B affine triangle records
 * with primIdx in T3.w, tinybvh_wrapper.cpp:53-75): a tree this library did
 * not upload is taken as tinybvh's, and its triangles are traced with the
 * watertight test on the vertices of the per-primIdx Primitiva copy
 * (primitivas_bvh, RenderKernel.h:55-61).  Converted scene data are
 * cached per scene buffer; call tfg_dropin_invalidate() after rewriting a
 * scene buffer in place.  At most 64 analytic primitives and 64 lights. */
int tfg_launch_render_tiny(const tfg_camara* d_camara, const tfg_primitiva* d_prims, int n_prims,
                           const tfg_luz_puntual* d_luces, int n_luces, const tfg_primitiva* d_prims_malla,
                           int n_prims_malla, int width, int height, int spp, int frame, tfg_cwbvh_gpu bvh,
                           const tfg_primitiva* d_prims_bvh, int n_prims_bvh, tfg_imagen_gpu imagen,
                           unsigned int* d_dev_counter, tfg_registro_entrenamiento* d_recs,
                           unsigned int* d_counter_train, int max_cap_train, tfg_datos_mlp* d_inference,
                           tfg_color* d_throughput, int usar_red_inferencia, int entrenar_red, tfg_stream stream);
/* Path length and Russian-roulette depth of the drop-in frames (defaults
 * 20 and 10, render.h:63,210). */
int tfg_dropin_configure(uint32_t max_depth, uint32_t rr_depth);
int tfg_dropin_invalidate(void);

#ifdef __cplusplus
}
#endif
#endif /* TFG_H */
