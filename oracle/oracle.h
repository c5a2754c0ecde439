/*
 * oracle.h - CPU restatement of the reference path tracer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library in tfg_amd/ is tested against
 * this library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  Product code never links it.
 *
 * Each function names the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed in
 * the header of oracle.cpp.
 */
#ifndef TFG_ORACLE_H
#define TFG_ORACLE_H

#include <stdint.h>
#include "../include/tfg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rng.h:12-35 */
uint32_t orc_pcg32(uint32_t* state);
float orc_pcg32_float(uint32_t* state);
uint32_t orc_seed(int px, int py, int frame);
/* Writes n draws of the stream seeded by orc_seed(px, py, frame). */
void orc_stream(int px, int py, int frame, int n, uint32_t* out_u32, float* out_f32);

/* RenderKernel.cu:68-90 + camara.h:29-38 (FP64 Vector3d, FP32 plane sizes) */
void orc_camera_default(int width, int height, tfg_camara* out);
void orc_camera_make(const double pos[3], const double front[3], const double up[3], double fov_deg,
                     int width, int height, tfg_camara* out);
/* camara.h:42-59: primary ray for pixel (x, y) with jitter (ox, oy). */
void orc_camera_ray(const tfg_camara* cam, int x, int y, int w, int h, float ox, float oy,
                    double origin[3], double dir[3]);

/* traverse_cwbvh.cuh:51-257 decode + this library's triangle tests.
 * format 0 = watertight vertex records, 1 = affine (reference) records.
 * counts (nullable): per ray {nodes decoded, triangles tested}. */
void orc_cwbvh_trace(const float* nodes, const float* tris, uint32_t format, const tfg_ray* rays,
                     uint32_t n, tfg_hit* hits, uint32_t* counts, int nthreads);
void orc_cwbvh_occluded(const float* nodes, const float* tris, uint32_t format, const tfg_ray* rays,
                        uint32_t n, uint32_t* occluded, uint32_t* counts, int nthreads);
/* Brute force over all triangles with the same watertight test. verts: 9
 * floats per triangle, prim id = triangle index. */
void orc_brute_trace(const float* verts, uint32_t n_tris, const tfg_ray* rays, uint32_t n, tfg_hit* hits);
/* Single watertight test (exposed for the closed-form unit tests). */
int orc_tri_watertight(const float v[9], const float o[3], const float d[3], float tmin, float tmax,
                       float* t, float* u, float* v_out);

/* bvh.cpp:14-115 (median split BVH2, leaf <= 4) over Primitiva triangles,
 * traversed as escenario.h:187-237 with primitiva.h:164-183 Moller-Trumbore
 * (FP64 vectors, FP32 scalars).  Ray directions are renormalised in FP64
 * as Rayo does (rayo.h:22-23).  Returns hit t in FP32, prim = original
 * index.  Hits must satisfy t > tmin (escenario.h:349 uses EPSILON). */
void* orc_bvh2_build(const tfg_primitiva* prims, uint32_t n);
void orc_bvh2_free(void* bvh);
uint32_t orc_bvh2_num_nodes(const void* bvh);
void orc_bvh2_trace(const void* bvh, const tfg_ray* rays, uint32_t n, tfg_hit* hits, int nthreads);

/* CWBVH structural check: every child box decoded in FP64 contains every
 * triangle below it; every triangle record appears exactly once.
 * Returns 0 when valid, otherwise a nonzero code; *depth gets the max
 * node depth. */
int orc_cwbvh_validate(const float* nodes, uint32_t num_nodes, const float* tris, uint32_t num_tris,
                       uint32_t format, const float* verts, uint32_t* depth);


/* render.h:33-466 + escenario.h:123-184 restated on the CPU: geometry and
 * directions in FP64 (Vector3d), colours in FP32 (Color), intersections
 * by the CWBVH decode above on FP32-cast rays.  Same extensions as the HIP
 * integrator: GGX lobe for roughness > 0, shadow rays against the mesh,
 * configurable max depth / Russian-roulette depth.
 * cam: tfg_camera (16 floats); shading: tfg_tri_shading per primIdx;
 * accum (W*H*3 doubles) += sum over frames frame0 .. frame0+spp-1;
 * nthreads host threads.  Returns total extension and shadow rays traced
 * through *rays_out[2] (nullable). */
typedef struct orc_scene {
    const float* nodes;
    const float* tris;
    const tfg_tri_shading* shading;
    const tfg_material* materials;
    const tfg_point_light* lights;
    uint32_t n_lights;
    const tfg_camera* cam;
    const tfg_analytic* analytic;   /* may be NULL; hit id 0x80000000 | index */
    uint32_t n_analytic;
} orc_scene;
void orc_render(const orc_scene* sc, int width, int height, int frame0, int spp, int max_depth, int rr_depth,
                double* accum, uint64_t* rays_out, int nthreads);
/* Rows [y0, y1) only; accum holds (y1 - y0) * width pixels. */
void orc_render_rows(const orc_scene* sc, int width, int height, int y0, int y1, int frame0, int spp, int max_depth,
                     int rr_depth, double* accum, uint64_t* rays_out, int nthreads);
/* Cache-mode paths (render.h:33-262 with usar_red_inferencia / entrenar_red
 * from mode bits 1 / 2) of every pixel of one frame.  out: W*H x 64 floats:
 * pixel(3), flags(1: bit0 training, bit1 head, bit2 query, bit3 tail),
 * head or query vertex(16), tail vertex(16), L_suffix(3), T_suffix(3),
 * throughput x reflectance at the head / query (3), pad. */
void orc_render_nrc(const orc_scene* sc, int width, int height, int frame, int max_depth, int rr_depth, int mode,
                    float train_frac, float* out, int nthreads);
/* One path of pixel (x, y) at `frame`; out_rgb[3]. */
void orc_render_pixel(const orc_scene* sc, int width, int height, int x, int y, int frame, int max_depth,
                      int rr_depth, float* out_rgb);
// Conditioning probe: the paths of n pixels (xy: x, y pairs) of one frame,
// each ray displaced by delta before its closest-hit query (test use).
void orc_render_pixels_perturbed(const orc_scene* sc, int width, int height, const int32_t* xy, uint32_t n, int frame,
                                 int max_depth, int rr_depth, double delta, uint32_t pseed, float* out_rgb,
                                 int nthreads);

/* Shading functions of the integrator above, for the closed-form tests:
 * cosine-weighted direction about +z (render.h:383-390); GGX D, Smith G1
 * and the microfacet BRDF times Schlick Fresnel in the local frame (n =
 * +z); the visible-normal draw (Heitz 2018); analytic sphere / plane /
 * triangle distance (primitiva.h:137-187, -1 for a miss). */
void orc_cosine_local(double r1, double r2, double out[3]);
double orc_ggx_D(double alpha, double nh);
double orc_ggx_G1(double alpha, double nv);
void orc_ggx_eval(const float ks[3], double alpha, const double wo[3], const double wi[3], float out[3]);
void orc_ggx_vndf(const double wo[3], double alpha, double u1, double u2, double out[3]);
float orc_analytic_t(const tfg_analytic* q, const double o[3], const double d[3]);

#ifdef __cplusplus
}
#endif
#endif
