/*
 * tfg_types.h - plain-old-data types shared by the host, the HIP kernels and
 * the C ABI.  The first group keeps the byte layout of the reference's host
 * structs so that buffers filled by a reference caller can be handed over
 * unchanged; the second group is this library's own compact device layout.
 *
 * Reference layouts (sizes and offsets fixed by static_asserts in
 * tfg_amd/csrc/layout_check.cpp):
 *   Vector3d              vector3d.h:14-19        24 B (double x,y,z)
 *   Color                 color.h:13-15           12 B (float r,g,b)
 *   Primitiva             primitiva.h:17-39      200 B (tagged union)
 *   LuzPuntual            luzPuntual.h:13-16      40 B
 *   Camara                camara.h:15-22         112 B
 *   DatosMLP              mlp_types.h:14-23      112 B
 *   DatosGeometricos      mlp_types.h:32-38       96 B
 *   RegistroEntrenamiento mlp_types.h:41-53      232 B
 *   SceneBounds           mlp_types.h:26-29       48 B
 *   ImagenGPU             imagen_gpu.h:14-17      16 B (passed by value)
 *   TinyBVHD_GPU          tinybvh_wrapper.h:15-21 24 B (passed by value)
 */
#ifndef TFG_TYPES_H
#define TFG_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-layout types ------------------------------------------- */

typedef struct tfg_vec3d { double x, y, z; } tfg_vec3d;
typedef struct tfg_color { float r, g, b; } tfg_color;

/* TipoPrimitiva (primitiva.h:15) */
enum { TFG_ESFERA = 0, TFG_PLANO = 1, TFG_TRIANGULO = 2 };

typedef struct tfg_primitiva {
    int32_t tipo;
    tfg_color emision;
    tfg_color difuso;
    tfg_color especular;
    tfg_color transmision;
    float indice_refraccion;
    union {
        struct { tfg_vec3d centro; float radio; } esfera;
        struct { tfg_vec3d normal; float distancia; } plano;
        struct { tfg_vec3d v0, v1, v2, v0_normal, v1_normal, v2_normal; } triangulo;
    } g;
} tfg_primitiva;

typedef struct tfg_luz_puntual { tfg_vec3d posicion; tfg_color intensidad; } tfg_luz_puntual;

typedef struct tfg_camara {
    tfg_vec3d posicion, frente, arriba, izquierda;
    float ancho_plano, alto_plano, distancia_focal;
} tfg_camara;

typedef struct tfg_datos_mlp {
    tfg_vec3d posicion, direccion, normal;
    tfg_color difuso, especular, color;
} tfg_datos_mlp;

typedef struct tfg_datos_geometricos {
    tfg_vec3d posicion, direccion, normal;
    tfg_color difuso, especular;
} tfg_datos_geometricos;

typedef struct tfg_registro_entrenamiento {
    tfg_datos_geometricos head;
    tfg_datos_geometricos tail;
    tfg_color luz_acumulada_sufijo;
    tfg_color throughput_sufijo;
    tfg_color factor_normalizacion;
    uint8_t valido;           /* C++ bool: one byte */
} tfg_registro_entrenamiento;

typedef struct tfg_scene_bounds { tfg_vec3d min, max; } tfg_scene_bounds;

typedef struct tfg_imagen_gpu { int32_t anchura, altura; tfg_color* datos; } tfg_imagen_gpu;

/* Device CWBVH descriptor, same fields and order as TinyBVHD_GPU.
 * d_nodos: 5 x float4 (80 B) per node, d_tris: 4 x float4 (64 B) per
 * triangle record. */
typedef struct tfg_cwbvh_gpu {
    void* d_nodos;
    void* d_tris;
    uint32_t num_nodos;
    uint32_t num_tris;
} tfg_cwbvh_gpu;

/* ---- this library's own compact layouts -------------------------------- */

/* One ray for the trace entry points: 32 B, two float4 loads. */
typedef struct tfg_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} tfg_ray;

/* Closest-hit result: 16 B.  prim = 0xFFFFFFFF on a miss (t = ray tmax). */
typedef struct tfg_hit {
    float t, u, v;
    uint32_t prim;
} tfg_hit;

/* Material table entry (64 B).  kd/ks/kt follow Primitiva's difuso /
 * especular / transmision; roughness > 0 turns the specular lobe into GGX,
 * roughness == 0 keeps the reference's ideal mirror. */
typedef struct tfg_material {
    float kd[3]; float roughness;
    float ks[3]; float ior;
    float kt[3]; float pad0;
    float le[3]; float pad1;
} tfg_material;

/* Per-triangle shading record (64 B), indexed by primIdx: vertex normals
 * (used when flags bit 0 is set, as escenario.h:126-134 does for triangles
 * whose vertex normals differ), material index and the unnormalised
 * geometric normal (v1 - v0) x (v2 - v0). */
typedef struct tfg_tri_shading {
    float n0[3]; uint32_t material;
    float n1[3]; uint32_t flags;
    float n2[3]; uint32_t pad0;
    float ng[3]; uint32_t pad1;
} tfg_tri_shading;

/* Analytic primitive (48 B) tested linearly next to the CWBVH, like the
 * reference's Primitiva list (escenario.h:323-334, primitiva.h:43-200).
 * Its hit id is 0x80000000 | index. */
enum { TFG_AN_SPHERE = 0, TFG_AN_PLANE = 1, TFG_AN_TRIANGLE = 2 };
typedef struct tfg_analytic {
    float a[3]; uint32_t type;       /* sphere: centre; plane: normal; triangle: v0 */
    float b[3]; float s;             /* sphere: s = radius; plane: s = distance; triangle: b = v1 */
    float c[3]; uint32_t material;   /* triangle: c = v2 */
} tfg_analytic;

typedef struct tfg_point_light {
    float pos[3]; float pad0;
    float intensity[3]; float pad1;
} tfg_point_light;

/* Pinhole camera in FP32, same parameterisation as Camara (camara.h:42-59). */
typedef struct tfg_camera {
    float pos[3];   float focal;
    float front[3]; float plane_w;
    float up[3];    float plane_h;
    float left[3];  float pad;
} tfg_camera;

#ifdef __cplusplus
}
#endif
#endif /* TFG_TYPES_H */
