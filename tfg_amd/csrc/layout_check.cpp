// layout_check.cpp - compile-time proof that the reference-layout structs
// of tfg_types.h have the sizes and field offsets of the reference's
// host structs (sizes from a host compile of the reference headers:
// Primitiva 200, LuzPuntual 40, Camara 112, DatosMLP 112,
// RegistroEntrenamiento 232, SceneBounds 48, ImagenGPU 16, TinyBVHD_GPU 24).
#include <cstddef>

#include "../../include/tfg.h"
#include "../../include/tfg_types.h"

static_assert(sizeof(tfg_vec3d) == 24, "Vector3d");
static_assert(sizeof(tfg_color) == 12, "Color");
static_assert(sizeof(tfg_primitiva) == 200, "Primitiva");
static_assert(offsetof(tfg_primitiva, emision) == 4, "Primitiva::emision");
static_assert(offsetof(tfg_primitiva, indice_refraccion) == 52, "Primitiva::indice_refraccion");
static_assert(offsetof(tfg_primitiva, g) == 56, "Primitiva union");
static_assert(sizeof(tfg_luz_puntual) == 40, "LuzPuntual");
static_assert(sizeof(tfg_camara) == 112, "Camara");
static_assert(offsetof(tfg_camara, ancho_plano) == 96, "Camara::ancho_plano");
static_assert(sizeof(tfg_datos_mlp) == 112, "DatosMLP");
static_assert(offsetof(tfg_datos_mlp, color) == 96, "DatosMLP::color");
static_assert(sizeof(tfg_datos_geometricos) == 96, "DatosGeometricos");
static_assert(sizeof(tfg_registro_entrenamiento) == 232, "RegistroEntrenamiento");
static_assert(offsetof(tfg_registro_entrenamiento, luz_acumulada_sufijo) == 192, "luz_acumulada_sufijo");
static_assert(offsetof(tfg_registro_entrenamiento, factor_normalizacion) == 216, "factor_normalizacion");
static_assert(offsetof(tfg_registro_entrenamiento, valido) == 228, "valido");
static_assert(sizeof(tfg_scene_bounds) == 48, "SceneBounds");
static_assert(sizeof(tfg_imagen_gpu) == 16, "ImagenGPU");
static_assert(sizeof(tfg_cwbvh_gpu) == 24, "TinyBVHD_GPU");
static_assert(sizeof(tfg_ray) == 32 && sizeof(tfg_hit) == 16, "ray/hit");
static_assert(sizeof(tfg_material) == 64 && sizeof(tfg_tri_shading) == 64, "material/shading");
static_assert(sizeof(tfg_point_light) == 32 && sizeof(tfg_camera) == 64, "light/camera");
static_assert(sizeof(tfg_analytic) == 48, "analytic");
static_assert(sizeof(tfg_pt_view) == 72 && offsetof(tfg_pt_view, camera) == 8, "tfg_pt_view");
