// NrcKernel.h - drop-in replacement of the reference's NrcKernel.h
// (NrcKernel.h:24-41) for callers built with hipcc; include after the
// caller's DatosMLP, SceneBounds and Color.  The (blocks, threads) arguments
// are ignored; hipStream_t replaces cudaStream_t.
#pragma once
#include <hip/hip_runtime.h>

#include "../tfg.h"

static_assert(sizeof(DatosMLP) == sizeof(tfg_datos_mlp), "DatosMLP layout (mlp_types.h:14-23)");
static_assert(sizeof(SceneBounds) == sizeof(tfg_scene_bounds), "SceneBounds layout (mlp_types.h:26-29)");

inline tfg_scene_bounds tfg_ref_bounds(const SceneBounds& b) {
    tfg_scene_bounds o;
    __builtin_memcpy(&o, &b, sizeof(o));
    return o;
}

inline void launchGenerarIndicesAleatorios(int, int, hipStream_t stream, int* indices, int max_range, int count,
                                           unsigned int seed) {
    (void)tfg_launch_generar_indices(indices, max_range, count, seed, stream);
}

inline void launchPrepararDatosEntrenamiento(int, int, hipStream_t stream, const DatosMLP* buffer,
                                             const int* indices_aleatorios, int batch_size, int n_in, int n_out,
                                             SceneBounds bounds, float* input_matrix, float* target_matrix) {
    (void)tfg_launch_preparar_entrenamiento(reinterpret_cast<const tfg_datos_mlp*>(buffer), indices_aleatorios,
                                            batch_size, n_in, n_out, tfg_ref_bounds(bounds), input_matrix,
                                            target_matrix, stream);
}

inline void launchPrepararDatosInferencia(int, int, hipStream_t stream, const DatosMLP* datos, uint32_t n_elements,
                                          uint32_t n_in, float* buffer_in, SceneBounds bounds) {
    (void)tfg_launch_preparar_inferencia(reinterpret_cast<const tfg_datos_mlp*>(datos), n_elements, n_in, buffer_in,
                                         tfg_ref_bounds(bounds), stream);
}

inline void launchGuardarSalidaInferencia(int, int, hipStream_t stream, float* network_output, Color* buffer_color,
                                          int n_elements) {
    (void)tfg_launch_guardar_salida(network_output, reinterpret_cast<tfg_color*>(buffer_color), n_elements, stream);
}
