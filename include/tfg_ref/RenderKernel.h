// RenderKernel.h - drop-in replacement of the reference's RenderKernel.h
// (RenderKernel.h:47-78) for callers built with hipcc.  Include it after the
// caller's own definitions of Camara, Primitiva, LuzPuntual, ImagenGPU,
// TinyBVHD_GPU, RegistroEntrenamiento, DatosMLP and Color (camara.h,
// primitiva.h, ... ported to HIP); every launcher forwards to libtfg_hip.so
// through the C ABI of tfg.h.  The grid/block arguments are accepted and
// ignored: the library picks its own launch shapes.  Errors are reported the
// way the reference reports them, through the runtime's error state: the
// launchers return void and store the message for tfg_last_error().
#pragma once
#include <hip/hip_runtime.h>

#include "../tfg.h"

static_assert(sizeof(Camara) == sizeof(tfg_camara), "Camara layout (camara.h:15-22)");
static_assert(sizeof(Primitiva) == sizeof(tfg_primitiva), "Primitiva layout (primitiva.h:17-39)");
static_assert(sizeof(LuzPuntual) == sizeof(tfg_luz_puntual), "LuzPuntual layout");
static_assert(sizeof(ImagenGPU) == sizeof(tfg_imagen_gpu), "ImagenGPU layout");
static_assert(sizeof(TinyBVHD_GPU) == sizeof(tfg_cwbvh_gpu), "TinyBVHD_GPU layout");
static_assert(sizeof(RegistroEntrenamiento) == sizeof(tfg_registro_entrenamiento), "RegistroEntrenamiento layout");
static_assert(sizeof(DatosMLP) == sizeof(tfg_datos_mlp), "DatosMLP layout");
static_assert(sizeof(Color) == sizeof(tfg_color), "Color layout");

namespace tfg_ref_detail {
template <typename To, typename From>
inline To pod_cast(const From& f) {
    static_assert(sizeof(To) == sizeof(From), "layout");
    To t;
    __builtin_memcpy(&t, &f, sizeof(To));
    return t;
}
}  // namespace tfg_ref_detail

// RenderKernel.h:55-61
inline void launchKernelRenderTiny(dim3, dim3, const Camara* camara, const Primitiva* primitivas, int num_primitivas,
                                   const LuzPuntual* luces, int num_luces, const Primitiva* primitivas_malla,
                                   int num_primitivas_malla, int ancho_img, int alto_img, int samples_per_pixel,
                                   int frame_number, TinyBVHD_GPU nodos_bvh, const Primitiva* primitivas_bvh,
                                   int num_nodos_bvh, ImagenGPU imagen_directa, unsigned int* dev_counter,
                                   RegistroEntrenamiento* buffer_registros, unsigned int* counter_train,
                                   int max_cap_train, DatosMLP* buffer_inference, Color* buffer_throughput,
                                   bool usar_red_inferencia, bool entrenar_red) {
    using tfg_ref_detail::pod_cast;
    (void)tfg_launch_render_tiny(
        reinterpret_cast<const tfg_camara*>(camara), reinterpret_cast<const tfg_primitiva*>(primitivas),
        num_primitivas, reinterpret_cast<const tfg_luz_puntual*>(luces), num_luces,
        reinterpret_cast<const tfg_primitiva*>(primitivas_malla), num_primitivas_malla, ancho_img, alto_img,
        samples_per_pixel, frame_number, pod_cast<tfg_cwbvh_gpu>(nodos_bvh),
        reinterpret_cast<const tfg_primitiva*>(primitivas_bvh), num_nodos_bvh, pod_cast<tfg_imagen_gpu>(imagen_directa),
        dev_counter, reinterpret_cast<tfg_registro_entrenamiento*>(buffer_registros), counter_train, max_cap_train,
        reinterpret_cast<tfg_datos_mlp*>(buffer_inference), reinterpret_cast<tfg_color*>(buffer_throughput),
        usar_red_inferencia ? 1 : 0, entrenar_red ? 1 : 0, nullptr);
}

// RenderKernel.h:63-65
inline void launchKernelComposite(dim3, dim3, Color* img_pt, Color* img_prediccion, Color* throughput_map, int ancho,
                                  int alto) {
    (void)tfg_launch_composite(reinterpret_cast<tfg_color*>(img_pt), reinterpret_cast<const tfg_color*>(img_prediccion),
                               reinterpret_cast<const tfg_color*>(throughput_map), ancho, alto, nullptr);
}

// RenderKernel.h:67
inline void launchInicializarCamara(Camara* d_camara, int ancho_imagen, int alto_imagen) {
    (void)tfg_launch_inicializar_camara(reinterpret_cast<tfg_camara*>(d_camara), ancho_imagen, alto_imagen, nullptr);
}

// RenderKernel.h:69-73
inline void launchKernelCalcularTargets(dim3, dim3, RegistroEntrenamiento* buffer_registros,
                                        Color* buffer_prediccion_tail, DatosMLP* buffer_training_final,
                                        int num_elementos) {
    (void)tfg_launch_calcular_targets(reinterpret_cast<const tfg_registro_entrenamiento*>(buffer_registros),
                                      reinterpret_cast<const tfg_color*>(buffer_prediccion_tail),
                                      reinterpret_cast<tfg_datos_mlp*>(buffer_training_final), num_elementos, nullptr);
}

// RenderKernel.h:75-78
inline void launchKernelPrepararInferenciaTail(dim3, dim3, const RegistroEntrenamiento* source_registros,
                                               DatosMLP* dest_inference_inputs, int num_elements) {
    (void)tfg_launch_preparar_inferencia_tail(reinterpret_cast<const tfg_registro_entrenamiento*>(source_registros),
                                              reinterpret_cast<tfg_datos_mlp*>(dest_inference_inputs), num_elements,
                                              nullptr);
}
