// ref_kernels.hip - the reference's small launchers on the reference's own
// struct layouts (DatosMLP, RegistroEntrenamiento, Camara, Color, SceneBounds):
//
//   tfg_launch_inicializar_camara        RenderKernel.cu:68-90
//   tfg_launch_composite                 RenderKernel.cu:49-66
//   tfg_launch_calcular_targets          RenderKernel.cu:92-137
//   tfg_launch_preparar_inferencia_tail  RenderKernel.cu:139-166
//   tfg_launch_generar_indices           NrcKernel.cu:10-16
//   tfg_launch_preparar_entrenamiento    NrcKernel.cu:18-68
//   tfg_launch_preparar_inferencia       NrcKernel.cu:70-97
//   tfg_launch_guardar_salida            NrcKernel.cu:99-126
//
// Grid and block come from the library (grid-stride loops over n), not from
// the caller: the reference's (blocks, threads) arguments are accepted by
// the C++ shims and ignored.
#include <hip/hip_runtime.h>

#include "host_common.h"
#include "nrc_device.h"
#include "shade_device.h"

namespace tfg {
namespace {

constexpr int kB = 256;

inline dim3 grid_for(uint64_t n) {
    uint64_t g = (n + kB - 1) / kB;
    if (g > 65535ull * 4) g = 65535ull * 4;
    return dim3((unsigned)(g ? g : 1));
}

__device__ __forceinline__ double d_len(tfg_vec3d a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ tfg_vec3d d_norm(tfg_vec3d a) {
    const double l = d_len(a);
    return l > 0.0 ? tfg_vec3d{a.x / l, a.y / l, a.z / l} : a;
}
__device__ __forceinline__ tfg_vec3d d_cross(tfg_vec3d a, tfg_vec3d b) {
    return tfg_vec3d{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__global__ void k_init_camera(tfg_camara* cam, int w, int h) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float focal = 1.0f;
    const float fov = 45.0f * 3.14159265358979323846 / 180.0f;
    const float alto = 2.0f * focal * tanf(fov / 2.0f);
    const float ancho = alto * ((float)w / (float)h);
    const tfg_vec3d pos{0.0, -0.75, 4.25};
    const tfg_vec3d f = d_norm(tfg_vec3d{0.0, 0.0, -1.0});
    const tfg_vec3d up0 = d_norm(tfg_vec3d{0.0, 1.0, 0.0});
    const tfg_vec3d left = d_norm(d_cross(f, up0));
    const tfg_vec3d up = d_norm(d_cross(left, f));
    cam->posicion = pos;
    cam->frente = f;
    cam->arriba = up;
    cam->izquierda = left;
    cam->ancho_plano = ancho;
    cam->alto_plano = alto;
    cam->distancia_focal = focal;
}

__global__ void k_composite(tfg_color* img, const tfg_color* pred, const tfg_color* thr, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_color th = thr[i];
        if (th.r > 0 || th.g > 0 || th.b > 0) {
            const tfg_color p = pred[i];
            tfg_color c = img[i];
            c.r += th.r * p.r;
            c.g += th.g * p.g;
            c.b += th.b * p.b;
            img[i] = c;
        }
    }
}

__device__ __forceinline__ float clean_target(float t) {
    t = fminf(t, 100000.0f);
    if (isnan(t) || isinf(t)) t = 0.0f;
    return t;
}

__global__ void k_targets(const tfg_registro_entrenamiento* recs, const tfg_color* pred_tail, tfg_datos_mlp* out,
                          uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_registro_entrenamiento& r = recs[i];
        if (!r.valido) {
            out[i].color = tfg_color{0.f, 0.f, 0.f};
            continue;
        }
        const tfg_color p = pred_tail[i];
        const float eps = 1e-8f;
        tfg_datos_mlp d;
        d.color.r = clean_target((r.luz_acumulada_sufijo.r + r.throughput_sufijo.r * p.r) / fmaxf(r.factor_normalizacion.r, eps));
        d.color.g = clean_target((r.luz_acumulada_sufijo.g + r.throughput_sufijo.g * p.g) / fmaxf(r.factor_normalizacion.g, eps));
        d.color.b = clean_target((r.luz_acumulada_sufijo.b + r.throughput_sufijo.b * p.b) / fmaxf(r.factor_normalizacion.b, eps));
        d.posicion = r.head.posicion;
        d.direccion = r.head.direccion;
        d.normal = r.head.normal;
        d.difuso = r.head.difuso;
        d.especular = r.head.especular;
        out[i] = d;
    }
}

__global__ void k_tail_inputs(const tfg_registro_entrenamiento* recs, tfg_datos_mlp* out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_registro_entrenamiento& r = recs[i];
        tfg_datos_mlp d{};
        if (r.valido) {
            d.posicion = r.tail.posicion;
            d.direccion = r.tail.direccion;
            d.normal = r.tail.normal;
            d.difuso = r.tail.difuso;
            d.especular = r.tail.especular;
        }
        out[i] = d;
    }
}

__global__ void k_indices(int* idx, int max_range, int count, uint32_t seed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        uint32_t s = seed ^ (uint32_t)i;
        idx[i] = (int)(pcg32(s) % (uint32_t)max_range);
    }
}

__device__ __forceinline__ float norm_coord(float v, float mn, float mx) {
    return fminf(fmaxf((v - mn) / (mx - mn), 0.0f), 1.0f);
}

__device__ __forceinline__ void pack15(const tfg_datos_mlp& d, const tfg_scene_bounds& b, float* o) {
    o[0] = norm_coord((float)d.posicion.x, (float)b.min.x, (float)b.max.x);
    o[1] = norm_coord((float)d.posicion.y, (float)b.min.y, (float)b.max.y);
    o[2] = norm_coord((float)d.posicion.z, (float)b.min.z, (float)b.max.z);
    o[3] = (float)d.direccion.x;
    o[4] = (float)d.direccion.y;
    o[5] = (float)d.direccion.z;
    o[6] = (float)d.normal.x;
    o[7] = (float)d.normal.y;
    o[8] = (float)d.normal.z;
    o[9] = d.difuso.r;
    o[10] = d.difuso.g;
    o[11] = d.difuso.b;
    o[12] = d.especular.r;
    o[13] = d.especular.g;
    o[14] = d.especular.b;
}

__global__ void k_prep_train(const tfg_datos_mlp* buf, const int* idx, int batch, int n_in, int n_out,
                             tfg_scene_bounds b, float* in, float* tgt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < batch; i += gridDim.x * blockDim.x) {
        const tfg_datos_mlp d = buf[idx[i]];
        pack15(d, b, in + (size_t)i * n_in);
        float r = d.color.r, g = d.color.g, bl = d.color.b;
        if (r > 100000.0f) r = 100000.0f;
        if (g > 100000.0f) g = 100000.0f;
        if (bl > 100000.0f) bl = 100000.0f;
        float* t = tgt + (size_t)i * n_out;
        t[0] = r;
        t[1] = g;
        t[2] = bl;
    }
}

__global__ void k_prep_infer(const tfg_datos_mlp* buf, uint32_t n, uint32_t n_in, float* in, tfg_scene_bounds b) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        pack15(buf[i], b, in + (size_t)i * n_in);
}

__global__ void k_save_output(const float* out, tfg_color* col, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* o = out + 3 * (size_t)i;
        col[i] = tfg_color{sanitize_out(o[0]), sanitize_out(o[1]), sanitize_out(o[2])};
    }
}

}  // namespace
}  // namespace tfg

using namespace tfg;

#define TFG_LAUNCH_CHECK(name)                         \
    do {                                               \
        hipError_t _e = hipGetLastError();             \
        if (_e != hipSuccess) return hip_fail(_e, name); \
        return TFG_OK;                                 \
    } while (0)

extern "C" {

int tfg_launch_inicializar_camara(tfg_camara* d_cam, int width, int height, tfg_stream stream) {
    if (!d_cam || width <= 0 || height <= 0) { set_error("tfg_launch_inicializar_camara: bad argument"); return TFG_E_INVALID; }
    hipLaunchKernelGGL(k_init_camera, dim3(1), dim3(1), 0, stream, d_cam, width, height);
    TFG_LAUNCH_CHECK("k_init_camera");
}

int tfg_launch_composite(tfg_color* d_img, const tfg_color* d_pred, const tfg_color* d_thr, int width, int height,
                         tfg_stream stream) {
    if (!d_img || !d_pred || !d_thr || width < 0 || height < 0) { set_error("tfg_launch_composite: bad argument"); return TFG_E_INVALID; }
    const uint32_t n = (uint32_t)width * (uint32_t)height;
    if (n == 0) return TFG_OK;
    hipLaunchKernelGGL(k_composite, grid_for(n), dim3(kB), 0, stream, d_img, d_pred, d_thr, n);
    TFG_LAUNCH_CHECK("k_composite");
}

int tfg_launch_calcular_targets(const tfg_registro_entrenamiento* d_recs, const tfg_color* d_pred_tail,
                                tfg_datos_mlp* d_out, int n, tfg_stream stream) {
    if (!d_recs || !d_pred_tail || !d_out || n < 0) { set_error("tfg_launch_calcular_targets: bad argument"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    hipLaunchKernelGGL(k_targets, grid_for(n), dim3(kB), 0, stream, d_recs, d_pred_tail, d_out, (uint32_t)n);
    TFG_LAUNCH_CHECK("k_targets");
}

int tfg_launch_preparar_inferencia_tail(const tfg_registro_entrenamiento* d_recs, tfg_datos_mlp* d_out, int n,
                                        tfg_stream stream) {
    if (!d_recs || !d_out || n < 0) { set_error("tfg_launch_preparar_inferencia_tail: bad argument"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    hipLaunchKernelGGL(k_tail_inputs, grid_for(n), dim3(kB), 0, stream, d_recs, d_out, (uint32_t)n);
    TFG_LAUNCH_CHECK("k_tail_inputs");
}

int tfg_launch_generar_indices(int* d_idx, int max_range, int count, uint32_t seed, tfg_stream stream) {
    if (!d_idx || max_range <= 0 || count < 0) { set_error("tfg_launch_generar_indices: bad argument"); return TFG_E_INVALID; }
    if (count == 0) return TFG_OK;
    hipLaunchKernelGGL(k_indices, grid_for(count), dim3(kB), 0, stream, d_idx, max_range, count, seed);
    TFG_LAUNCH_CHECK("k_indices");
}

int tfg_launch_preparar_entrenamiento(const tfg_datos_mlp* d_buf, const int* d_idx, int batch, int n_in, int n_out,
                                      tfg_scene_bounds bounds, float* d_in, float* d_tgt, tfg_stream stream) {
    if (!d_buf || !d_idx || !d_in || !d_tgt || batch < 0 || n_in < 15 || n_out < 3) {
        set_error("tfg_launch_preparar_entrenamiento: bad argument (n_in >= 15, n_out >= 3)");
        return TFG_E_INVALID;
    }
    if (batch == 0) return TFG_OK;
    hipLaunchKernelGGL(k_prep_train, grid_for(batch), dim3(kB), 0, stream, d_buf, d_idx, batch, n_in, n_out, bounds,
                       d_in, d_tgt);
    TFG_LAUNCH_CHECK("k_prep_train");
}

int tfg_launch_preparar_inferencia(const tfg_datos_mlp* d_buf, uint32_t n, uint32_t n_in, float* d_in,
                                   tfg_scene_bounds bounds, tfg_stream stream) {
    if (!d_buf || !d_in || n_in < 15) { set_error("tfg_launch_preparar_inferencia: bad argument (n_in >= 15)"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    hipLaunchKernelGGL(k_prep_infer, grid_for(n), dim3(kB), 0, stream, d_buf, n, n_in, d_in, bounds);
    TFG_LAUNCH_CHECK("k_prep_infer");
}

int tfg_launch_guardar_salida(const float* d_out, tfg_color* d_col, int n, tfg_stream stream) {
    if (!d_out || !d_col || n < 0) { set_error("tfg_launch_guardar_salida: bad argument"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    hipLaunchKernelGGL(k_save_output, grid_for(n), dim3(kB), 0, stream, d_out, d_col, (uint32_t)n);
    TFG_LAUNCH_CHECK("k_save_output");
}

}  // extern "C"
