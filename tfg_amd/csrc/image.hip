// image.hip - frame output on the GPU: mean of the FP64 accumulation buffer,
// ACES filmic curve and 8-bit quantisation (tfg_image_resolve).
//
// Replaces the host side of the reference's frame loop: the per-frame copy
// and FP32 host accumulation (main.cpp:1070-1071, 1109-1111), the filmic
// tone curve Imagen::acesFilmic / filmic (imagen.h:194-215) and the byte
// conversion of guardarPNG (imagen_utils.h:58-62).  Compiled with
// -ffp-contract=off so every FP32 operation rounds as written (the oracle
// restates it operation by operation).
#include <hip/hip_runtime.h>

#include "host_common.h"

namespace tfg {
namespace {

// imagen.h:194-205: x scaled by 0.6, then (x (a x + b)) / (x (c x + d) + e)
// clamped to [0, 1].
__device__ __forceinline__ float aces_filmic(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    x = x * 0.6f;
    const float y = (x * (a * x + b)) / (x * (c * x + d) + e);
    return fminf(fmaxf(y, 0.0f), 1.0f);
}

// imagen_utils.h:58-62: round(clamp(v, 0, 1) * 255), halves away from zero.
__device__ __forceinline__ uint8_t to_byte(float v) {
    const float c = fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f;
    return (uint8_t)roundf(c);
}

__global__ __launch_bounds__(256) void k_resolve(const double* __restrict__ accum, uint32_t n, double inv_samples,
                                                 uint8_t* __restrict__ rgb8, float* __restrict__ mean_out,
                                                 int tonemap) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float m[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) m[k] = (float)(accum[3 * (size_t)i + k] * inv_samples);
        if (mean_out) {
#pragma unroll
            for (int k = 0; k < 3; ++k) mean_out[3 * (size_t)i + k] = m[k];
        }
        if (rgb8) {
#pragma unroll
            for (int k = 0; k < 3; ++k) rgb8[3 * (size_t)i + k] = to_byte(tonemap ? aces_filmic(m[k]) : m[k]);
        }
    }
}

}  // namespace
}  // namespace tfg

extern "C" int tfg_image_resolve(const double* d_accum, uint32_t n_pixels, uint32_t samples, int tonemap,
                                 uint8_t* d_rgb8, float* d_mean, tfg_stream stream) {
    if (!d_accum || (!d_rgb8 && !d_mean)) {
        tfg::set_error("tfg_image_resolve: NULL argument");
        return TFG_E_INVALID;
    }
    if (samples == 0) {
        tfg::set_error("tfg_image_resolve: samples must be > 0");
        return TFG_E_INVALID;
    }
    if (n_pixels == 0) return TFG_OK;
    const uint32_t blocks = (n_pixels + 255) / 256 < 4096 ? (n_pixels + 255) / 256 : 4096;
    hipLaunchKernelGGL(tfg::k_resolve, dim3(blocks), dim3(256), 0, stream, d_accum, n_pixels, 1.0 / (double)samples,
                       d_rgb8, d_mean, tonemap);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}
