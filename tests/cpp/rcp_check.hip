// rcp_check.hip - the traversal's ray set-up reciprocal (tfg::rcp_rn in
// tfg_amd/csrc/trace_device.h) against the IEEE division 1.0f / d on the
// GPU: every normal d (both signs, all 2^23 significands of every exponent
// -126..127, so both the v_rcp_f32 + Newton path and the division taken for
// |d| >= 2^126), plus zeros, subnormals, infinities and NaN.  Prints the
// number of values whose bits differ (NaN: both results must be NaN) and
// exits 0 only if there are none.  Built by tests/cpp/Makefile; run by
// tests/test_trace_gpu.py.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "trace_device.h"

#define CHECK(x)                                                                          \
    do {                                                                                  \
        const hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                           \
            std::printf("%s failed: %s\n", #x, hipGetErrorString(e_));                    \
            return 2;                                                                     \
        }                                                                                 \
    } while (0)

__device__ bool same(float a, float b) {
    if (a != a) return b != b;
    return __float_as_uint(a) == __float_as_uint(b);
}

// Thread w: significand w & (2^23 - 1), exponent field 1 + (w >> 24), sign
// (w >> 23) & 1: 254 exponent fields x 2 signs x 2^23 values.
__global__ void check_normals(unsigned long long total, unsigned long long* bad) {
    unsigned long long mine = 0;
    for (unsigned long long w = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; w < total;
         w += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t bits = (uint32_t)(((w >> 23) & 1ull) << 31) | (uint32_t)((1ull + (w >> 24)) << 23) |
                              (uint32_t)(w & 0x7FFFFFull);
        const float d = __uint_as_float(bits);
        if (!same(tfg::rcp_rn(d), 1.0f / d)) ++mine;
    }
    if (mine) atomicAdd(bad, mine);
}

__global__ void check_list(const uint32_t* bits, uint32_t n, unsigned long long* bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = __uint_as_float(bits[i]);
        if (!same(tfg::rcp_rn(d), 1.0f / d)) atomicAdd(bad, 1ull);
    }
}

int main() {
    const unsigned long long total = 254ull * 2ull << 23;
    std::vector<uint32_t> special = {0x00000000u, 0x80000000u, 0x7F800000u, 0xFF800000u, 0x7FC00000u, 0xFFC00000u,
                                     0x7F800001u};
    for (uint32_t m = 1; m < (1u << 23); m += 4099u) {   // subnormals, both signs
        special.push_back(m);
        special.push_back(0x80000000u | m);
    }
    special.push_back(0x007FFFFFu);
    special.push_back(0x807FFFFFu);
    unsigned long long* d_bad = nullptr;
    uint32_t* d_special = nullptr;
    CHECK(hipMalloc(&d_bad, 2 * sizeof(unsigned long long)));
    CHECK(hipMalloc(&d_special, special.size() * sizeof(uint32_t)));
    CHECK(hipMemset(d_bad, 0, 2 * sizeof(unsigned long long)));
    CHECK(hipMemcpy(d_special, special.data(), special.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(check_normals, dim3(8192), dim3(256), 0, 0, total, d_bad);
    CHECK(hipGetLastError());
    hipLaunchKernelGGL(check_list, dim3(64), dim3(256), 0, 0, d_special, (uint32_t)special.size(), d_bad + 1);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    unsigned long long bad[2] = {0, 0};
    CHECK(hipMemcpy(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost));
    CHECK(hipFree(d_bad));
    CHECK(hipFree(d_special));
    std::printf("rcp_rn vs 1.0f / d: %llu normal values, %llu differ; %zu special values, %llu differ\n", total, bad[0],
                special.size(), bad[1]);
    return bad[0] == 0 && bad[1] == 0 ? 0 : 1;
}
