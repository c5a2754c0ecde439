// node_check.hip - the traversal's node test (tfg::node_hitmask in
// tfg_amd/csrc/trace_device.h), which reads the quantised child bounds as
// FP16 subnormals through v_fma_mix_f32 and compares slab distances scaled
// by 2^-24, against the unscaled test restated below (bounds converted with
// v_cvt_f32_ubyte, fma(b, idir, orig) compared with tmin and tcur) on the
// GPU: pseudo-random nodes and rays, a quarter of them with the ray origin
// on the node's origin planes (slab distances that are exact multiples of
// idir, zero included) and a quarter with a zero direction component (the
// reciprocal clamped to 1e30).  Prints the number of hit masks that differ
// and exits 0 only if there are none.  Built by tests/cpp/Makefile; run by
// tests/test_trace_gpu.py.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "trace_device.h"

#define CHECK(x)                                                                          \
    do {                                                                                  \
        const hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                           \
            std::printf("%s failed: %s\n", #x, hipGetErrorString(e_));                    \
            return 2;                                                                     \
        }                                                                                 \
    } while (0)

// The node test with unscaled slab distances (the form the oracle restates).
__device__ uint32_t hitmask_unscaled(const float4 n0, const float4 n1, const float4 n2, const float4 n3,
                                     const float4 n4, const tfg::RayPre& r, float tcur) {
    const uint32_t w = __float_as_uint(n0.w);
    const int ex = (int)(int8_t)(w & 0xFFu), ey = (int)(int8_t)((w >> 8) & 0xFFu), ez = (int)(int8_t)((w >> 16) & 0xFFu);
    const float idir[3] = {__uint_as_float((uint32_t)(ex + 127) << 23) * r.rD[0],
                           __uint_as_float((uint32_t)(ey + 127) << 23) * r.rD[1],
                           __uint_as_float((uint32_t)(ez + 127) << 23) * r.rD[2]};
    const float orig[3] = {(n0.x - r.O[0]) * r.rD[0], (n0.y - r.O[1]) * r.rD[1], (n0.z - r.O[2]) * r.rD[2]};
    const uint32_t meta[2] = {__float_as_uint(n1.z), __float_as_uint(n1.w)};
    const uint32_t lo[3][2] = {{__float_as_uint(n2.x), __float_as_uint(n2.y)},
                               {__float_as_uint(n2.z), __float_as_uint(n2.w)},
                               {__float_as_uint(n3.x), __float_as_uint(n3.y)}};
    const uint32_t hi[3][2] = {{__float_as_uint(n3.z), __float_as_uint(n3.w)},
                               {__float_as_uint(n4.x), __float_as_uint(n4.y)},
                               {__float_as_uint(n4.z), __float_as_uint(n4.w)}};
    uint32_t mask = 0;
    for (int h = 0; h < 2; ++h)
        for (int j = 0; j < 4; ++j) {
            float tnear = r.tmin, tfar = tcur;
            for (int a = 0; a < 3; ++a) {
                const bool neg = r.rD[a] < 0.0f;
                const float bn = (float)(((neg ? hi[a][h] : lo[a][h]) >> (8 * j)) & 0xFFu);
                const float bf = (float)(((neg ? lo[a][h] : hi[a][h]) >> (8 * j)) & 0xFFu);
                tnear = fmaxf(fmaf(bn, idir[a], orig[a]), tnear);
                tfar = fminf(fmaf(bf, idir[a], orig[a]), tfar);
            }
            if (!(tnear <= tfar * tfg::kBoxSlack)) continue;
            const uint32_t m = (meta[h] >> (8 * j)) & 0xFFu;
            const uint32_t inner = (m & (m << 1)) & 0x10u;
            const uint32_t bit = (inner ? (m ^ r.octinv) : m) & 0x1Fu;
            mask |= ((m >> 5) & 7u) << bit;
        }
    return mask;
}

__device__ uint32_t mix(uint32_t x) {   // integer hash (xorshift-multiply rounds)
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ float unit(uint32_t& s) {   // [-1, 1)
    s = mix(s + 0x9e3779b9u);
    return (float)(int32_t)s * 0x1p-31f;
}

__global__ void check(uint32_t base, uint32_t n, unsigned long long* bad, unsigned long long* hits) {
    unsigned long long mine = 0, nonzero = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t s = mix(base + i);
        const uint32_t kind = (base + i) & 3u;
        // node: origin in [-8, 8), per-axis exponents -9..-6 (extents 0.5-4), random metadata, and per
        // child a low bound byte and a high one at or above it (sometimes equal: flat boxes)
        const float nx = 8.0f * unit(s), ny = 8.0f * unit(s), nz = 8.0f * unit(s);
        s = mix(s);
        const int ex = -9 + (int)(s & 3u), ey = -9 + (int)((s >> 2) & 3u), ez = -9 + (int)((s >> 4) & 3u);
        const uint32_t e = (uint32_t)(uint8_t)(int8_t)ex | ((uint32_t)(uint8_t)(int8_t)ey << 8) |
                           ((uint32_t)(uint8_t)(int8_t)ez << 16) | 0xFF000000u;
        uint32_t wds[16];   // n1, then the low bound words (n2, n3.xy), then the high ones (n3.zw, n4)
        for (int k = 0; k < 4; ++k) wds[k] = s = mix(s);
        for (int a = 0; a < 3; ++a)
            for (int h = 0; h < 2; ++h) {
                uint32_t lo = 0, hi = 0;
                for (int j = 0; j < 4; ++j) {
                    s = mix(s);
                    const uint32_t l = s & 0xFFu, span = (s >> 8) & 0xFFu;
                    lo |= l << (8 * j);
                    hi |= (l + span > 255u ? 255u : l + span) << (8 * j);
                }
                wds[4 + 2 * a + h] = lo;
                wds[10 + 2 * a + h] = hi;
            }
        const float4 n0 = make_float4(nx, ny, nz, __uint_as_float(e));
        const float4 n1 = make_float4(__uint_as_float(wds[0]), __uint_as_float(wds[1]), __uint_as_float(wds[2]),
                                      __uint_as_float(wds[3]));
        const float4 n2 = make_float4(__uint_as_float(wds[4]), __uint_as_float(wds[5]), __uint_as_float(wds[6]),
                                      __uint_as_float(wds[7]));
        const float4 n3 = make_float4(__uint_as_float(wds[8]), __uint_as_float(wds[9]), __uint_as_float(wds[10]),
                                      __uint_as_float(wds[11]));
        const float4 n4 = make_float4(__uint_as_float(wds[12]), __uint_as_float(wds[13]), __uint_as_float(wds[14]),
                                      __uint_as_float(wds[15]));
        // ray: origin within 2 of the node's centre (on its origin planes for kind 1), direction with a zero
        // component for kind 2
        float ox = nx + 128.0f * ldexpf(1.0f, ex) + 2.0f * unit(s);
        float oy = ny + 128.0f * ldexpf(1.0f, ey) + 2.0f * unit(s);
        float oz = nz + 128.0f * ldexpf(1.0f, ez) + 2.0f * unit(s);
        if (kind == 1) {
            ox = nx;
            oy = ny;
            oz = nz;
        }
        float dx = unit(s), dy = unit(s), dz = unit(s);
        if (kind == 2) {
            const uint32_t a = mix(s) % 3u;
            if (a == 0) dx = 0.0f;
            else if (a == 1) dy = 0.0f;
            else dz = 0.0f;
        }
        const float tmax = 64.0f * (0.5f + 0.5f * unit(s)) + 0x1p-10f;
        const float tmin = kind == 3 ? 0x1p-8f * (1.0f + unit(s)) : 0.0f;
        const tfg::RayPre r = tfg::make_ray(ox, oy, oz, dx, dy, dz, tmin, tmax);
        const float tcur = tmax * (0.25f + 0.75f * (0.5f + 0.5f * unit(s)));
        const uint32_t a = tfg::node_hitmask(n0, n1, n2, n3, n4, r, tcur);
        const uint32_t b = hitmask_unscaled(n0, n1, n2, n3, n4, r, tcur);
        if (a != b) ++mine;
        if (b) ++nonzero;
    }
    if (mine) atomicAdd(bad, mine);
    if (nonzero) atomicAdd(hits, nonzero);
}

int main() {
    unsigned long long *bad = nullptr, *hits = nullptr;
    CHECK(hipMalloc(&bad, sizeof(unsigned long long)));
    CHECK(hipMalloc(&hits, sizeof(unsigned long long)));
    CHECK(hipMemset(bad, 0, sizeof(unsigned long long)));
    CHECK(hipMemset(hits, 0, sizeof(unsigned long long)));
    const uint32_t per = 1u << 24, launches = 8;
    for (uint32_t k = 0; k < launches; ++k) {
        hipLaunchKernelGGL(check, dim3(4096), dim3(256), 0, 0, k * per, per, bad, hits);
        CHECK(hipGetLastError());
    }
    CHECK(hipDeviceSynchronize());
    unsigned long long nb = 0, nh = 0;
    CHECK(hipMemcpy(&nb, bad, sizeof(nb), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&nh, hits, sizeof(nh), hipMemcpyDeviceToHost));
    CHECK(hipFree(bad));
    CHECK(hipFree(hits));
    std::printf("%u nodes x rays, %llu with a nonzero hit mask; %llu differ\n", per * launches, nh, nb);
    return nb == 0 ? 0 : 1;
}
