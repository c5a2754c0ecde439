// wave_ops.h - wave64 helpers shared by the frame kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfg {

// One append slot per requesting lane: one ballot, one atomic per wave,
// lane offsets by popcount.  All lanes of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    if (m == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

}  // namespace tfg
