// wave_ops.h - wave64 / block helpers shared by the frame kernels.
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

// XCD-aware order of `chunks` work chunks over a grid.  The GPU hands
// workgroups to its 8 XCDs round robin (workgroup b runs on XCD b % 8) and
// each XCD has its own L2.  Chunks are grouped into super-chunks of `group`
// consecutive chunks; super-chunk j belongs to XCD j % 8, whose workgroups
// share out its chunks.  Neighbouring rays (neighbouring pixels, for the
// queues here) thus meet on one L2, while every XCD still gets work from the
// whole frame.  Usage:
//     for (ChunkIter it(chunks, group); it.valid(); it.next()) ... it.chunk ...
// Every chunk is visited exactly once whatever the real placement (with
// enable false, or a grid that is not a multiple of 8, the plain grid stride).
constexpr uint32_t kXcds = 8;
struct ChunkIter {
    uint32_t chunks, group, xcd, step, k, chunk;
    bool xcd_order;
    __device__ __forceinline__ ChunkIter(uint32_t n_chunks, uint32_t group_chunks, bool enable = true)
        : chunks(n_chunks), group(group_chunks ? group_chunks : 1u) {
        xcd_order = enable && gridDim.x % kXcds == 0;
        xcd = blockIdx.x % kXcds;
        step = xcd_order ? gridDim.x / kXcds : gridDim.x;
        k = xcd_order ? blockIdx.x / kXcds : blockIdx.x;
        map();
    }
    __device__ __forceinline__ void map() {
        chunk = xcd_order ? (xcd + kXcds * (k / group)) * group + k % group : k;
    }
    __device__ __forceinline__ bool valid() const { return chunk < chunks; }
    __device__ __forceinline__ void next() {
        k += step;
        map();
    }
};

// Two appends aggregated over the whole block: one ballot per wave, wave
// offsets by an LDS scan, one atomic per block and counter (kBlock / 64
// times fewer atomics on the two counters than wave_append).  Every thread
// of the block must call it: it synchronises the block.
template <int kBlock>
__device__ __forceinline__ void block_append2(bool want_a, uint32_t* counter_a, uint32_t& slot_a, bool want_b,
                                              uint32_t* counter_b, uint32_t& slot_b) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t wave_off[2][kWaves];
    __shared__ uint32_t block_base[2];
    const uint64_t ma = __ballot(want_a), mb = __ballot(want_b);
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
    if (lane == 0) {
        wave_off[0][wave] = (uint32_t)__popcll(ma);
        wave_off[1][wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wave_off[threadIdx.x][w];
            wave_off[threadIdx.x][w] = tot;
            tot += c;
        }
        block_base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? counter_b : counter_a, tot) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    slot_a = block_base[0] + wave_off[0][wave] + (uint32_t)__popcll(ma & below);
    slot_b = block_base[1] + wave_off[1][wave] + (uint32_t)__popcll(mb & below);
    __syncthreads();   // the LDS words are reused by the next call
}

// block_append2, plus each counter's first slot for the block and the
// block's count: the block's slots of a counter are base .. base + n - 1,
// ranked in lane order.  Every thread of the block must call it.
template <int kBlock>
__device__ __forceinline__ void block_append2_range(bool want_a, uint32_t* counter_a, uint32_t& slot_a,
                                                    uint32_t& base_a, uint32_t& n_a, bool want_b,
                                                    uint32_t* counter_b, uint32_t& slot_b, uint32_t& base_b,
                                                    uint32_t& n_b) {
    constexpr int kWaves = kBlock / 64;
    __shared__ uint32_t wave_off[2][kWaves];
    __shared__ uint32_t block_base[2], block_n[2];
    const uint64_t ma = __ballot(want_a), mb = __ballot(want_b);
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
    if (lane == 0) {
        wave_off[0][wave] = (uint32_t)__popcll(ma);
        wave_off[1][wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wave_off[threadIdx.x][w];
            wave_off[threadIdx.x][w] = tot;
            tot += c;
        }
        block_n[threadIdx.x] = tot;
        block_base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? counter_b : counter_a, tot) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    base_a = block_base[0];
    base_b = block_base[1];
    n_a = block_n[0];
    n_b = block_n[1];
    slot_a = base_a + wave_off[0][wave] + (uint32_t)__popcll(ma & below);
    slot_b = base_b + wave_off[1][wave] + (uint32_t)__popcll(mb & below);
    __syncthreads();   // the LDS words are reused by the next call
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Adds the block's sum of v to *counter with one atomic.  Every thread of
// the block must call it: it synchronises the block.
template <int kBlock>
__device__ __forceinline__ void block_add(uint32_t v, uint32_t* counter) {
    __shared__ uint32_t partial[kBlock / 64];
    const uint32_t s = wave_sum(v);
    if (__lane_id() == 0) partial[threadIdx.x / 64u] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += partial[w];
        if (t) atomicAdd(counter, t);
    }
    __syncthreads();
}

}  // namespace tfg
