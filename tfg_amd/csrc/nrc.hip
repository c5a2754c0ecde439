// nrc.hip - online-trained radiance cache: fused MFMA inference, fused
// training step (tail inference + targets + forward + backward + per-block
// weight gradients), gradient reduction and Adam/EMA on FP32 master weights.
//
// Restates mlp.cpp:95-170 (train_step / inference), NrcKernel.cu:10-126
// (batch indices, input packing, output sanitising) and
// RenderKernel.cu:92-166 (tail inference and training targets) as one
// graph-capturable sequence per frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "host_common.h"
#include "nrc_device.h"
#include "nrc_iface.h"
#include "pt_internal.h"
#include "shade_device.h"

#include <mutex>

using namespace tfg;

// Scratch memory of tfg_nrc_infer_datos_noalloc: per query one vertex record
// (kVertF4 float4) followed by the 3 raw network outputs of all queries.
// Every GPU the library runs on holds its own copy from module load until
// the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_ptr below).
constexpr uint32_t kNoAllocMaxQueries = TFG_NRC_NOALLOC_MAX_QUERIES;
constexpr size_t kScratchVertF4 = (size_t)kNoAllocMaxQueries * kVertF4;
constexpr size_t kScratchF4 = kScratchVertF4 + (3 * (size_t)kNoAllocMaxQueries + 3) / 4;
static __device__ __align__(256) float4 scratch[kScratchF4];

struct tfg_nrc {
    tfg_nrc_config cfg{};
    NrcBounds bounds{};
    // FP32 optimiser state (kNrcParams each)
    float* master = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* ema = nullptr;
    float* grad = nullptr;
    float* partial = nullptr;      // one kNrcParams gradient per training workgroup
    float* block_loss = nullptr;   // one loss per training workgroup
    float* loss_sum = nullptr;     // their sum, for the step's closing
    // prepared samples (head vertex + target) of one call's steps
    float4* prep = nullptr;
    size_t prep_cap = 0;           // samples
    // operand layouts, 16-bit (bf16 or fp16 by cfg.precision)
    void* fop_train = nullptr;
    void* bop_train = nullptr;
    void* fop_infer = nullptr;
    // device scalars: [0] step count t, [1] seed counter, [2] flags of the
    // current step, [3] skipped (non-finite) steps, [4] no-data steps,
    // [5] loss of the last step (float bits), [6] steps taken
    uint32_t* state = nullptr;
    int n_cu = 0;
    // DatosMLP entry points (ColorMLP::train_step / inference): converted
    // inputs, outputs and the sample count, grown on demand
    float4* dm_buf = nullptr;
    float* dm_out = nullptr;
    uint32_t* dm_count = nullptr;
    size_t dm_cap = 0;
};

namespace tfg {
namespace {

constexpr int kStLt = 0, kStSeed = 1, kStFlags = 2, kStSkipNan = 3, kStSkipData = 4, kStLoss = 5, kStSteps = 6,
              kStDone = 7;
constexpr size_t kOpBytes16 = 16;   // one operand tile entry: 8 x 16-bit per lane

struct Hyper {
    float lr, beta1, beta2, eps, ema_decay, inv_loss_scale;
};

__device__ __forceinline__ uint32_t pcg32_step(uint32_t s) {
    s = s * 747796405u + 2891336453u;
    const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (word >> 22u) ^ word;
}

// Element offsets (in 16-bit units) of parameter p in the operand layouts.
__host__ __device__ inline void slot_of(int idx, int& c, int& g, int& j) {
    c = idx / 32;
    const int r = idx % 32;
    if (r < 16) { g = r / 4; j = r % 4; }
    else { g = (r - 16) / 4; j = 4 + (r - 16) % 4; }
}
__host__ __device__ inline int param_layer(int p) { return p < 5 * 4096 ? p / 4096 : 5; }
__host__ __device__ inline size_t fop_pos(int p) {
    const int l = param_layer(p);
    const int r = p - nrc_layer_offset(l);
    const int n = r / 64, k = r % 64;
    int c, g, j;
    slot_of(k, c, g, j);
    const int lane = 16 * g + n % 16;
    return ((size_t)(fop_tile_base(l) + 2 * (n / 16) + c) * 64 + lane) * 8 + j;
}
__host__ __device__ inline size_t bop_pos(int p) {   // layers 1..5
    const int l = param_layer(p);
    const int r = p - nrc_layer_offset(l);
    const int n = r / 64, k = r % 64;
    int c, g, j;
    slot_of(n, c, g, j);
    const int lane = 16 * g + k % 16;
    const int u = k / 16;
    const int tile = bop_tile_base(l) + (l < 5 ? 2 * u + c : u);
    return ((size_t)tile * 64 + lane) * 8 + j;
}

// Copies `kTiles` operand tiles (64 x 16 B each) into LDS with every load of
// a thread in flight before its first store: one memory latency per copy
// instead of one per 16 B entry.
// (The entries move as plain 16-byte words: copied as vectors of 16-bit
// floats they were split into halves and re-packed.)
template <typename T, int kTiles, int kThreads>
__device__ __forceinline__ void load_ops(typename MfmaOps<T>::vec* dst, const typename MfmaOps<T>::vec* src) {
    constexpr int kN = kTiles * 64, kPer = (kN + kThreads - 1) / kThreads;
    const uint4v* s = reinterpret_cast<const uint4v*>(src);
    uint4v* d = reinterpret_cast<uint4v*>(dst);
    uint4v tmp[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + k * kThreads;
        if (i < kN) tmp[k] = s[i];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + k * kThreads;
        if (i < kN) d[i] = tmp[k];
    }
}

// ------------------------------------------------------------- inference
// Queries (kQueryF4 float4: vertex + (throughput, path id)) -> cache
// prediction, sanitised, composited into the image (RenderKernel.cu:49-66).
// Workgroups of 8 waves share one LDS copy of the weights (42 KB: 3
// workgroups per CU); a wave infers 16 queries at a time (G = 1: 68 VGPRs,
// 6 waves per SIMD), which measured faster than 2 or 4 groups of 16 per wave
// with their fewer waves: 505 vs 417 / 376 TFLOP/s on a 1080p frame's queries.
constexpr int kInferBlock = 512;
constexpr int kInferGroups = 1;
template <typename T, int G>
__global__ __launch_bounds__(kInferBlock) void k_nrc_infer(const float4* __restrict__ queries, const uint32_t* __restrict__ count,
                                                   tfg_color* __restrict__ image,
                                                   const typename MfmaOps<T>::vec* __restrict__ fop, NrcBounds b) {
    typedef typename MfmaOps<T>::vec V;
    __shared__ V W[kFopTiles * 64];
    load_ops<T, kFopTiles, kInferBlock>(W, fop);
    __syncthreads();
    const uint32_t n = *count;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, ml = lane & 15;
    constexpr uint32_t kWaves = kInferBlock / 64;
    for (uint32_t base = (blockIdx.x * kWaves + w) * 16u * G; base < n; base += gridDim.x * kWaves * 16u * G) {
        V x[G][2];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const uint32_t s = base + 16 * q + ml;
            if (s < n) {
                const float4* Q = queries + (size_t)kQueryF4 * s;
                encode<T>(Q[0], Q[1], Q[2], Q[3], b, g, x[q][0], x[q][1]);
            } else {
                x[q][0] = V{};
                x[q][1] = V{};
            }
        }
        floatx4 out[G];
        mlp_forward<T, G>(W, lane, x, out);
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const uint32_t s = base + 16 * q + ml;
                if (s >= n) continue;
                const float4 th = queries[(size_t)kQueryF4 * s + 4];
                if (th.x > 0.0f || th.y > 0.0f || th.z > 0.0f) {
                    const uint32_t pid = __float_as_uint(th.w);
                    tfg_color c = image[pid];
                    c.r += th.x * sanitize_out(out[q][0]);
                    c.g += th.y * sanitize_out(out[q][1]);
                    c.b += th.z * sanitize_out(out[q][2]);
                    image[pid] = c;
                }
            }
        }
    }
}

// Vertices (4 float4 each) -> raw outputs (3 floats each); optionally the
// pre-activation of every layer (n x 6 x 64 floats) for the oracle checks.
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_eval(const float4* __restrict__ verts, uint32_t n, float* __restrict__ out,
                                                  float* __restrict__ dump,
                                                  const typename MfmaOps<T>::vec* __restrict__ fop, NrcBounds b) {
    typedef MfmaOps<T> M;
    typedef typename M::vec V;
    __shared__ V W[kFopTiles * 64];
    load_ops<T, kFopTiles, 256>(W, fop);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, ml = lane & 15;
    for (uint32_t base = (blockIdx.x * 4 + w) * 16u; base < n; base += gridDim.x * 64u) {
        const uint32_t s = base + ml;
        V x[2] = {V{}, V{}};
        if (s < n) encode<T>(verts[4 * (size_t)s], verts[4 * (size_t)s + 1], verts[4 * (size_t)s + 2],
                             verts[4 * (size_t)s + 3], b, g, x[0], x[1]);
        for (int l = 0; l < kNrcLayers; ++l) {
            const int ntiles = l < 5 ? 4 : 1;
            floatx4 acc[4];
            for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int t = 0; t < ntiles; ++t)
                for (int c = 0; c < 2; ++c) acc[t] = M::mfma(W[(fop_tile_base(l) + 2 * t + c) * 64 + lane], x[c], acc[t]);
            if (dump && s < n)
                for (int t = 0; t < ntiles; ++t)
                    for (int i = 0; i < 4; ++i) dump[((size_t)s * kNrcLayers + l) * 64 + 16 * t + 4 * g + i] = acc[t][i];
            if (l < 5) {
                x[0] = pack_relu<T>(acc[0], acc[1]);
                x[1] = pack_relu<T>(acc[2], acc[3]);
            } else if (g == 0 && s < n) {
                for (int i = 0; i < 3; ++i) out[3 * (size_t)s + i] = acc[0][i];
            }
        }
    }
}

// ------------------------------------------------------------- training
// Prepared training sample: the head vertex (kVertF4 float4) and its target
// (rgb, 0), what an optimiser step reads per sample.
constexpr int kPrepF4 = kVertF4 + 1;

// Training records -> the prepared samples of `steps` optimiser steps.  Step
// k (seed = the seed counter + k: every step advances the counter by one,
// mlp.cpp:102) draws records[pcg32(seed ^ i) % n] for its sample i
// (generarIndicesAleatorios, NrcKernel.cu:10-16).  The sample's target
// (L_suf + T_suf * cache(tail)) / max(normaliser, 1e-8), clamped to 1e5,
// non-finite -> 0 (kernelPrepararInferenciaTail + ColorMLP::inference +
// kernelCalcularTargets, RenderKernel.cu:92-166, main.cpp:1044-1052), uses
// the inference (EMA) weights as they are when this kernel runs: once per
// call, before its steps, as the reference computes a frame's targets once
// before its step.  Workgroups of 8 waves share one LDS copy of the weights
// and loop over the samples (the weights are staged once per workgroup, not
// once per 128 samples); a wave prepares 16 samples at a time, lane group g
// copying head float4 g.
struct PrepArgs {
    const float4* records;
    const uint32_t* n_rec;
    uint32_t min_records;
    const uint32_t* state;
    uint32_t batch, steps;
    NrcBounds b;
    const void* fop;
    float4* out;       // kPrepF4 float4 per sample, samples of step k at k * batch
    uint32_t* out_n;   // nullable: receives the record count
};

constexpr int kPrepBlock = 512;
template <typename T>
__global__ __launch_bounds__(kPrepBlock) void k_nrc_prepare(PrepArgs A) {
    typedef typename MfmaOps<T>::vec V;
    __shared__ V W[kFopTiles * 64];
    const uint32_t n = *A.n_rec;
    if (A.out_n && blockIdx.x == 0 && threadIdx.x == 0) *A.out_n = n;
    if (n <= A.min_records) return;   // the steps will skip for lack of data
    load_ops<T, kFopTiles, kPrepBlock>(W, static_cast<const V*>(A.fop));
    __syncthreads();
    const uint32_t seed0 = A.state[kStSeed];
    const uint32_t total = A.batch * A.steps;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, ml = lane & 15;
    constexpr uint32_t kWaves = kPrepBlock / 64;
    for (uint32_t base = (blockIdx.x * kWaves + w) * 16u; base < total; base += gridDim.x * kWaves * 16u) {
        const uint32_t j = base + ml;
        const bool valid = j < total;
        const float4* R = A.records;
        V x[1][2];
        x[0][0] = V{};
        x[0][1] = V{};
        if (valid) {
            const uint32_t k = j / A.batch, i = j - k * A.batch;
            R += (size_t)kRecF4 * (pcg32_step((seed0 + k) ^ i) % n);
            encode<T>(R[4], R[5], R[6], R[7], A.b, g, x[0][0], x[0][1]);
        }
        floatx4 out[1];
        mlp_forward<T, 1>(W, lane, x, out);   // the whole wave: matrix-core operations
        if (!valid) continue;
        float4* o = A.out + (size_t)kPrepF4 * j;
        o[g] = R[g];
        if (g == 0) {
            const float4 ls = R[8], ts = R[9], nm = R[10];
            const float lsv[3] = {ls.x, ls.y, ls.z}, tsv[3] = {ts.x, ts.y, ts.z}, nmv[3] = {nm.x, nm.y, nm.z};
            float tgt[3];
            for (int c = 0; c < 3; ++c) {
                float t = (lsv[c] + tsv[c] * sanitize_out(out[0][c])) / fmaxf(nmv[c], 1e-8f);
                t = fminf(t, 100000.0f);
                if (isnan(t) || isinf(t)) t = 0.0f;
                tgt[c] = t;
            }
            o[kVertF4] = make_float4(tgt[0], tgt[1], tgt[2], 0.0f);
        }
    }
}

// One optimiser step's forward and backward passes.  A workgroup holds 128
// samples (8 waves of 16), one LDS copy of the training weights (forward and
// transposed operands, 78 KB) and the transposed dy / x tiles of its samples
// (34 KB): one workgroup per CU, 2 waves per SIMD.  Per layer, from the
// output back, each wave forms its samples' input gradient in registers,
// then after a barrier two 16 x 16 tiles of the weight gradient summed over
// the workgroup's 128 samples (4 matrix-core steps of 32 samples each),
// written as this workgroup's partial gradient (batch / 128 partials per
// step).  Measured at 2^14 samples: 4 steps in 0.131 ms; 16 waves of 256
// samples per workgroup (64 workgroups, 4 waves per SIMD) took 0.146 ms.
constexpr int kTrainWaves = 8;
static_assert(kTrainWaves >= 4 && kTrainWaves <= 16 && 16 % kTrainWaves == 0, "training waves per workgroup");
constexpr int kTrainThreads = 64 * kTrainWaves;
constexpr int kTrainSamples = 16 * kTrainWaves;
// Row stride (16-bit elements) of the transposed dy / x tiles in LDS: the
// workgroup's samples plus 8, so that the 4 lane groups writing rows 4 apart
// fall on different banks; rows stay 16-byte aligned for the operand reads.
constexpr int kTS = kTrainSamples + 8;
constexpr size_t kTrainLds =
    (size_t)(kFopTiles + kBopTiles) * 64 * kOpBytes16 + 2 * 64 * kTS * 2 + kTrainWaves * sizeof(float);
static_assert(kTrainLds <= 160 * 1024, "training workgroup exceeds the CU's LDS");

inline uint32_t train_blocks(uint32_t batch) { return (batch + kTrainSamples - 1) / kTrainSamples; }

struct TrainArgs {
    const float4* prep;   // this step's prepared samples, in sample order
    const uint32_t* n_rec;
    uint32_t min_records;
    uint32_t batch;
    float grad_scale;   // loss_scale / (3 * batch)
    float loss_norm;    // 1 / (3 * batch)
    NrcBounds b;
    const void* fop_train;
    const void* bop_train;
    float* partial;      // kNrcParams per workgroup
    float* block_loss;   // one per workgroup
};

template <typename T>
__device__ __forceinline__ floatx4 wgrad_tile(const T* s_dyT, const T* s_xT, int row0, int col0, int ml, int g) {
    typedef MfmaOps<T> M;
    typedef typename M::vec V;
    floatx4 a0 = floatx4{0.f, 0.f, 0.f, 0.f}, a1 = floatx4{0.f, 0.f, 0.f, 0.f};
    // two k-steps per iteration (the operands of all of them in flight at
    // once cost registers the saved activations need)
#pragma unroll 1
    for (int c = 0; c < kTrainSamples / 32; c += 2) {
        a0 = M::mfma(*reinterpret_cast<const V*>(&s_dyT[(row0 + ml) * kTS + 32 * c + 8 * g]),
                     *reinterpret_cast<const V*>(&s_xT[(col0 + ml) * kTS + 32 * c + 8 * g]), a0);
        a1 = M::mfma(*reinterpret_cast<const V*>(&s_dyT[(row0 + ml) * kTS + 32 * (c + 1) + 8 * g]),
                     *reinterpret_cast<const V*>(&s_xT[(col0 + ml) * kTS + 32 * (c + 1) + 8 * g]), a1);
    }
    return a0 + a1;
}

// A packed pair of 16-bit values into column `at` of two consecutive rows
// (`stride` elements apart) of a transposed tile: low half first.
template <typename T>
__device__ __forceinline__ void put_pair(T* tile, int at, int stride, uint32_t pair) {
    uint16_t* p = reinterpret_cast<uint16_t*>(tile);
    p[at] = (uint16_t)(pair & 0xFFFFu);
    p[at + stride] = (uint16_t)(pair >> 16);
}

template <typename T>
__global__ __launch_bounds__(kTrainThreads) void k_nrc_train(TrainArgs A) {
    typedef MfmaOps<T> M;
    typedef typename M::vec V;
    extern __shared__ __align__(16) unsigned char smem[];
    V* s_fop = reinterpret_cast<V*>(smem);
    V* s_bop = s_fop + kFopTiles * 64;
    T* s_dyT = reinterpret_cast<T*>(s_bop + kBopTiles * 64);   // [64 neurons][kTS: samples]
    T* s_xT = s_dyT + 64 * kTS;                                // [64 inputs][kTS: samples]
    float* s_loss = reinterpret_cast<float*>(s_xT + 64 * kTS);
    if (*A.n_rec <= A.min_records) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, ml = lane & 15;
    const int m = 16 * w + ml;   // sample within the workgroup
    const uint32_t i = blockIdx.x * kTrainSamples + m;
    const bool valid = i < A.batch;
    float4 smp[kPrepF4];   // in flight while the weights are staged
#pragma unroll
    for (int k = 0; k < kPrepF4; ++k) smp[k] = valid ? A.prep[(size_t)kPrepF4 * i + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    load_ops<T, kFopTiles, kTrainThreads>(s_fop, static_cast<const V*>(A.fop_train));
    load_ops<T, kBopTiles, kTrainThreads>(s_bop, static_cast<const V*>(A.bop_train));
    __syncthreads();

    V x[1][2];
    encode<T>(smp[0], smp[1], smp[2], smp[3], A.b, g, x[0][0], x[0][1]);
    V xs[kNrcLayers][2];
    floatx4 out[1];
    mlp_forward_keep<T>(s_fop, lane, x, out, xs);
    floatx4 dy5 = floatx4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.0f;
    if (g == 0 && valid) {   // RelativeL2Luminance: (p - t)^2 / (lum(p)^2 + 0.01), lum held constant
        const float tgt[3] = {smp[kVertF4].x, smp[kVertF4].y, smp[kVertF4].z};
        const float lum = kLumR * out[0][0] + kLumG * out[0][1] + kLumB * out[0][2];
        const float inv = 1.0f / (lum * lum + 0.01f);
        for (int c = 0; c < 3; ++c) {
            const float d = out[0][c] - tgt[c];
            loss += d * d * inv;
            dy5[c] = 2.0f * d * inv * A.grad_scale;
        }
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) s_loss[w] = loss;

    float* part = A.partial + (size_t)blockIdx.x * kNrcParams;
    // ---- layer 5 (16 x 64): every row of the 16 is written (rows 3..15 zero) ----
    const uint32_t d5[2] = {cvt2<T>(dy5[0], dy5[1]), cvt2<T>(dy5[2], dy5[3])};
    for (int h = 0; h < 2; ++h) {
        put_pair(s_dyT, (4 * g + 2 * h) * kTS + m, kTS, d5[h]);
    }
    for (int c = 0; c < 2; ++c)
        for (int j = 0; j < 8; ++j) s_xT[nrc_perm(c, g, j) * kTS + m] = xs[5][c][j];
    floatx4 dy[4];
    {   // input gradient of layer 5 (registers only), before the barrier
        const V bdy = __builtin_bit_cast(V, (uint4v){d5[0], d5[1], 0u, 0u});
        for (int u = 0; u < 4; ++u) {
            dy[u] = M::mfma(s_bop[(bop_tile_base(5) + u) * 64 + lane], bdy, floatx4{0.f, 0.f, 0.f, 0.f});
            for (int ii = 0; ii < 4; ++ii)
                if (!((float)xs[5][u >> 1][(u & 1) * 4 + ii] > 0.0f)) dy[u][ii] = 0.0f;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // the workgroup's loss, waves summed in order
        float L = 0.0f;
        for (int k = 0; k < kTrainWaves; ++k) L += s_loss[k];
        A.block_loss[blockIdx.x] = L * A.loss_norm;
    }
    if (w < 4) {   // 4 tiles of 16 inputs
        const floatx4 acc = wgrad_tile<T>(s_dyT, s_xT, 0, 16 * w, ml, g);
        for (int ii = 0; ii < 4; ++ii) part[nrc_layer_offset(5) + (4 * g + ii) * 64 + 16 * w + ml] = acc[ii];
    }
    __syncthreads();
    // ---- layers 4..0 (unrolled: the saved operands stay in registers) ----
#pragma unroll
    for (int l = 4; l >= 0; --l) {
        // dy in 16 bits, converted once: the transposed tile and the input
        // gradient's operands (tiles 0-1 and 2-3) read the same pairs
        uint32_t pd[4][2];
        for (int t = 0; t < 4; ++t)
            for (int h = 0; h < 2; ++h) {
                pd[t][h] = cvt2<T>(dy[t][2 * h], dy[t][2 * h + 1]);
                put_pair(s_dyT, (16 * t + 4 * g + 2 * h) * kTS + m, kTS, pd[t][h]);
            }
        for (int c = 0; c < 2; ++c)
            for (int j = 0; j < 8; ++j) s_xT[nrc_perm(c, g, j) * kTS + m] = xs[l][c][j];
        floatx4 ndy[4];
        if (l > 0) {   // input gradient of layer l (registers only), before the barrier
            const V b0 = __builtin_bit_cast(V, (uint4v){pd[0][0], pd[0][1], pd[1][0], pd[1][1]});
            const V b1 = __builtin_bit_cast(V, (uint4v){pd[2][0], pd[2][1], pd[3][0], pd[3][1]});
            for (int u = 0; u < 4; ++u) {
                floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
                acc = M::mfma(s_bop[(bop_tile_base(l) + 2 * u) * 64 + lane], b0, acc);
                acc = M::mfma(s_bop[(bop_tile_base(l) + 2 * u + 1) * 64 + lane], b1, acc);
                for (int ii = 0; ii < 4; ++ii)
                    if (!((float)xs[l][u >> 1][(u & 1) * 4 + ii] > 0.0f)) acc[ii] = 0.0f;
                ndy[u] = acc;
            }
        }
        __syncthreads();
        for (int tile = w; tile < 16; tile += kTrainWaves) {   // neurons 16 (tile / 4) .., inputs 16 (tile % 4) ..
            const int rt = tile >> 2, ct = tile & 3;
            const floatx4 acc = wgrad_tile<T>(s_dyT, s_xT, 16 * rt, 16 * ct, ml, g);
            for (int ii = 0; ii < 4; ++ii)
                part[nrc_layer_offset(l) + (16 * rt + 4 * g + ii) * 64 + 16 * ct + ml] = acc[ii];
        }
        if (l > 0)
            for (int u = 0; u < 4; ++u) dy[u] = ndy[u];
        __syncthreads();
    }
}

// Sum of the per-workgroup partial gradients; flags non-finite values or a
// skipped step.  A block covers 64 parameters with 4 waves; wave k sums the
// partials of workgroups b = k, k + 4, ... (lanes read 64 consecutive
// parameters), and the 4 sums are added in wave order through LDS, so the
// result is deterministic.  One more block sums the workgroups' losses (each
// thread a fixed strided subset, then a fixed tree) into *loss_sum.
constexpr int kReduceParams = 64;
constexpr int kReduceGrid = (kNrcParams + kReduceParams - 1) / kReduceParams + 1;
__global__ __launch_bounds__(256) void k_nrc_grad_reduce(const float* __restrict__ partial, uint32_t nblk,
                                                         float* __restrict__ grad, uint32_t* state,
                                                         const uint32_t* __restrict__ n_rec, uint32_t min_records,
                                                         float inv_scale, const float* __restrict__ block_loss,
                                                         float* __restrict__ loss_sum) {
    __shared__ float sums[4][kReduceParams];
    const int lane = (int)(threadIdx.x & 63u), slice = (int)(threadIdx.x >> 6);
    const int p = blockIdx.x * kReduceParams + lane;
    if (*n_rec <= min_records) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&state[kStFlags], 2u);
        return;
    }
    if (blockIdx.x == gridDim.x - 1) {
        float* tree = &sums[0][0];   // 256 floats
        float L = 0.0f;
        for (uint32_t b = threadIdx.x; b < nblk; b += 256) L += block_loss[b];
        tree[threadIdx.x] = L;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) tree[threadIdx.x] += tree[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss_sum = tree[0];
        return;
    }
    float s = 0.0f;
    if (p < kNrcParams) {
#pragma unroll 4
        for (uint32_t b = (uint32_t)slice; b < nblk; b += 4) s += partial[(size_t)b * kNrcParams + p];
    }
    sums[slice][lane] = s;
    __syncthreads();
    if (slice == 0 && p < kNrcParams) {
        const float t = ((sums[0][lane] + sums[1][lane]) + sums[2][lane]) + sums[3][lane];
        const float g = t * inv_scale;
        grad[p] = g;
        if (!isfinite(g)) atomicOr(&state[kStFlags], 1u);
    }
}

template <typename T>
__device__ __forceinline__ void put16(void* base, size_t pos, float v) { static_cast<T*>(base)[pos] = (T)v; }

// Adam (Kingma & Ba 2015) on the FP32 master weights, EMA of the weights for
// inference (mlp.cpp:61-69: EMA(0.99) over Adam(lr 1e-2)), then the 16-bit
// operand layouts; no parameter changes when the step was flagged.  The
// last block to finish (a done counter) closes the step: it advances the
// batch seed (mlp.cpp:102 step_counter++ on every call), counts the step as
// taken or skipped, keeps its loss (summed by the reduction) and clears the
// flags for the next step, so a step is three launches (train, reduce, Adam).
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_adam(float* __restrict__ master, float* __restrict__ am,
                                                  float* __restrict__ av, float* __restrict__ ema,
                                                  const float* __restrict__ grad, uint32_t* state,
                                                  const float* __restrict__ loss_sum, Hyper h, void* fop_train,
                                                  void* bop_train, void* fop_infer) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    // the parameter's operands are loaded together with the flags (one
    // memory latency, not two); nothing is written unless the step is taken
    const int q = p < kNrcParams ? p : 0;
    const float g = grad[q], m0 = am[q], v0 = av[q], w0 = master[q], e0 = ema[q];
    const uint32_t flags = state[kStFlags];
    const float t = (float)(state[kStLt] + 1u);
    if (p < kNrcParams && flags == 0u) {
        const float m = h.beta1 * m0 + (1.0f - h.beta1) * g;
        const float v = h.beta2 * v0 + (1.0f - h.beta2) * g * g;
        const float mh = m / (1.0f - powf(h.beta1, t));
        const float vh = v / (1.0f - powf(h.beta2, t));
        const float w = w0 - h.lr * mh / (sqrtf(vh) + h.eps);
        const float e = h.ema_decay * e0 + (1.0f - h.ema_decay) * w;
        am[p] = m;
        av[p] = v;
        master[p] = w;
        ema[p] = e;
        put16<T>(fop_train, fop_pos(p), w);
        if (param_layer(p) >= 1) put16<T>(bop_train, bop_pos(p), w);
        put16<T>(fop_infer, fop_pos(p), e);
    }
    __syncthreads();   // every read of the state in this block is done
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&state[kStDone], 1u) == gridDim.x - 1u) {
            state[kStSeed] += 1u;
            if (flags == 0u) {
                state[kStLt] += 1u;
                state[kStSteps] += 1u;
                state[kStLoss] = __float_as_uint(*loss_sum);
            } else if (flags & 1u) {
                state[kStSkipNan] += 1u;
            } else {
                state[kStSkipData] += 1u;
            }
            state[kStFlags] = 0u;
            state[kStDone] = 0u;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_nrc_publish(const float* __restrict__ master, const float* __restrict__ ema,
                                                     void* fop_train, void* bop_train, void* fop_infer) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNrcParams) return;
    put16<T>(fop_train, fop_pos(p), master[p]);
    if (param_layer(p) >= 1) put16<T>(bop_train, bop_pos(p), master[p]);
    put16<T>(fop_infer, fop_pos(p), ema[p]);
}

// Prepared samples of `steps` steps from records, on stream s (see k_nrc_prepare).
template <typename T>
int enqueue_prepare(tfg_nrc* nc, const float4* records, const uint32_t* n_rec, uint32_t steps, uint32_t min_records,
                    float4* out, uint32_t* out_n, hipStream_t s) {
    const PrepArgs P{records, n_rec, min_records, nc->state, nc->cfg.batch, steps, nc->bounds, nc->fop_infer, out,
                     out_n};
    const uint64_t total = (uint64_t)nc->cfg.batch * steps;
    // one workgroup per CU: 65,536 samples are 2 groups of 16 per wave
    const uint32_t grid = (uint32_t)std::min<uint64_t>((total + 127) / 128, (uint64_t)nc->n_cu);
    hipLaunchKernelGGL(k_nrc_prepare<T>, dim3(grid > 0 ? grid : 1), dim3(kPrepBlock), 0, s, P);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

// `steps` optimiser steps on prepared samples (steps x batch of them, step k's
// at k * batch), three launches each.
template <typename T>
int enqueue_steps(tfg_nrc* nc, const float4* prep, const uint32_t* n_rec, uint32_t steps, uint32_t min_records,
                  hipStream_t s) {
    const uint32_t nblk = train_blocks(nc->cfg.batch);
    TrainArgs A;
    A.prep = prep;
    A.n_rec = n_rec;
    A.min_records = min_records;
    A.batch = nc->cfg.batch;
    A.grad_scale = nc->cfg.loss_scale / (3.0f * (float)nc->cfg.batch);
    A.loss_norm = 1.0f / (3.0f * (float)nc->cfg.batch);
    A.b = nc->bounds;
    A.fop_train = nc->fop_train;
    A.bop_train = nc->bop_train;
    A.partial = nc->partial;
    A.block_loss = nc->block_loss;
    const Hyper h{nc->cfg.lr, nc->cfg.beta1, nc->cfg.beta2, nc->cfg.eps, nc->cfg.ema_decay, 1.0f / nc->cfg.loss_scale};
    const int pblocks = (kNrcParams + 255) / 256;
    for (uint32_t k = 0; k < steps; ++k) {
        A.prep = prep + (size_t)k * nc->cfg.batch * kPrepF4;
        hipLaunchKernelGGL(k_nrc_train<T>, dim3(nblk), dim3(kTrainThreads), kTrainLds, s, A);
        hipLaunchKernelGGL(k_nrc_grad_reduce, dim3(kReduceGrid), dim3(256), 0, s, nc->partial, nblk, nc->grad,
                           nc->state, n_rec, min_records, h.inv_loss_scale, nc->block_loss, nc->loss_sum);
        hipLaunchKernelGGL(k_nrc_adam<T>, dim3(pblocks), dim3(256), 0, s, nc->master, nc->adam_m, nc->adam_v, nc->ema,
                           nc->grad, nc->state, nc->loss_sum, h, nc->fop_train, nc->bop_train, nc->fop_infer);
    }
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

// One call's steps on training records: their samples and targets are
// prepared first, with the inference weights of the call's start, then the
// steps run on them.  The caller has reserved prep for steps x batch samples.
template <typename T>
int enqueue_train_steps(tfg_nrc* nc, const float4* records, const uint32_t* n_rec, uint32_t steps, hipStream_t s,
                        uint32_t min_records) {
    if (steps == 0) return TFG_OK;
    if ((size_t)steps * nc->cfg.batch > nc->prep_cap) {
        set_error("tfg_nrc: %u steps of %u samples exceed the prepared-sample buffer (%zu)", steps, nc->cfg.batch,
                  nc->prep_cap);
        return TFG_E_LIMIT;
    }
    const int rc = enqueue_prepare<T>(nc, records, n_rec, steps, min_records, nc->prep, nullptr, s);
    if (rc != TFG_OK) return rc;
    return enqueue_steps<T>(nc, nc->prep, n_rec, steps, min_records, s);
}

template <typename T>
int enqueue_infer(tfg_nrc* nc, const float4* queries, const uint32_t* count, tfg_color* image, hipStream_t s) {
    typedef typename MfmaOps<T>::vec V;
    hipLaunchKernelGGL((k_nrc_infer<T, kInferGroups>), dim3(nc->n_cu * 4), dim3(kInferBlock), 0, s, queries, count, image,
                       static_cast<const V*>(nc->fop_infer), nc->bounds);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

template <typename T>
int setup_lds() {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nrc_train<T>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrainLds);
    return e == hipSuccess ? TFG_OK : hip_fail(e, "hipFuncSetAttribute(k_nrc_train)");
}

}  // namespace

// DatosMLP (mlp_types.h:14-23) -> vertex record (roughness 0) or training
// record (L_suffix = colour, T_suffix = 0, normaliser 1: target = colour).
__global__ void k_from_datos(const tfg_datos_mlp* __restrict__ d, uint32_t n, float4* out, int as_record) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_datos_mlp x = d[i];
        float4* o = out + (size_t)(as_record ? kRecF4 : kVertF4) * i;
        o[0] = make_float4((float)x.posicion.x, (float)x.posicion.y, (float)x.posicion.z, 0.0f);
        o[1] = make_float4((float)x.direccion.x, (float)x.direccion.y, (float)x.direccion.z, x.difuso.r);
        o[2] = make_float4((float)x.normal.x, (float)x.normal.y, (float)x.normal.z, x.difuso.g);
        o[3] = make_float4(x.difuso.b, x.especular.r, x.especular.g, x.especular.b);
        if (as_record) {
            for (int k = 4; k < 8; ++k) o[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            o[8] = make_float4(x.color.r, x.color.g, x.color.b, 1.0f);
            o[9] = make_float4(0.f, 0.f, 0.f, 0.f);
            o[10] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        }
    }
}

__global__ void k_to_color(const float* __restrict__ raw, uint32_t n, tfg_color* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = tfg_color{sanitize_out(raw[3 * (size_t)i]), sanitize_out(raw[3 * (size_t)i + 1]),
                           sanitize_out(raw[3 * (size_t)i + 2])};
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

int dm_reserve(tfg_nrc* nc, size_t n) {
    if (n <= nc->dm_cap) return TFG_OK;
    if (nc->dm_buf) (void)hipFree(nc->dm_buf);
    if (nc->dm_out) (void)hipFree(nc->dm_out);
    nc->dm_buf = nullptr;
    nc->dm_out = nullptr;
    nc->dm_cap = 0;
    TFG_HIP_CHECK(hipMalloc((void**)&nc->dm_buf, n * kRecF4 * sizeof(float4)));
    TFG_HIP_CHECK(hipMalloc((void**)&nc->dm_out, n * 3 * sizeof(float)));
    if (!nc->dm_count) TFG_HIP_CHECK(hipMalloc((void**)&nc->dm_count, sizeof(uint32_t)));
    nc->dm_cap = n;
    return TFG_OK;
}

// Grows the prepared-sample buffer to `samples` (a tfg_nrc_train call with
// more steps than steps_per_frame); work already enqueued may still read the
// old buffer, so the device is idle first.
int prep_reserve(tfg_nrc* nc, size_t samples) {
    if (samples <= nc->prep_cap) return TFG_OK;
    TFG_HIP_CHECK(hipDeviceSynchronize());
    if (nc->prep) (void)hipFree(nc->prep);
    nc->prep = nullptr;
    nc->prep_cap = 0;
    TFG_HIP_CHECK(hipMalloc((void**)&nc->prep, samples * kPrepF4 * sizeof(float4)));
    nc->prep_cap = samples;
    return TFG_OK;
}

size_t nrc_gather_f4(const tfg_nrc* nc) {
    return (size_t)nc->cfg.batch * nc->cfg.steps_per_frame * kPrepF4;
}

int nrc_gather_enqueue(tfg_nrc* nc, const tfg_nrc_frame& f, float4* out, uint32_t* out_n, hipStream_t s) {
    return nc->cfg.precision == TFG_NRC_FP16
               ? enqueue_prepare<_Float16>(nc, f.records, f.counters + CNT_RECORDS, nc->cfg.steps_per_frame,
                                           nc->cfg.min_records, out, out_n, s)
               : enqueue_prepare<__bf16>(nc, f.records, f.counters + CNT_RECORDS, nc->cfg.steps_per_frame,
                                         nc->cfg.min_records, out, out_n, s);
}

int nrc_train_gathered_enqueue(tfg_nrc* nc, const float4* gathered, const uint32_t* n_rec, hipStream_t s) {
    if (nc->cfg.steps_per_frame == 0) return TFG_OK;
    return nc->cfg.precision == TFG_NRC_FP16
               ? enqueue_steps<_Float16>(nc, gathered, n_rec, nc->cfg.steps_per_frame, nc->cfg.min_records, s)
               : enqueue_steps<__bf16>(nc, gathered, n_rec, nc->cfg.steps_per_frame, nc->cfg.min_records, s);
}

int nrc_frame_enqueue(tfg_nrc* nc, const tfg_nrc_frame& f, tfg_color* image, hipStream_t s) {
    const bool fp16 = nc->cfg.precision == TFG_NRC_FP16;
    int rc = TFG_OK;
    if (f.mode & MODE_INFER) {
        rc = fp16 ? enqueue_infer<_Float16>(nc, f.queries, f.counters + CNT_QUERIES, image, s)
                  : enqueue_infer<__bf16>(nc, f.queries, f.counters + CNT_QUERIES, image, s);
        if (rc != TFG_OK) return rc;
    }
    if ((f.mode & MODE_TRAIN) && nc->cfg.steps_per_frame > 0) {
        rc = fp16 ? enqueue_train_steps<_Float16>(nc, f.records, f.counters + CNT_RECORDS, nc->cfg.steps_per_frame, s,
                                                  nc->cfg.min_records)
                  : enqueue_train_steps<__bf16>(nc, f.records, f.counters + CNT_RECORDS, nc->cfg.steps_per_frame, s,
                                                nc->cfg.min_records);
    }
    return rc;
}

}  // namespace tfg

namespace {
std::mutex g_scratch_mu;
void* g_scratch[64] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int scratch_ptr(float4** out) {
    int dev = -1;
    TFG_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("tfg_nrc_infer_datos_noalloc: device %d out of range", dev);
        return TFG_E_LIMIT;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = p;
    }
    *out = static_cast<float4*>(g_scratch[dev]);
    return TFG_OK;
}
}  // namespace

// ===================================================================== C ABI
extern "C" {

int tfg_nrc_default_config(tfg_nrc_config* c) {
    if (!c) { set_error("tfg_nrc_default_config: NULL"); return TFG_E_INVALID; }
    std::memset(c, 0, sizeof(*c));
    c->batch = 16384;            // main.cpp:939
    c->steps_per_frame = 4;      // 4 x 2^14 = 2^16 training samples per frame
    c->min_records = 1024;       // main.cpp:1043
    c->lr = 1e-2f;               // mlp.cpp:66
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->eps = 1e-8f;
    c->ema_decay = 0.99f;        // mlp.cpp:62
    c->loss_scale = 128.0f;
    c->bounds_min[0] = -6.0f; c->bounds_min[1] = -4.0f; c->bounds_min[2] = -10.0f;   // main.cpp:870-872
    c->bounds_max[0] = 6.0f;  c->bounds_max[1] = 6.0f;  c->bounds_max[2] = 2.0f;
    c->seed = 1337u;
    c->precision = TFG_NRC_BF16;
    return TFG_OK;
}

int tfg_nrc_create(const tfg_nrc_config* cfg, tfg_nrc** out) {
    if (!cfg || !out) { set_error("tfg_nrc_create: NULL argument"); return TFG_E_INVALID; }
    *out = nullptr;
    if (cfg->batch == 0 || cfg->batch % 64 != 0 || cfg->batch > (1u << 22)) {
        set_error("tfg_nrc_create: batch must be a positive multiple of 64 (<= 2^22)");
        return TFG_E_INVALID;
    }
    if (cfg->precision != TFG_NRC_BF16 && cfg->precision != TFG_NRC_FP16) {
        set_error("tfg_nrc_create: unknown precision %u", cfg->precision);
        return TFG_E_INVALID;
    }
    if (!(cfg->loss_scale > 0.0f) || !(cfg->lr > 0.0f)) { set_error("tfg_nrc_create: bad hyper-parameters"); return TFG_E_INVALID; }
    for (int a = 0; a < 3; ++a)
        if (!(cfg->bounds_max[a] > cfg->bounds_min[a])) { set_error("tfg_nrc_create: empty scene bounds"); return TFG_E_INVALID; }
    auto* nc = new (std::nothrow) tfg_nrc();
    if (!nc) { set_error("tfg_nrc_create: out of memory"); return TFG_E_NOMEM; }
    nc->cfg = *cfg;
    for (int a = 0; a < 3; ++a) {
        nc->bounds.mn[a] = cfg->bounds_min[a];
        nc->bounds.inv_ext[a] = 1.0f / (cfg->bounds_max[a] - cfg->bounds_min[a]);
    }
    const size_t pb = kNrcParams * sizeof(float);
    hipError_t e = hipSuccess;
    void** bufs[] = {(void**)&nc->master, (void**)&nc->adam_m, (void**)&nc->adam_v, (void**)&nc->ema, (void**)&nc->grad};
    for (void** b : bufs)
        if (e == hipSuccess) e = hipMalloc(b, pb);
    const size_t nblk = train_blocks(cfg->batch);
    if (e == hipSuccess) e = hipMalloc((void**)&nc->partial, nblk * pb);
    if (e == hipSuccess) e = hipMalloc((void**)&nc->block_loss, nblk * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&nc->loss_sum, sizeof(float));
    nc->prep_cap = (size_t)cfg->batch * std::max<uint32_t>(cfg->steps_per_frame, 1u);
    if (e == hipSuccess) e = hipMalloc((void**)&nc->prep, nc->prep_cap * kPrepF4 * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&nc->fop_train, (size_t)kFopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc(&nc->bop_train, (size_t)kBopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc(&nc->fop_infer, (size_t)kFopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc((void**)&nc->state, 16 * sizeof(uint32_t));
    int dev = 0;
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
        tfg_nrc_destroy(nc);
        return hip_fail(e, "tfg_nrc_create");
    }
    nc->n_cu = prop.multiProcessorCount;
    int rc = cfg->precision == TFG_NRC_FP16 ? setup_lds<_Float16>() : setup_lds<__bf16>();
    if (rc != TFG_OK) {
        tfg_nrc_destroy(nc);
        return rc;
    }
    // Xavier-uniform initialisation (seeded); output rows 3..15 are padding.
    std::vector<float> w(kNrcParams, 0.0f);
    std::mt19937 gen(cfg->seed);
    for (int l = 0; l < kNrcLayers; ++l) {
        const int rows = l < 5 ? 64 : kNrcOut;
        const float a = std::sqrt(6.0f / (64.0f + (float)(l < 5 ? 64 : kNrcOut)));
        std::uniform_real_distribution<float> U(-a, a);
        for (int n = 0; n < rows; ++n)
            for (int k = 0; k < 64; ++k) w[nrc_layer_offset(l) + n * 64 + k] = U(gen);
    }
    rc = tfg_nrc_set_weights(nc, w.data(), w.data());
    if (rc != TFG_OK) {
        tfg_nrc_destroy(nc);
        return rc;
    }
    *out = nc;
    return TFG_OK;
}

int tfg_nrc_set_weights(tfg_nrc* nc, const float* master, const float* ema) {
    if (!nc || !master) { set_error("tfg_nrc_set_weights: NULL argument"); return TFG_E_INVALID; }
    const size_t pb = kNrcParams * sizeof(float);
    TFG_HIP_CHECK(hipMemcpy(nc->master, master, pb, hipMemcpyHostToDevice));
    TFG_HIP_CHECK(hipMemcpy(nc->ema, ema ? ema : master, pb, hipMemcpyHostToDevice));
    TFG_HIP_CHECK(hipMemset(nc->adam_m, 0, pb));
    TFG_HIP_CHECK(hipMemset(nc->adam_v, 0, pb));
    TFG_HIP_CHECK(hipMemset(nc->fop_train, 0, (size_t)kFopTiles * 64 * kOpBytes16));
    TFG_HIP_CHECK(hipMemset(nc->bop_train, 0, (size_t)kBopTiles * 64 * kOpBytes16));
    TFG_HIP_CHECK(hipMemset(nc->fop_infer, 0, (size_t)kFopTiles * 64 * kOpBytes16));
    TFG_HIP_CHECK(hipMemset(nc->state, 0, 16 * sizeof(uint32_t)));
    const int pblocks = (kNrcParams + 255) / 256;
    if (nc->cfg.precision == TFG_NRC_FP16)
        hipLaunchKernelGGL(k_nrc_publish<_Float16>, dim3(pblocks), dim3(256), 0, nullptr, nc->master, nc->ema,
                           nc->fop_train, nc->bop_train, nc->fop_infer);
    else
        hipLaunchKernelGGL(k_nrc_publish<__bf16>, dim3(pblocks), dim3(256), 0, nullptr, nc->master, nc->ema,
                           nc->fop_train, nc->bop_train, nc->fop_infer);
    TFG_HIP_CHECK(hipGetLastError());
    TFG_HIP_CHECK(hipDeviceSynchronize());
    return TFG_OK;
}

int tfg_nrc_get_weights(tfg_nrc* nc, float* master, float* ema) {
    if (!nc) { set_error("tfg_nrc_get_weights: NULL"); return TFG_E_INVALID; }
    TFG_HIP_CHECK(hipDeviceSynchronize());
    const size_t pb = kNrcParams * sizeof(float);
    if (master) TFG_HIP_CHECK(hipMemcpy(master, nc->master, pb, hipMemcpyDeviceToHost));
    if (ema) TFG_HIP_CHECK(hipMemcpy(ema, nc->ema, pb, hipMemcpyDeviceToHost));
    return TFG_OK;
}

int tfg_nrc_get_stats(tfg_nrc* nc, tfg_nrc_stats* st) {
    if (!nc || !st) { set_error("tfg_nrc_get_stats: NULL argument"); return TFG_E_INVALID; }
    uint32_t s[16];
    TFG_HIP_CHECK(hipDeviceSynchronize());
    TFG_HIP_CHECK(hipMemcpy(s, nc->state, sizeof(s), hipMemcpyDeviceToHost));
    st->steps = s[kStSteps];
    st->skipped_nonfinite = s[kStSkipNan];
    st->skipped_no_data = s[kStSkipData];
    std::memcpy(&st->last_loss, &s[kStLoss], sizeof(float));
    st->seed_counter = s[kStSeed];
    return TFG_OK;
}

int tfg_nrc_eval(tfg_nrc* nc, const float* d_verts, uint32_t n, float* d_out, float* d_dump, int use_train_weights,
                 tfg_stream stream) {
    if (!nc || !d_verts || !d_out) { set_error("tfg_nrc_eval: NULL argument"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    const void* fop = use_train_weights ? nc->fop_train : nc->fop_infer;
    const dim3 grid((n + 63) / 64);
    if (nc->cfg.precision == TFG_NRC_FP16)
        hipLaunchKernelGGL(k_nrc_eval<_Float16>, grid, dim3(256), 0, stream, reinterpret_cast<const float4*>(d_verts),
                           n, d_out, d_dump, static_cast<const f16x8*>(fop), nc->bounds);
    else
        hipLaunchKernelGGL(k_nrc_eval<__bf16>, grid, dim3(256), 0, stream, reinterpret_cast<const float4*>(d_verts), n,
                           d_out, d_dump, static_cast<const bf16x8*>(fop), nc->bounds);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

int tfg_nrc_infer_queries(tfg_nrc* nc, const float* d_queries, const uint32_t* d_count, tfg_color* d_image,
                          tfg_stream stream) {
    if (!nc || !d_queries || !d_count || !d_image) { set_error("tfg_nrc_infer_queries: NULL argument"); return TFG_E_INVALID; }
    const float4* q = reinterpret_cast<const float4*>(d_queries);
    return nc->cfg.precision == TFG_NRC_FP16 ? enqueue_infer<_Float16>(nc, q, d_count, d_image, stream)
                                             : enqueue_infer<__bf16>(nc, q, d_count, d_image, stream);
}

int tfg_nrc_train(tfg_nrc* nc, const float* d_records, const uint32_t* d_n_records, uint32_t steps, tfg_stream stream) {
    if (!nc || !d_records || !d_n_records) { set_error("tfg_nrc_train: NULL argument"); return TFG_E_INVALID; }
    if ((uint64_t)steps * nc->cfg.batch > (1ull << 27)) {
        set_error("tfg_nrc_train: %u steps of %u samples: at most 2^27 samples per call", steps, nc->cfg.batch);
        return TFG_E_LIMIT;
    }
    const int rc = prep_reserve(nc, (size_t)steps * nc->cfg.batch);
    if (rc != TFG_OK) return rc;
    const float4* r = reinterpret_cast<const float4*>(d_records);
    return nc->cfg.precision == TFG_NRC_FP16
               ? enqueue_train_steps<_Float16>(nc, r, d_n_records, steps, stream, nc->cfg.min_records)
               : enqueue_train_steps<__bf16>(nc, r, d_n_records, steps, stream, nc->cfg.min_records);
}

int tfg_nrc_last_gradient(tfg_nrc* nc, float* host_grad) {
    if (!nc || !host_grad) { set_error("tfg_nrc_last_gradient: NULL argument"); return TFG_E_INVALID; }
    TFG_HIP_CHECK(hipDeviceSynchronize());
    TFG_HIP_CHECK(hipMemcpy(host_grad, nc->grad, kNrcParams * sizeof(float), hipMemcpyDeviceToHost));
    return TFG_OK;
}

// ColorMLP::inference (mlp.cpp:127-170): DatosMLP -> sanitised Color, synchronous.
int tfg_nrc_infer_datos(tfg_nrc* nc, const tfg_datos_mlp* d_in, uint32_t n, tfg_color* d_out, tfg_stream stream) {
    if (!nc || !d_in || !d_out) { set_error("tfg_nrc_infer_datos: NULL argument"); return TFG_E_INVALID; }
    if (n == 0) return TFG_OK;
    int rc = dm_reserve(nc, n);
    if (rc != TFG_OK) return rc;
    hipLaunchKernelGGL(k_from_datos, dim3(1024), dim3(256), 0, stream, d_in, n, nc->dm_buf, 0);
    rc = tfg_nrc_eval(nc, reinterpret_cast<const float*>(nc->dm_buf), n, nc->dm_out, nullptr, 0, stream);
    if (rc != TFG_OK) return rc;
    hipLaunchKernelGGL(k_to_color, dim3(1024), dim3(256), 0, stream, nc->dm_out, n, d_out);
    TFG_HIP_CHECK(hipGetLastError());
    TFG_HIP_CHECK(hipStreamSynchronize(stream));
    return TFG_OK;
}


// tfg_nrc_infer_datos without allocation: the converted inputs and raw
// outputs live in the library's scratch array instead of buffers grown with
// hipMalloc.  Same kernels, same results; asynchronous on `stream`.
int tfg_nrc_infer_datos_noalloc(tfg_nrc* nc, const tfg_datos_mlp* d_in, uint32_t n, tfg_color* d_out,
                                tfg_stream stream) {
    if (!nc || !d_in || !d_out) { set_error("tfg_nrc_infer_datos_noalloc: NULL argument"); return TFG_E_INVALID; }
    if (n > kNoAllocMaxQueries) {
        set_error("tfg_nrc_infer_datos_noalloc: %u queries exceed the scratch capacity of %u per call", n,
                  kNoAllocMaxQueries);
        return TFG_E_LIMIT;
    }
    if (n == 0) return TFG_OK;
    float4* base = nullptr;
    const int rc = scratch_ptr(&base);
    if (rc != TFG_OK) return rc;
    float4* verts = base;
    float* raw = reinterpret_cast<float*>(base + kScratchVertF4);
    // every entry read below is written first: nothing needs clearing
    const uint32_t grid = std::min<uint32_t>((n + 255) / 256, 1024u);
    hipLaunchKernelGGL(k_from_datos, dim3(grid), dim3(256), 0, stream, d_in, n, verts, 0);
    const void* fop = nc->fop_infer;
    const dim3 egrid((n + 63) / 64);
    if (nc->cfg.precision == TFG_NRC_FP16)
        hipLaunchKernelGGL(k_nrc_eval<_Float16>, egrid, dim3(256), 0, stream, verts, n, raw, nullptr,
                           static_cast<const f16x8*>(fop), nc->bounds);
    else
        hipLaunchKernelGGL(k_nrc_eval<__bf16>, egrid, dim3(256), 0, stream, verts, n, raw, nullptr,
                           static_cast<const bf16x8*>(fop), nc->bounds);
    hipLaunchKernelGGL(k_to_color, dim3(grid), dim3(256), 0, stream, raw, n, d_out);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

// ColorMLP::train_step (mlp.cpp:95-125): one optimiser step on `batch`
// samples drawn from the first n entries; returns 0 loss and does nothing
// when n < 100 (mlp.cpp:96).  Synchronous.
int tfg_nrc_train_datos(tfg_nrc* nc, const tfg_datos_mlp* d_in, uint32_t n, float* loss, tfg_stream stream) {
    if (!nc || !d_in) { set_error("tfg_nrc_train_datos: NULL argument"); return TFG_E_INVALID; }
    if (loss) *loss = 0.0f;
    if (n < 100) return TFG_OK;
    int rc = dm_reserve(nc, n);
    if (rc != TFG_OK) return rc;
    hipLaunchKernelGGL(k_from_datos, dim3(1024), dim3(256), 0, stream, d_in, n, nc->dm_buf, 1);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, stream, nc->dm_count, n);
    TFG_HIP_CHECK(hipGetLastError());
    rc = nc->cfg.precision == TFG_NRC_FP16
             ? enqueue_train_steps<_Float16>(nc, nc->dm_buf, nc->dm_count, 1, stream, 99u)
             : enqueue_train_steps<__bf16>(nc, nc->dm_buf, nc->dm_count, 1, stream, 99u);
    if (rc != TFG_OK) return rc;
    TFG_HIP_CHECK(hipStreamSynchronize(stream));
    if (loss) {
        uint32_t s[16];
        TFG_HIP_CHECK(hipMemcpy(s, nc->state, sizeof(s), hipMemcpyDeviceToHost));
        std::memcpy(loss, &s[kStLoss], sizeof(float));
    }
    return TFG_OK;
}

int tfg_nrc_destroy(tfg_nrc* nc) {
    if (!nc) return TFG_OK;
    (void)hipDeviceSynchronize();
    void* bufs[] = {nc->master, nc->adam_m, nc->adam_v, nc->ema, nc->grad, nc->partial, nc->block_loss, nc->loss_sum, nc->prep,
                    nc->fop_train, nc->bop_train, nc->fop_infer, nc->state, nc->dm_buf, nc->dm_out, nc->dm_count};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete nc;
    return TFG_OK;
}

}  // extern "C"
