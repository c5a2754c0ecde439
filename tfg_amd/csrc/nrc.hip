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
    float* partial = nullptr;   // (batch / 64) x kNrcParams
    // operand layouts, 16-bit (bf16 or fp16 by cfg.precision)
    void* fop_train = nullptr;
    void* bop_train = nullptr;
    void* fop_infer = nullptr;
    // the inference (EMA) operands as they were when a call's optimiser
    // steps were enqueued: every step of the call forms its targets with
    // them (the reference computes a frame's targets once, before its step)
    void* fop_tail = nullptr;
    // device scalars: [0] step count t, [1] seed counter, [2] flags of the
    // current step, [3] skipped (non-finite) steps, [4] no-data steps,
    // [5] loss of the last step (float bits), [6] steps taken
    uint32_t* state = nullptr;
    float* loss_accum = nullptr;
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
template <typename T, int kTiles, int kThreads>
__device__ __forceinline__ void load_ops(typename MfmaOps<T>::vec* dst, const typename MfmaOps<T>::vec* src) {
    typedef typename MfmaOps<T>::vec V;
    constexpr int kN = kTiles * 64, kPer = (kN + kThreads - 1) / kThreads;
    V tmp[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + k * kThreads;
        if (i < kN) tmp[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + k * kThreads;
        if (i < kN) dst[i] = tmp[k];
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
struct TrainArgs {
    const float4* records;   // all records, or (gathered) this step's batch in sample order
    int gathered;
    const uint32_t* n_rec;
    uint32_t min_records;
    uint32_t* state;
    uint32_t batch;
    float grad_scale;   // loss_scale / (3 * batch)
    float loss_norm;    // 1 / (3 * batch)
    NrcBounds b;
    const void* fop_train;
    const void* bop_train;
    const void* fop_tail;
    float* partial;
    float* loss_accum;
};

// Row stride (16-bit elements) of the transposed dy / x tiles in LDS: 72,
// not 64, so that the 4 lane groups writing rows 4 apart hit different banks
// (a 128-byte stride put them on one bank); rows stay 16-byte aligned for the
// operand reads.
constexpr int kTS = 72;
constexpr size_t kTrainLds = (size_t)(2 * kFopTiles + kBopTiles) * 64 * kOpBytes16 + 2 * 64 * kTS * 2;

// One block = 64 training samples (4 waves x 16).  Tail inference with the
// inference (EMA) weights of the call's start (kernelPrepararInferenciaTail
// + ColorMLP::inference, RenderKernel.cu:139-166, main.cpp:1047-1050),
// targets (RenderKernel.cu:92-137), head forward
// with the training weights, RelativeL2Luminance gradient, backward through
// all layers and this block's weight gradient (summed over its samples).
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_train(TrainArgs A) {
    typedef MfmaOps<T> M;
    typedef typename M::vec V;
    extern __shared__ __align__(16) unsigned char smem[];
    V* s_fop = reinterpret_cast<V*>(smem);
    V* s_inf = s_fop + kFopTiles * 64;
    V* s_bop = s_inf + kFopTiles * 64;
    T* s_dyT = reinterpret_cast<T*>(s_bop + kBopTiles * 64);   // [64 neurons][kTS: 64 samples]
    T* s_xT = s_dyT + 64 * kTS;                                // [64 inputs][kTS: 64 samples]
    const uint32_t n_rec = *A.n_rec;
    if (n_rec <= A.min_records) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, ml = lane & 15;
    const int m = 16 * w + ml;   // sample within the block
    const uint32_t i = blockIdx.x * 64 + m;
    // generarIndicesAleatorios (NrcKernel.cu:10-16): seed = step counter
    const uint32_t idx = A.gathered ? i : pcg32_step(A.state[kStSeed] ^ i) % n_rec;
    const float4* R = A.records + (size_t)kRecF4 * idx;
    float4 rec[kRecF4];   // the sample's record, in flight while the weights are staged
#pragma unroll
    for (int k = 0; k < kRecF4; ++k) rec[k] = R[k];
    load_ops<T, kFopTiles, 256>(s_fop, static_cast<const V*>(A.fop_train));
    load_ops<T, kFopTiles, 256>(s_inf, static_cast<const V*>(A.fop_tail));
    load_ops<T, kBopTiles, 256>(s_bop, static_cast<const V*>(A.bop_train));
    __syncthreads();

    // tail inference (EMA weights) and head forward (training weights, layer
    // inputs kept for the backward pass), one layer of each at a time: the
    // two chains are independent and fill each other's latencies
    V xt[1][2], x[1][2];
    encode<T>(rec[4], rec[5], rec[6], rec[7], A.b, g, xt[0][0], xt[0][1]);
    encode<T>(rec[0], rec[1], rec[2], rec[3], A.b, g, x[0][0], x[0][1]);
    V xs[kNrcLayers][2];
#pragma unroll
    for (int l = 0; l < kNrcHidden; ++l) {
        xs[l][0] = x[0][0];
        xs[l][1] = x[0][1];
        mlp_layer<T, 1>(s_inf, lane, l, xt);
        mlp_layer<T, 1>(s_fop, lane, l, x);
    }
    xs[5][0] = x[0][0];
    xs[5][1] = x[0][1];
    floatx4 out_t[1], out[1];
    mlp_output<T, 1>(s_inf, lane, xt, out_t);
    mlp_output<T, 1>(s_fop, lane, x, out);
    float tgt[3] = {0.f, 0.f, 0.f};
    if (g == 0) {
        const float4 ls = rec[8], ts = rec[9], nm = rec[10];
        const float lsv[3] = {ls.x, ls.y, ls.z}, tsv[3] = {ts.x, ts.y, ts.z}, nmv[3] = {nm.x, nm.y, nm.z};
        for (int c = 0; c < 3; ++c) {
            float t = (lsv[c] + tsv[c] * sanitize_out(out_t[0][c])) / fmaxf(nmv[c], 1e-8f);
            t = fminf(t, 100000.0f);
            if (isnan(t) || isinf(t)) t = 0.0f;
            tgt[c] = t;
        }
    }
    floatx4 dy5 = floatx4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.0f;
    if (g == 0) {   // RelativeL2Luminance: (p - t)^2 / (lum(p)^2 + 0.01), lum held constant
        const float lum = kLumR * out[0][0] + kLumG * out[0][1] + kLumB * out[0][2];
        const float inv = 1.0f / (lum * lum + 0.01f);
        for (int c = 0; c < 3; ++c) {
            const float d = out[0][c] - tgt[c];
            loss += d * d * inv;
            dy5[c] = 2.0f * d * inv * A.grad_scale;
        }
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) atomicAdd(A.loss_accum, loss * A.loss_norm);

    float* part = A.partial + (size_t)blockIdx.x * kNrcParams;
    // ---- layer 5 (16 x 64): every row of the 16 is written (rows 3..15 zero) ----
    for (int ii = 0; ii < 4; ++ii) s_dyT[(4 * g + ii) * kTS + m] = (T)dy5[ii];
    for (int c = 0; c < 2; ++c)
        for (int j = 0; j < 8; ++j) s_xT[nrc_perm(c, g, j) * kTS + m] = xs[5][c][j];
    floatx4 dy[4];
    {   // input gradient of layer 5 (registers only), before the barrier
        V bdy;
        for (int j = 0; j < 4; ++j) { bdy[j] = (T)dy5[j]; bdy[4 + j] = (T)0.0f; }
        for (int u = 0; u < 4; ++u) {
            dy[u] = M::mfma(s_bop[(bop_tile_base(5) + u) * 64 + lane], bdy, floatx4{0.f, 0.f, 0.f, 0.f});
            for (int ii = 0; ii < 4; ++ii)
                if (!((float)xs[5][u >> 1][(u & 1) * 4 + ii] > 0.0f)) dy[u][ii] = 0.0f;
        }
    }
    __syncthreads();
    {
        const int u = w;   // k-tile of this wave
        floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < 2; ++c) {
            const V a = *reinterpret_cast<const V*>(&s_dyT[ml * kTS + 32 * c + 8 * g]);
            const V bb = *reinterpret_cast<const V*>(&s_xT[(16 * u + ml) * kTS + 32 * c + 8 * g]);
            acc = M::mfma(a, bb, acc);
        }
        for (int ii = 0; ii < 4; ++ii) part[nrc_layer_offset(5) + (4 * g + ii) * 64 + 16 * u + ml] = acc[ii];
    }
    __syncthreads();
    // ---- layers 4..0 (unrolled: the saved operands stay in registers) ----
#pragma unroll
    for (int l = 4; l >= 0; --l) {
        for (int t = 0; t < 4; ++t)
            for (int ii = 0; ii < 4; ++ii) s_dyT[(16 * t + 4 * g + ii) * kTS + m] = (T)dy[t][ii];
        for (int c = 0; c < 2; ++c)
            for (int j = 0; j < 8; ++j) s_xT[nrc_perm(c, g, j) * kTS + m] = xs[l][c][j];
        floatx4 ndy[4];
        if (l > 0) {   // input gradient of layer l (registers only), before the barrier
            V b0, b1;
            for (int ii = 0; ii < 4; ++ii) {
                b0[ii] = (T)dy[0][ii];
                b0[4 + ii] = (T)dy[1][ii];
                b1[ii] = (T)dy[2][ii];
                b1[4 + ii] = (T)dy[3][ii];
            }
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
        for (int u = 0; u < 4; ++u) {
            floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 2; ++c) {
                const V a = *reinterpret_cast<const V*>(&s_dyT[(16 * w + ml) * kTS + 32 * c + 8 * g]);
                const V bb = *reinterpret_cast<const V*>(&s_xT[(16 * u + ml) * kTS + 32 * c + 8 * g]);
                acc = M::mfma(a, bb, acc);
            }
            for (int ii = 0; ii < 4; ++ii) part[nrc_layer_offset(l) + (16 * w + 4 * g + ii) * 64 + 16 * u + ml] = acc[ii];
        }
        if (l > 0)
            for (int u = 0; u < 4; ++u) dy[u] = ndy[u];
        __syncthreads();
    }
}

// Sums the per-block gradients; flags non-finite values or a skipped step.
// Sum of the per-block partial gradients.  A block covers 64 parameters
// with 4 waves; wave k sums the partials of blocks b = k, k + 4, ... (lanes
// read 64 consecutive parameters), and the 4 sums are added in wave order
// through LDS, so the result is deterministic.
constexpr int kReduceParams = 64;
__global__ __launch_bounds__(256) void k_nrc_grad_reduce(const float* __restrict__ partial, uint32_t nblk,
                                                         float* __restrict__ grad, uint32_t* state,
                                                         const uint32_t* __restrict__ n_rec, uint32_t min_records,
                                                         float inv_scale) {
    __shared__ float sums[4][kReduceParams];
    const int lane = (int)(threadIdx.x & 63u), slice = (int)(threadIdx.x >> 6);
    const int p = blockIdx.x * kReduceParams + lane;
    if (*n_rec <= min_records) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&state[kStFlags], 2u);
        return;
    }
    float s = 0.0f;
    if (p < kNrcParams) {
#pragma unroll 8
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
// taken or skipped, keeps its loss, and clears the flags and the loss sum
// for the next step, so a step is three launches (train, reduce, Adam).
template <typename T>
__global__ __launch_bounds__(256) void k_nrc_adam(float* __restrict__ master, float* __restrict__ am,
                                                  float* __restrict__ av, float* __restrict__ ema,
                                                  const float* __restrict__ grad, uint32_t* state, float* loss_accum,
                                                  Hyper h, void* fop_train, void* bop_train, void* fop_infer) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t flags = state[kStFlags];
    if (p < kNrcParams && flags == 0u) {
        const float t = (float)(state[kStLt] + 1u);
        const float g = grad[p];
        const float m = h.beta1 * am[p] + (1.0f - h.beta1) * g;
        const float v = h.beta2 * av[p] + (1.0f - h.beta2) * g * g;
        const float mh = m / (1.0f - powf(h.beta1, t));
        const float vh = v / (1.0f - powf(h.beta2, t));
        const float w = master[p] - h.lr * mh / (sqrtf(vh) + h.eps);
        const float e = h.ema_decay * ema[p] + (1.0f - h.ema_decay) * w;
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
                state[kStLoss] = __float_as_uint(*loss_accum);
            } else if (flags & 1u) {
                state[kStSkipNan] += 1u;
            } else {
                state[kStSkipData] += 1u;
            }
            state[kStFlags] = 0u;
            *loss_accum = 0.0f;
            state[kStDone] = 0u;
        }
    }
}

// fop_tail = fop_infer (16-byte entries), once per call, before its steps
__global__ __launch_bounds__(256) void k_nrc_snapshot(const uint4* __restrict__ src, uint4* __restrict__ dst, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
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

// The records the next `steps` optimiser steps will draw, copied in sample
// order: step k (seed = the seed counter + k, as each step advances it by
// one) takes records[pcg32(seed ^ i) % n] for sample i, the draw
// k_nrc_train makes itself; plus the record count.  A copy lets the steps
// run after the tracer has started appending the next frame's records.
__global__ void k_nrc_gather(const float4* __restrict__ records, const uint32_t* __restrict__ n_rec, uint32_t min_records,
                             const uint32_t* __restrict__ state, uint32_t batch, uint32_t steps, float4* __restrict__ out,
                             uint32_t* __restrict__ out_n) {
    const uint32_t n = *n_rec;
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_n = n;
    if (n <= min_records) return;   // the steps will skip for lack of data
    const uint32_t seed0 = state[kStSeed];
    const uint64_t total = (uint64_t)batch * steps * kRecF4;
    for (uint64_t w = blockIdx.x * blockDim.x + threadIdx.x; w < total; w += gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(w / kRecF4), c = (uint32_t)(w - (uint64_t)j * kRecF4);
        const uint32_t k = j / batch, i = j - k * batch;
        const uint32_t idx = pcg32_step((seed0 + k) ^ i) % n;
        out[w] = records[(size_t)kRecF4 * idx + c];
    }
}

template <typename T>
int enqueue_train_steps(tfg_nrc* nc, const float4* records, const uint32_t* n_rec, uint32_t steps, hipStream_t s,
                        uint32_t min_records, bool gathered = false) {
    const uint32_t nblk = nc->cfg.batch / 64;
    TrainArgs A;
    A.records = records;
    A.gathered = gathered ? 1 : 0;
    A.n_rec = n_rec;
    A.min_records = min_records;
    A.state = nc->state;
    A.batch = nc->cfg.batch;
    A.grad_scale = nc->cfg.loss_scale / (3.0f * (float)nc->cfg.batch);
    A.loss_norm = 1.0f / (3.0f * (float)nc->cfg.batch);
    A.b = nc->bounds;
    A.fop_train = nc->fop_train;
    A.bop_train = nc->bop_train;
    A.fop_tail = nc->fop_tail;
    A.partial = nc->partial;
    A.loss_accum = nc->loss_accum;
    const Hyper h{nc->cfg.lr, nc->cfg.beta1, nc->cfg.beta2, nc->cfg.eps, nc->cfg.ema_decay, 1.0f / nc->cfg.loss_scale};
    const int pblocks = (kNrcParams + 255) / 256;
    const uint32_t op_entries = (uint32_t)kFopTiles * 64;
    if (steps > 0)
        hipLaunchKernelGGL(k_nrc_snapshot, dim3((op_entries + 255) / 256), dim3(256), 0, s,
                           static_cast<const uint4*>(nc->fop_infer), static_cast<uint4*>(nc->fop_tail), op_entries);
    for (uint32_t k = 0; k < steps; ++k) {
        if (gathered) A.records = records + (size_t)k * nc->cfg.batch * kRecF4;
        hipLaunchKernelGGL(k_nrc_train<T>, dim3(nblk), dim3(256), kTrainLds, s, A);
        hipLaunchKernelGGL(k_nrc_grad_reduce, dim3((kNrcParams + kReduceParams - 1) / kReduceParams), dim3(256), 0,
                           s, nc->partial, nblk, nc->grad, nc->state, n_rec, min_records, h.inv_loss_scale);
        hipLaunchKernelGGL(k_nrc_adam<T>, dim3(pblocks), dim3(256), 0, s, nc->master, nc->adam_m, nc->adam_v, nc->ema,
                           nc->grad, nc->state, nc->loss_accum, h, nc->fop_train, nc->bop_train, nc->fop_infer);
    }
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
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

size_t nrc_gather_records(const tfg_nrc* nc) {
    return (size_t)nc->cfg.batch * nc->cfg.steps_per_frame;
}

int nrc_gather_enqueue(tfg_nrc* nc, const tfg_nrc_frame& f, float4* out, uint32_t* out_n, hipStream_t s) {
    const size_t total = nrc_gather_records(nc) * kRecF4;
    const uint32_t grid = (uint32_t)std::min<size_t>((total + 255) / 256, (size_t)nc->n_cu * 16);
    hipLaunchKernelGGL(k_nrc_gather, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, f.records, f.counters + CNT_RECORDS,
                       nc->cfg.min_records, nc->state, nc->cfg.batch, nc->cfg.steps_per_frame, out, out_n);
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

int nrc_train_gathered_enqueue(tfg_nrc* nc, const float4* gathered, const uint32_t* n_rec, hipStream_t s) {
    if (nc->cfg.steps_per_frame == 0) return TFG_OK;
    return nc->cfg.precision == TFG_NRC_FP16
               ? enqueue_train_steps<_Float16>(nc, gathered, n_rec, nc->cfg.steps_per_frame, s, nc->cfg.min_records,
                                               true)
               : enqueue_train_steps<__bf16>(nc, gathered, n_rec, nc->cfg.steps_per_frame, s, nc->cfg.min_records, true);
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
    if (e == hipSuccess) e = hipMalloc((void**)&nc->partial, (size_t)(cfg->batch / 64) * pb);
    if (e == hipSuccess) e = hipMalloc(&nc->fop_train, (size_t)kFopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc(&nc->bop_train, (size_t)kBopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc(&nc->fop_infer, (size_t)kFopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc(&nc->fop_tail, (size_t)kFopTiles * 64 * kOpBytes16);
    if (e == hipSuccess) e = hipMalloc((void**)&nc->state, 16 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&nc->loss_accum, sizeof(float));
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
    TFG_HIP_CHECK(hipMemset(nc->loss_accum, 0, sizeof(float)));   // the steps' loss sum starts cleared
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
    void* bufs[] = {nc->master, nc->adam_m, nc->adam_v, nc->ema, nc->grad, nc->partial, nc->fop_train,
                    nc->bop_train, nc->fop_infer, nc->fop_tail, nc->state, nc->loss_accum, nc->dm_buf, nc->dm_out, nc->dm_count};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete nc;
    return TFG_OK;
}

}  // extern "C"
