// nrc_device.h - radiance-cache MLP on gfx950 matrix cores.
//
// Network (restating the reference's tcnn model, mlp.cpp:22-57, with the
// 12-frequency position encoding the reference keeps as its commented-out
// alternative, mlp.cpp:26-30): 64 encoded inputs -> 5 hidden layers of 64
// ReLU neurons -> 3 linear outputs, no biases (a constant-1 input feature
// gives the first layer an offset).  Weights are held in FP32 (master),
// used as BF16 or FP16 operands with FP32 accumulation.
//
// Matrix-core mapping.  Every layer is computed transposed, Y^T = W X^T,
// with v_mfma_f32_16x16x32_{bf16,f16}: A = a 16-neuron x 32-input weight
// tile, B = a 32-input x 16-sample activation tile.  The result tile leaves
// lane l holding sample l%16, neurons 4*(l/16)+i.  Reading the input slots of
// the next layer in the permuted order
//     slot (c, g, j) -> neuron 32c + (j < 4 ? 4g + j : 16 + 4g + j - 4)
// makes two result tiles exactly the next B operand of the same lane, so
// activations never leave registers between layers; the weight operands are
// stored pre-permuted (see nrc.hip) and staged once per workgroup in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kNrcWidth = 64;
constexpr int kNrcHidden = 5;                         // hidden 64x64 matrices: layers 0..4
constexpr int kNrcLayers = 6;                         // + output layer 5 (16 x 64, 3 used)
constexpr int kNrcOut = 3;
constexpr int kNrcParams = 5 * 64 * 64 + 16 * 64;     // 21504 (output rows padded to 16)
constexpr int kFopTiles = 5 * 8 + 2;                  // forward operand tiles (16 B per lane each)
constexpr int kBopTiles = 4 * 8 + 4;                  // backward (transposed) operand tiles, layers 1..5
constexpr float kLumR = 0.299f, kLumG = 0.587f, kLumB = 0.114f;

__host__ __device__ constexpr int nrc_layer_offset(int l) { return l * 64 * 64; }
__host__ __device__ constexpr int nrc_layer_rows(int l) { return l < 5 ? 64 : 16; }
__host__ __device__ constexpr int fop_tile_base(int l) { return l * 8; }
__host__ __device__ constexpr int bop_tile_base(int l) { return (l - 1) * 8; }   // l = 1..5

// input slot (c, g, j) -> input neuron / feature index
__host__ __device__ constexpr int nrc_perm(int c, int g, int j) { return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

template <typename T>
struct MfmaOps;
template <>
struct MfmaOps<__bf16> {
    typedef bf16x8 vec;
    static __device__ __forceinline__ floatx4 mfma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOps<_Float16> {
    typedef f16x8 vec;
    static __device__ __forceinline__ floatx4 mfma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

typedef float float2v __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));
typedef uint32_t uint4v __attribute__((ext_vector_type(4)));
template <typename T>
struct Pair16;
template <>
struct Pair16<__bf16> {
    typedef __bf16 type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair16<_Float16> {
    typedef _Float16 type __attribute__((ext_vector_type(2)));
};
// (a, b) -> one packed conversion to 16 bits (a in the low half): one
// instruction, where element-wise inserts into an 8-vector cost a conversion
// each plus a byte permute per pair
template <typename T>
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((float2v){a, b}, typename Pair16<T>::type));
}
// the same with ReLU (see pack_relu)
template <typename T>
__device__ __forceinline__ uint32_t cvt_relu2(float a, float b) {
    const short2v s = __builtin_elementwise_max(__builtin_bit_cast(short2v, cvt2<T>(a, b)), (short2v)0);
    return __builtin_bit_cast(uint32_t, s);
}

// ---------------------------------------------------------------- encoding
// Vertex record (4 float4): (pos, roughness), (dir mapped to [0,1], kd.r),
// (normal mapped to [0,1], kd.g), (kd.b, ks.rgb).
struct NrcBounds {
    float mn[3], inv_ext[3];
};

__device__ __forceinline__ float oneblob(float s, int bin) {
    const float d = (s - ((float)bin + 0.5f) * 0.25f) * 4.0f;
    return __expf(-0.5f * d * d);
}
__device__ __forceinline__ float sin_pi_pow2(float x, int k) {
    // sin(2^k * pi * x) for x in [0, 1]: 2^(k-1) x revolutions, reduced
    // exactly by fract (scaling by a power of two is exact).
    float r = x * __uint_as_float((uint32_t)(126 + k) << 23);
    r = r - floorf(r);
    return __builtin_amdgcn_sinf(r);
}
__device__ __forceinline__ void sph01(float mx, float my, float mz, float& th, float& ph) {
    const float x = 2.0f * mx - 1.0f, y = 2.0f * my - 1.0f, z = 2.0f * mz - 1.0f;
    th = acosf(fminf(1.0f, fmaxf(-1.0f, z))) * 0.31830988618379067f;
    ph = atan2f(y, x) * 0.15915494309189535f + 0.5f;
}

// The 16 input features lane-group g of a sample needs: chunk c, slot j.
// Slot meaning by (c, j), parameter g (0..3):
//  c0: j0 pos.x f=g, j1 pos.x f=4+g, j2 pos.x f=8+g, j3 pos.y f=g, j4 pos.y f=4+g,
//      j5 pos.y f=8+g, j6 pos.z f=g, j7 pos.z f=4+g
//  c1: j0 pos.z f=8+g, j1 dir theta blob g, j2 dir phi blob g, j3 normal theta
//      blob g, j4 normal phi blob g, j5 roughness blob g, j6 kd[g] (g=3: 1),
//      j7 ks[g] (g=3: 0)
// with pos features sin(2^f pi x), x = position normalised to [0,1] by the
// scene bounds (NrcKernel.cu:5-8).
template <typename T>
__device__ __forceinline__ void encode(const float4 v0, const float4 v1, const float4 v2, const float4 v3,
                                       const NrcBounds& b, int g, typename MfmaOps<T>::vec& x0,
                                       typename MfmaOps<T>::vec& x1) {
    const float px = fminf(fmaxf((v0.x - b.mn[0]) * b.inv_ext[0], 0.0f), 1.0f);
    const float py = fminf(fmaxf((v0.y - b.mn[1]) * b.inv_ext[1], 0.0f), 1.0f);
    const float pz = fminf(fmaxf((v0.z - b.mn[2]) * b.inv_ext[2], 0.0f), 1.0f);
    typedef typename MfmaOps<T>::vec V;
    const uint4v w0 = {cvt2<T>(sin_pi_pow2(px, g), sin_pi_pow2(px, 4 + g)),
                       cvt2<T>(sin_pi_pow2(px, 8 + g), sin_pi_pow2(py, g)),
                       cvt2<T>(sin_pi_pow2(py, 4 + g), sin_pi_pow2(py, 8 + g)),
                       cvt2<T>(sin_pi_pow2(pz, g), sin_pi_pow2(pz, 4 + g))};
    x0 = __builtin_bit_cast(V, w0);
    // Each lane converts one of the two unit vectors to angles (lane groups 0
    // and 1 the direction, 2 and 3 the normal) and takes the other vector's
    // angles from lane l ^ 32, which holds the same sample: one acos and one
    // atan2 per lane instead of two of each.  Callers that skip encode() for
    // some samples skip both lanes of a pair (same sample), so every lane
    // read here is active.
    const bool is_dir = g < 2;
    float th, ph;
    sph01(is_dir ? v1.x : v2.x, is_dir ? v1.y : v2.y, is_dir ? v1.z : v2.z, th, ph);
    const float th_o = __shfl_xor(th, 32), ph_o = __shfl_xor(ph, 32);
    const float dth = is_dir ? th : th_o, dph = is_dir ? ph : ph_o;
    const float nth = is_dir ? th_o : th, nph = is_dir ? ph_o : ph;
    const float kd[3] = {v1.w, v2.w, v3.x};
    const float ks[3] = {v3.y, v3.z, v3.w};
    const uint4v w1 = {cvt2<T>(sin_pi_pow2(pz, 8 + g), oneblob(dth, g)), cvt2<T>(oneblob(dph, g), oneblob(nth, g)),
                       cvt2<T>(oneblob(nph, g), oneblob(v0.w, g)),
                       cvt2<T>(g < 3 ? kd[g] : 1.0f, g < 3 ? ks[g] : 0.0f)};
    x1 = __builtin_bit_cast(V, w1);
}

// D tiles (4 per 64 neurons) -> B operands of the next layer, with ReLU.
// The ReLU is taken after the conversion, on the 16-bit patterns as signed
// integers (two per packed integer max): a negative bf16 / fp16 value, -0
// included, is a negative integer, a positive one keeps its order, so every
// finite pre-activation gives the same operand as converting max(x, 0), in 16
// VALU instead of 40 (a float max also canonicalises its input first).  A
// NaN pre-activation stays NaN when its sign bit is clear.
template <typename T>
__device__ __forceinline__ typename MfmaOps<T>::vec pack_relu(floatx4 lo, floatx4 hi) {
    const uint4v w = {cvt_relu2<T>(lo[0], lo[1]), cvt_relu2<T>(lo[2], lo[3]), cvt_relu2<T>(hi[0], hi[1]),
                      cvt_relu2<T>(hi[2], hi[3])};
    return __builtin_bit_cast(typename MfmaOps<T>::vec, w);
}

// One hidden layer for G 16-sample groups: x <- relu(W_l x).
template <typename T, int G>
__device__ __forceinline__ void mlp_layer(const typename MfmaOps<T>::vec* __restrict__ W, int lane, int l,
                                          typename MfmaOps<T>::vec (&x)[G][2]) {
    typedef MfmaOps<T> M;
    floatx4 acc[G][4];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[g][t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const typename M::vec a = W[(fop_tile_base(l) + 2 * t + c) * 64 + lane];
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g][t] = M::mfma(a, x[g][c], acc[g][t]);
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        x[g][0] = pack_relu<T>(acc[g][0], acc[g][1]);
        x[g][1] = pack_relu<T>(acc[g][2], acc[g][3]);
    }
}

template <typename T, int G>
__device__ __forceinline__ void mlp_output(const typename MfmaOps<T>::vec* __restrict__ W, int lane,
                                           const typename MfmaOps<T>::vec (&x)[G][2], floatx4 (&out)[G]) {
    typedef MfmaOps<T> M;
#pragma unroll
    for (int g = 0; g < G; ++g) out[g] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const typename M::vec a = W[(fop_tile_base(5) + c) * 64 + lane];
#pragma unroll
        for (int g = 0; g < G; ++g) out[g] = M::mfma(a, x[g][c], out[g]);
    }
}

// Forward through all layers for `G` 16-sample groups of one wave.  W is the
// forward operand array in LDS (kFopTiles x 64 lanes).  On return out[g] holds
// the output tile: lanes with l/16 == 0 own outputs 0..2 of sample l%16.
// The layer loop is kept rolled so the weights of one layer at a time are
// live.
template <typename T, int G>
__device__ __forceinline__ void mlp_forward(const typename MfmaOps<T>::vec* __restrict__ W, int lane,
                                            typename MfmaOps<T>::vec (&x)[G][2], floatx4 (&out)[G]) {
#pragma unroll 1
    for (int l = 0; l < kNrcHidden; ++l) mlp_layer<T, G>(W, lane, l, x);
    mlp_output<T, G>(W, lane, x, out);
}

// Training forward for one group: also keeps the B operands (layer inputs)
// of every layer in xs; unrolled so xs stays in registers.
template <typename T>
__device__ __forceinline__ void mlp_forward_keep(const typename MfmaOps<T>::vec* __restrict__ W, int lane,
                                                 typename MfmaOps<T>::vec (&x)[1][2], floatx4 (&out)[1],
                                                 typename MfmaOps<T>::vec (&xs)[kNrcLayers][2]) {
#pragma unroll
    for (int l = 0; l < kNrcHidden; ++l) {
        xs[l][0] = x[0][0];
        xs[l][1] = x[0][1];
        mlp_layer<T, 1>(W, lane, l, x);
    }
    xs[5][0] = x[0][0];
    xs[5][1] = x[0][1];
    mlp_output<T, 1>(W, lane, x, out);
}

// guardarSalidaInferencia (NrcKernel.cu:99-126)
__device__ __forceinline__ float sanitize_out(float r) {
    if (r < 0.0f) r = 0.0f;
    if (isnan(r)) r = 0.0f;
    if (r > 100000.0f) r = 100000.0f;
    if (isinf(r)) r = 100000.0f;
    return r;
}

}  // namespace tfg
