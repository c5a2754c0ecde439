// trace_device.h - CWBVH closest-hit / any-hit traversal for gfx950.
//
// Decodes the 80 B node format of traverse_cwbvh.cuh:75-185 with CDNA4
// integer ops (v_ffbh_u32 for the highest set bit, v_bcnt_u32_b32 for the
// child rank, v_cvt_f32_ubyteN for the quantised bounds) and intersects the
// 64 B triangle records with the FP32 watertight test of Woop, Benthin and
// Wald (JCGT 2013) or, for the affine record format, the unit-triangle test
// of traverse_cwbvh.cuh:189-209.
//
// Every FP32 operation is written out in the order the CPU oracle uses and
// the translation unit is compiled with -ffp-contract=off, so hit ids, t, u
// and v are bit-identical to the oracle on the same rays.
//
// Result semantics: closest hit with tmin < t <= tmax; equal t resolved to
// the smaller primitive index (so the result does not depend on traversal
// order); any-hit reports a hit with tmin < t < tmax.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfg {

constexpr float kBoxSlack = 1.00000095367431640625f;   // 1 + 2^-20

struct RayPre {
    float O[3], D[3], rD[3];
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    uint32_t octinv;
};

// 1 / d, correctly rounded, in 3 instructions instead of the division's 10:
// v_rcp_f32 and one Newton step on its residual give the correctly rounded
// reciprocal of every normal d with |d| < 2^126 (checked exhaustively on
// MI355X against the IEEE division: both signs, all 2^23 significands of
// every exponent -126..125, no difference).  Other d (a subnormal
// reciprocal, which v_rcp_f32 returns as 0; subnormal d; infinities; NaN)
// take the division.
__device__ __forceinline__ float rcp_rn(float d) {
    const float a = fabsf(d);
    if (__builtin_expect(!(a >= 0x1p-126f && a < 0x1p126f), 0)) return 1.0f / d;
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
}

__device__ __forceinline__ float safe_rcp(float d) {
    if (fabsf(d) < 1e-30f) d = copysignf(1e-30f, d);
    return rcp_rn(d);
}

// Component k of (x, y, z) as two selects on values (no branches, no
// indexed array: the compares are per ray and leave the traversal loop).
__device__ __forceinline__ float sel3(float x, float y, float z, int k) {
    const float r = k == 1 ? y : x;
    return k == 2 ? z : r;
}
__device__ __forceinline__ float sel3(const float* v, int k) { return sel3(v[0], v[1], v[2], k); }

__device__ __forceinline__ RayPre make_ray(float ox, float oy, float oz, float dx, float dy, float dz, float tmin,
                                           float tmax) {
    RayPre p;
    p.O[0] = ox; p.O[1] = oy; p.O[2] = oz;
    p.D[0] = dx; p.D[1] = dy; p.D[2] = dz;
    p.rD[0] = safe_rcp(dx);
    p.rD[1] = safe_rcp(dy);
    p.rD[2] = safe_rcp(dz);
    p.tmin = tmin;
    p.tmax = tmax;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dkz = sel3(p.D, kz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    // the shear constants of the watertight test from the reciprocal already
    // computed (kz is the dominant axis); any constants shared by all of a
    // ray's triangles keep the test watertight
    p.Sz = sel3(p.rD, kz);
    p.Sx = sel3(p.D, kx) * p.Sz;
    p.Sy = sel3(p.D, ky) * p.Sz;
    p.octinv = 7u - ((p.rD[0] < 0.0f ? 4u : 0u) | (p.rD[1] < 0.0f ? 2u : 0u) | (p.rD[2] < 0.0f ? 1u : 0u));
    return p;
}

// Watertight test.  Vertices arrive as float4 (w ignored).  Accepts the hit
// when t = T / det lies in (tmin, tcur]; the division is done only for
// candidates that pass a division-free screen of T against tmin * det and
// tcur * det (widened by 2^-20 relative, far beyond the 2^-24 rounding of
// either side, so the screen never rejects a hit the exact test accepts;
// |det| < 2^-100 always takes the exact path), and u = V / det, v = W / det
// only for accepted hits and when kUV.  The accepted t, u, v are the same
// divisions as always, so results do not depend on the screen.
template <bool kUV>
__device__ __forceinline__ bool watertight(const float4 a, const float4 b, const float4 c, const RayPre& r, float tcur,
                                           float& t, float& u, float& v) {
    const float ax = a.x - r.O[0], ay = a.y - r.O[1], az = a.z - r.O[2];
    const float bx = b.x - r.O[0], by = b.y - r.O[1], bz = b.z - r.O[2];
    const float cx = c.x - r.O[0], cy = c.y - r.O[1], cz = c.z - r.O[2];
    const float Akz = sel3(ax, ay, az, r.kz), Bkz = sel3(bx, by, bz, r.kz), Ckz = sel3(cx, cy, cz, r.kz);
    const float Ax = sel3(ax, ay, az, r.kx) - r.Sx * Akz;
    const float Ay = sel3(ax, ay, az, r.ky) - r.Sy * Akz;
    const float Bx = sel3(bx, by, bz, r.kx) - r.Sx * Bkz;
    const float By = sel3(bx, by, bz, r.ky) - r.Sy * Bkz;
    const float Cx = sel3(cx, cy, cz, r.kx) - r.Sx * Ckz;
    const float Cy = sel3(cx, cy, cz, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    // bitwise (not short-circuit) conditions: one branch each instead of a chain
    if ((U == 0.0f) | (V == 0.0f) | (W == 0.0f)) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if (((U < 0.0f) | (V < 0.0f) | (W < 0.0f)) & ((U > 0.0f) | (V > 0.0f) | (W > 0.0f))) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float ad = fabsf(det);
    if (ad >= 0x1p-100f) {
        const float Ts = det < 0.0f ? -T : T;
        if (Ts <= r.tmin * ad * (1.0f - 0x1p-20f)) return false;
        if (Ts > tcur * ad * (1.0f + 0x1p-20f)) return false;
    }
    t = T / det;
    if (!(t > r.tmin && t <= tcur)) return false;
    if (kUV) {
        u = V / det;
        v = W / det;
    }
    return true;
}

// Affine unit-triangle test (traverse_cwbvh.cuh:189-209), T2 first so a
// plane miss costs one 16 B load.
__device__ __forceinline__ bool affine(const float4* __restrict__ tri, const RayPre& r, float tcur, float& t,
                                       float& u, float& v) {
    const float4 T2 = tri[2];
    const float transS = T2.x * r.O[0] + T2.y * r.O[1] + T2.z * r.O[2] + T2.w;
    const float transD = T2.x * r.D[0] + T2.y * r.D[1] + T2.z * r.D[2];
    const float d = -transS / transD;
    if (!(d > r.tmin) || d > tcur) return false;
    const float4 T0 = tri[0];
    const float4 T1 = tri[1];
    const float Ix = r.O[0] + d * r.D[0];
    const float Iy = r.O[1] + d * r.D[1];
    const float Iz = r.O[2] + d * r.D[2];
    const float uu = T0.x * Ix + T0.y * Iy + T0.z * Iz + T0.w;
    const float vv = T1.x * Ix + T1.y * Iy + T1.z * Iz + T1.w;
    if (uu >= 0.0f && vv >= 0.0f && uu + vv < 1.0f) {
        t = d;
        u = uu;
        v = vv;
        return true;
    }
    return false;
}

__device__ __forceinline__ float ubyte_f(uint32_t w, int j) { return (float)((w >> (8 * j)) & 0xFFu); }

// Quantised bounds as FP16 subnormals: the half with bit pattern b is
// b * 2^-24, exactly, so v_perm_b32 turns a word of 4 bounds into two
// registers of two halves each, and v_fma_mix_f32 reads either half as an
// f32 operand: 2 + 4 instructions per 4 bounds instead of 4 v_cvt_f32_ubyte
// and 4 v_fma_f32.  Needs FP16 subnormals kept, not flushed (the kernels'
// default FP16/FP64 denormal mode; the traversal parity tests fail at once
// otherwise).
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
struct SubBytes {
    uint32_t p01, p23;   // halves {b0, b1}, {b2, b3}
    __device__ __forceinline__ explicit SubBytes(uint32_t w)
        : p01(__builtin_amdgcn_perm(0u, w, 0x0c010c00u)), p23(__builtin_amdgcn_perm(0u, w, 0x0c030c02u)) {}
    __device__ __forceinline__ float operator[](int j) const {   // byte j * 2^-24
        const half2v h = __builtin_bit_cast(half2v, j < 2 ? p01 : p23);
        return (float)((j & 1) ? h.y : h.x);
    }
};
constexpr float kSlabScale = 0x1p-24f;

// One child box test, the oracle's node_hitmask loop body: lo* / hi* are the
// child's near and far quantised bounds per axis (the low or high bound by
// the sign of the ray direction on that axis).
__device__ __forceinline__ uint32_t child_bits(uint32_t meta, float lox, float loy, float loz, float hix, float hiy,
                                               float hiz, float idirx, float idiry, float idirz, float origx,
                                               float origy, float origz, const RayPre& r, float tcur) {
    const float tminx = fmaf(lox, idirx, origx), tmaxx = fmaf(hix, idirx, origx);
    const float tminy = fmaf(loy, idiry, origy), tmaxy = fmaf(hiy, idiry, origy);
    const float tminz = fmaf(loz, idirz, origz), tmaxz = fmaf(hiz, idirz, origz);
    const float tnear = fmaxf(fmaxf(fmaxf(tminx, tminy), tminz), r.tmin);
    const float tfar = fminf(fminf(fminf(tmaxx, tmaxy), tmaxz), tcur);
    if (!(tnear <= tfar * kBoxSlack)) return 0u;
    const uint32_t is_inner = (meta & (meta << 1)) & 0x10u;
    const uint32_t bit_index = (is_inner ? (meta ^ r.octinv) : meta) & 0x1Fu;
    return ((meta >> 5) & 7u) << bit_index;
}

// Decodes node (n0..n4) and returns the 32-bit hit mask.  kHalf: the bounds
// as FP16 subnormals through v_fma_mix_f32 with slab distances scaled by
// 2^-24 (fewer instructions, 6 more registers), else v_cvt_f32_ubyte and
// unscaled distances; both give the same mask (tests/cpp/node_check.hip), so
// each kernel takes the form that leaves it the better occupancy.
template <bool kHalf = true>
__device__ __forceinline__ uint32_t node_hitmask(const float4 n0, const float4 n1, const float4 n2, const float4 n3,
                                                 const float4 n4, const RayPre& r, float tcur) {
    const uint32_t w = __float_as_uint(n0.w);
    const int ex = (int)(int8_t)(w & 0xFFu), ey = (int)(int8_t)((w >> 8) & 0xFFu), ez = (int)(int8_t)((w >> 16) & 0xFFu);
    const float idirx = __uint_as_float((uint32_t)(ex + 127) << 23) * r.rD[0];
    const float idiry = __uint_as_float((uint32_t)(ey + 127) << 23) * r.rD[1];
    const float idirz = __uint_as_float((uint32_t)(ez + 127) << 23) * r.rD[2];
    // Slab distances scaled by 2^-24: fma(b * 2^-24, idir, orig * 2^-24) is
    // 2^-24 * fma(b, idir, orig) (the same rounding), and the test compares
    // them only with each other and with tmin and tcur scaled alike, so every
    // decision is the unscaled test's while the scaled values stay normal
    // (|t| >= 2^-102).
    constexpr float kS = kHalf ? kSlabScale : 1.0f;
    const float origx = (n0.x - r.O[0]) * r.rD[0] * kS;
    const float origy = (n0.y - r.O[1]) * r.rD[1] * kS;
    const float origz = (n0.z - r.O[2]) * r.rD[2] * kS;
    const float tmin_s = r.tmin * kS, tcur_s = tcur * kS;
    const uint32_t m[2] = {__float_as_uint(n1.z), __float_as_uint(n1.w)};
    // near / far bound words per axis, chosen once per node by the direction signs
    const bool nx = r.rD[0] < 0.0f, ny = r.rD[1] < 0.0f, nz = r.rD[2] < 0.0f;
    const uint32_t qlx[2] = {__float_as_uint(n2.x), __float_as_uint(n2.y)};
    const uint32_t qly[2] = {__float_as_uint(n2.z), __float_as_uint(n2.w)};
    const uint32_t qlz[2] = {__float_as_uint(n3.x), __float_as_uint(n3.y)};
    const uint32_t qhx[2] = {__float_as_uint(n3.z), __float_as_uint(n3.w)};
    const uint32_t qhy[2] = {__float_as_uint(n4.x), __float_as_uint(n4.y)};
    const uint32_t qhz[2] = {__float_as_uint(n4.z), __float_as_uint(n4.w)};
    uint32_t hitmask = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t lx = nx ? qhx[h] : qlx[h], hx = nx ? qlx[h] : qhx[h];
        const uint32_t ly = ny ? qhy[h] : qly[h], hy = ny ? qly[h] : qhy[h];
        const uint32_t lz = nz ? qhz[h] : qlz[h], hz = nz ? qlz[h] : qhz[h];
        // child_bits' bit index and triangle bits for the word's 4 children at
        // once: inner children (meta bits 3 and 4 set) get the octant xor
        const uint32_t mw = m[h];
        const uint32_t inner = (mw & (mw << 1)) & 0x10101010u;
        const uint32_t idx4 = (mw ^ ((r.octinv * 0x01010101u) & ((inner >> 4) * 0x1Fu))) & 0x1F1F1F1Fu;
        const uint32_t cnt4 = (mw >> 5) & 0x07070707u;
        const SubBytes blx(lx), bhx(hx), bly(ly), bhy(hy), blz(lz), bhz(hz);
        // bound j of word w: b * 2^-24 (kHalf) or b
        auto bnd = [&](const SubBytes& sb, uint32_t w, int j) { return kHalf ? sb[j] : ubyte_f(w, j); };
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float tminx = fmaf(bnd(blx, lx, j), idirx, origx), tmaxx = fmaf(bnd(bhx, hx, j), idirx, origx);
            const float tminy = fmaf(bnd(bly, ly, j), idiry, origy), tmaxy = fmaf(bnd(bhy, hy, j), idiry, origy);
            const float tminz = fmaf(bnd(blz, lz, j), idirz, origz), tmaxz = fmaf(bnd(bhz, hz, j), idirz, origz);
            const float tnear = fmaxf(fmaxf(fmaxf(tminx, tminy), tminz), tmin_s);
            const float tfar = fminf(fminf(fminf(tmaxx, tmaxy), tmaxz), tcur_s);
            const uint32_t cnt = tnear <= tfar * kBoxSlack ? (cnt4 >> (8 * j)) & 7u : 0u;
            hitmask |= cnt << ((idx4 >> (8 * j)) & 0x1Fu);
        }
    }
    return hitmask;
}

__device__ __forceinline__ int highest_bit(uint32_t x) { return 31 - __clz((int)x); }

struct TraceStats {
    uint32_t nodes = 0, tris = 0, overflow = 0;
};

// Triangles of one node's leaf children (tg: base, 24-bit hit mask) against
// the ray.  Watertight records: the vertices of two triangles are fetched
// together, so a ray that tests many triangles pays one memory latency per
// pair.  The accept rule (nearer t, then smaller primIdx) does not depend on
// the test order.  Returns true for an any-hit ray that found an occluder.
template <int kFormat, bool kAnyHit>
__device__ __forceinline__ bool test_triangles(const float4* __restrict__ tris, uint32_t tg_x, uint32_t tg_y,
                                               const RayPre& r, float& tcur, float& bu, float& bv,
                                               uint32_t& best_prim, bool& found, TraceStats* stats) {
    if (kFormat == 0) {
        while (tg_y != 0) {
            const int b0 = highest_bit(tg_y);
            tg_y &= ~(1u << b0);
            const bool two = tg_y != 0;
            int b1 = b0;
            if (two) {
                b1 = highest_bit(tg_y);
                tg_y &= ~(1u << b1);
            }
            const float4* T0 = tris + (size_t)(tg_x + 4u * (uint32_t)b0);
            const float4* T1 = tris + (size_t)(tg_x + 4u * (uint32_t)b1);
            const float4 a0 = T0[0], a1 = T0[1], a2 = T0[2];
            const float4 c0 = T1[0], c1 = T1[1], c2 = T1[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (k == 1 && !two) break;
                if (stats) stats->tris++;
                float t, u, v;
                const bool ok = k == 0 ? watertight<!kAnyHit>(a0, a1, a2, r, tcur, t, u, v)
                                       : watertight<!kAnyHit>(c0, c1, c2, r, tcur, t, u, v);
                if (!ok) continue;
                const uint32_t prim = __float_as_uint((k == 0 ? T0 : T1)[3].w);
                if (kAnyHit) {
                    if (t < r.tmax) return true;
                    continue;
                }
                if (t < tcur || prim < best_prim) {
                    tcur = t;
                    bu = u;
                    bv = v;
                    best_prim = prim;
                    found = true;
                }
            }
        }
    } else {
        while (tg_y != 0) {
            const int b = highest_bit(tg_y);
            tg_y &= ~(1u << b);
            const float4* T = tris + (size_t)(tg_x + 4u * (uint32_t)b);
            if (stats) stats->tris++;
            float t, u, v;
            if (!affine(T, r, tcur, t, u, v)) continue;
            const uint32_t prim = __float_as_uint(T[3].w);
            if (kAnyHit) {
                if (t < r.tmax) return true;
                continue;
            }
            if (t < tcur || prim < best_prim) {
                tcur = t;
                bu = u;
                bv = v;
                best_prim = prim;
                found = true;
            }
        }
    }
    return false;
}

// Node ni from the LDS copy of the top nodes or from global memory, as two
// separate paths so that each compiles to its own load kind (LDS reads,
// global loads) rather than generic-address loads that wait on both.
__device__ __forceinline__ void load_node(const float4* __restrict__ nodes, const float4* __restrict__ top_nodes,
                                          uint32_t n_top, uint32_t ni, float4& n0, float4& n1, float4& n2,
                                          float4& n3, float4& n4) {
    if (ni < n_top) {
        const float4* n = top_nodes + 5 * ni;
        n0 = n[0];
        n1 = n[1];
        n2 = n[2];
        n3 = n[3];
        n4 = n[4];
    } else {
        const float4* n = nodes + 5 * (size_t)ni;
        n0 = n[0];
        n1 = n[1];
        n2 = n[2];
        n3 = n[3];
        n4 = n[4];
    }
}

// Per-lane LDS stack: entry k of this lane lives at base[k * stride]; it
// holds at most stack_depth entries (tree depth - 1 suffices: one pending
// node group per level).  Nodes with index < n_top are read from the LDS
// copy top_nodes (the first levels of the breadth-first node array, shared
// by every ray of the block).
//
// Each step first selects the next node (from the current node group, or the
// group on top of the stack) and issues its fetch, then tests the triangles
// of the previous node while that fetch is in flight, then intersects the
// new node's children with the ray.  The node order does not depend on the
// triangle results, and each node's child test sees every triangle hit found
// before it, so visits and results are those of testing a node's triangles
// before selecting the next node.
template <int kFormat, bool kAnyHit, bool kHalf = true>
__device__ __forceinline__ bool traverse(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                         const float4* __restrict__ top_nodes, uint32_t n_top, const RayPre& r,
                                         uint2* stack, int stride, int stack_depth, float& t_out, float& u_out,
                                         float& v_out, uint32_t& prim_out, TraceStats* stats) {
    int sp = 0;
    uint32_t ng_x = 0, ng_y = 0x80000000u;   // node group: child base, (hits << 24) | imask
    uint32_t tg_x = 0, tg_y = 0;             // triangles of the previous node
    float tcur = r.tmax;
    uint32_t best_prim = 0xFFFFFFFFu;
    float bu = 0.f, bv = 0.f;
    bool found = false;
    for (;;) {
        if (ng_y <= 0x00FFFFFFu && sp > 0) {
            --sp;
            const uint2 e = stack[sp * stride];
            ng_x = e.x;
            ng_y = e.y;
        }
        const bool have = ng_y > 0x00FFFFFFu;
        float4 n0, n1, n2, n3, n4;
        if (have) {
            const uint32_t imask = ng_y;
            const int bit = highest_bit(ng_y & 0xFF000000u);
            ng_y &= ~(1u << bit);
            if (ng_y > 0x00FFFFFFu) {
                if (sp < stack_depth) {
                    stack[sp * stride] = make_uint2(ng_x, ng_y);
                    ++sp;
                } else if (stats) {
                    stats->overflow = 1;   // host refuses trees this deep
                }
            }
            const uint32_t slot = ((uint32_t)bit - 24u) ^ r.octinv;
            const uint32_t rel = (uint32_t)__popc(imask & ~(0xFFFFFFFFu << slot));
            const uint32_t ni = ng_x + rel;
            load_node(nodes, top_nodes, n_top, ni, n0, n1, n2, n3, n4);
        }
        if (test_triangles<kFormat, kAnyHit>(tris, tg_x, tg_y, r, tcur, bu, bv, best_prim, found, stats))
            return true;
        if (!have) break;
        if (stats) stats->nodes++;
        const uint32_t hm = node_hitmask<kHalf>(n0, n1, n2, n3, n4, r, tcur);
        ng_x = __float_as_uint(n1.x);
        ng_y = (hm & 0xFF000000u) | (__float_as_uint(n0.w) >> 24);
        tg_x = __float_as_uint(n1.y);
        tg_y = hm & 0x00FFFFFFu;
    }
    if (kAnyHit) return false;
    t_out = tcur;
    u_out = found ? bu : 0.f;
    v_out = found ? bv : 0.f;
    prim_out = found ? best_prim : 0xFFFFFFFFu;
    return found;
}

// ------------------------------------------------------------------------
// Group traversal: 8 adjacent lanes (an aligned octet of the wave) trace one
// ray together.  Lane j of the octet tests child slot j of each node and, per
// round, the j-th triangle of the node's pending leaf triangles; the octet
// combines the child bits by OR and the triangle hits by a (t, primIdx)
// minimum over shuffles.  The octet's traversal state is uniform (every lane
// holds the same copy); its stack lives in LDS, stack_depth entries of
// `stride` uint2, written identically by all 8 lanes.  A ray's sequence of
// dependent steps is the same as in traverse(), but each step does an eighth
// of the box tests and up to 8 triangle tests at once, so a ray's latency is
// shorter: the form for launches with too few rays to fill the GPU.
// Results (t, u, v, primIdx; occlusion) equal traverse()'s: the accept rule
// does not depend on the test order.
__device__ __forceinline__ uint32_t octet_or(uint32_t x) {
    x |= __shfl_xor(x, 1);
    x |= __shfl_xor(x, 2);
    x |= __shfl_xor(x, 4);
    return x;
}

template <bool kAnyHit>
__device__ __forceinline__ bool traverse_octet(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                               const float4* __restrict__ top_nodes, uint32_t n_top, const RayPre& r,
                                               uint2* stack, int stride, int stack_depth, float& t_out, float& u_out,
                                               float& v_out, uint32_t& prim_out, TraceStats* stats) {
    const uint32_t j = __lane_id() & 7u;
    int sp = 0;
    uint32_t ng_x = 0, ng_y = 0x80000000u;
    uint32_t tg_x = 0, tg_y = 0;
    float tcur = r.tmax;
    uint32_t best_prim = 0xFFFFFFFFu;
    float bu = 0.f, bv = 0.f;
    bool found = false;
    for (;;) {
        if (ng_y <= 0x00FFFFFFu && sp > 0) {
            --sp;
            const uint2 e = stack[sp * stride];
            ng_x = e.x;
            ng_y = e.y;
        }
        const bool have = ng_y > 0x00FFFFFFu;
        float4 n0, n1, n2, n3, n4;
        if (have) {
            const uint32_t imask = ng_y;
            const int bit = highest_bit(ng_y & 0xFF000000u);
            ng_y &= ~(1u << bit);
            if (ng_y > 0x00FFFFFFu) {
                if (sp < stack_depth) {
                    stack[sp * stride] = make_uint2(ng_x, ng_y);
                    ++sp;
                } else if (stats && j == 0) {
                    stats->overflow = 1;
                }
            }
            const uint32_t slot = ((uint32_t)bit - 24u) ^ r.octinv;
            const uint32_t rel = (uint32_t)__popc(imask & ~(0xFFFFFFFFu << slot));
            const uint32_t ni = ng_x + rel;
            load_node(nodes, top_nodes, n_top, ni, n0, n1, n2, n3, n4);
        }
        // triangles of the previous node, up to 8 per round: lane j takes
        // the (j+1)-th highest pending bit
        while (tg_y != 0) {
            uint32_t m = tg_y;
            for (uint32_t k = 0; k < j && m != 0; ++k) m &= ~(1u << highest_bit(m));
            const bool mine = m != 0;
            bool ok = false;
            float t = 0.f, u = 0.f, v = 0.f;
            uint32_t prim = 0xFFFFFFFFu;
            if (mine) {
                const int b = highest_bit(m);
                const float4* T = tris + (size_t)(tg_x + 4u * (uint32_t)b);
                if (stats) stats->tris++;
                ok = watertight<!kAnyHit>(T[0], T[1], T[2], r, tcur, t, u, v);
                if (ok) prim = __float_as_uint(T[3].w);
            }
            // drop the (up to) 8 highest pending bits
            for (int k = 0; k < 8 && tg_y != 0; ++k) tg_y &= ~(1u << highest_bit(tg_y));
            if (kAnyHit) {
                if (octet_or((ok && t < r.tmax) ? 1u : 0u)) return true;
                continue;
            }
            // octet minimum of (t, primIdx) over accepted candidates
            float bt = ok ? t : __int_as_float(0x7f800000);
            uint32_t bp = ok ? prim : 0xFFFFFFFFu;
            float uu = u, vv = v;
#pragma unroll
            for (int o = 1; o < 8; o <<= 1) {
                const float ot = __shfl_xor(bt, o);
                const uint32_t op = __shfl_xor(bp, o);
                const float ou = __shfl_xor(uu, o), ov = __shfl_xor(vv, o);
                if (ot < bt || (ot == bt && op < bp)) {
                    bt = ot;
                    bp = op;
                    uu = ou;
                    vv = ov;
                }
            }
            if (bp != 0xFFFFFFFFu && (bt < tcur || bp < best_prim)) {
                tcur = bt;
                bu = uu;
                bv = vv;
                best_prim = bp;
                found = true;
            }
        }
        if (!have) break;
        if (stats && j == 0) stats->nodes++;
        // lane j: child slot j
        const uint32_t w = __float_as_uint(n0.w);
        const int ex = (int)(int8_t)(w & 0xFFu), ey = (int)(int8_t)((w >> 8) & 0xFFu),
                  ez = (int)(int8_t)((w >> 16) & 0xFFu);
        const float idirx = __uint_as_float((uint32_t)(ex + 127) << 23) * r.rD[0];
        const float idiry = __uint_as_float((uint32_t)(ey + 127) << 23) * r.rD[1];
        const float idirz = __uint_as_float((uint32_t)(ez + 127) << 23) * r.rD[2];
        const float origx = (n0.x - r.O[0]) * r.rD[0];
        const float origy = (n0.y - r.O[1]) * r.rD[1];
        const float origz = (n0.z - r.O[2]) * r.rD[2];
        const int h = (int)(j >> 2), jj = (int)(j & 3u);
        const uint32_t meta = ((h ? __float_as_uint(n1.w) : __float_as_uint(n1.z)) >> (8 * jj)) & 0xFFu;
        const uint32_t lx = h ? __float_as_uint(n2.y) : __float_as_uint(n2.x);
        const uint32_t ly = h ? __float_as_uint(n2.w) : __float_as_uint(n2.z);
        const uint32_t lz = h ? __float_as_uint(n3.y) : __float_as_uint(n3.x);
        const uint32_t hx = h ? __float_as_uint(n3.w) : __float_as_uint(n3.z);
        const uint32_t hy = h ? __float_as_uint(n4.y) : __float_as_uint(n4.x);
        const uint32_t hz = h ? __float_as_uint(n4.w) : __float_as_uint(n4.z);
        const bool nx = r.rD[0] < 0.0f, ny = r.rD[1] < 0.0f, nz = r.rD[2] < 0.0f;   // near / far by sign
        const uint32_t bits = child_bits(meta, ubyte_f(nx ? hx : lx, jj), ubyte_f(ny ? hy : ly, jj),
                                         ubyte_f(nz ? hz : lz, jj), ubyte_f(nx ? lx : hx, jj), ubyte_f(ny ? ly : hy, jj),
                                         ubyte_f(nz ? lz : hz, jj), idirx, idiry, idirz, origx, origy, origz, r, tcur);
        const uint32_t hm = octet_or(bits);
        ng_x = __float_as_uint(n1.x);
        ng_y = (hm & 0xFF000000u) | (__float_as_uint(n0.w) >> 24);
        tg_x = __float_as_uint(n1.y);
        tg_y = hm & 0x00FFFFFFu;
    }
    if (kAnyHit) return false;
    t_out = tcur;
    u_out = found ? bu : 0.f;
    v_out = found ? bv : 0.f;
    prim_out = found ? best_prim : 0xFFFFFFFFu;
    return found;
}

}  // namespace tfg
