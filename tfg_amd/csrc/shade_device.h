// shade_device.h - surface interaction for the wavefront path tracer.
//
// Restates render.h:33-413 and escenario.h:123-184 in FP32:
//   * pixel seed and PCG32 stream             rng.h:12-35
//   * pinhole primary ray                     camara.h:42-59
//   * lobe choice proportional to max(kd), max(ks), max(kt)  render.h:268-288
//   * cosine hemisphere sample + tangent frame  render.h:379-413
//   * mirror reflection / refraction          render.h:337-376
//   * point-light NEE with distance clamp     escenario.h:144-184
// and adds a GGX microfacet specular lobe (roughness > 0), which the
// reference does not have, sampled by visible normals (Heitz, JCGT 2018).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tfg_types.h"

namespace tfg {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kEps = 1e-3f;   // render.h:27, escenario.h:40

// ------------------------------------------------------------ rng.h
__device__ __forceinline__ uint32_t pcg32(uint32_t& s) {
    s = s * 747796405u + 2891336453u;
    const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float pcg32_float(uint32_t& s) { return (float)pcg32(s) * (1.0f / 4294967296.0f); }
__device__ __forceinline__ uint32_t pixel_seed(int px, int py, int frame) {
    uint32_t s = (uint32_t)(py << 16) | (uint32_t)(px & 0xFFFF);
    s ^= (uint32_t)frame * 719393u;
    s = (s ^ 61u) ^ (s >> 16u);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15u);
    return s == 0u ? 1u : s;
}

// ------------------------------------------------------------ vectors
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Shading arithmetic uses the hardware reciprocal, square root and sine
// (v_rcp_f32, v_sqrt_f32, v_sin_f32 / v_cos_f32 on revolutions, about 1 ulp
// or 2^-20 absolute) instead of the correctly rounded sequences the build
// gives "/" and sqrtf (8-10 instructions each).  Shading values are compared
// with the oracle within a tolerance; the ray set-up and the triangle test of
// the traversal, whose hits must equal the oracle's bit for bit, keep the
// correctly rounded forms.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * frcp(b); }
// sin and cos of 2 pi r for r in [0, 1)
__device__ __forceinline__ void sincos_rev(float r, float& s, float& c) {
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}
__device__ __forceinline__ V3 normalize(V3 a) {
    const float l = fsqrt(dot(a, a));
    return l > 0.0f ? a * frcp(l) : a;
}
__device__ __forceinline__ float vmax(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
// Colour divided by a scalar with the reference's rule that a zero divisor
// gives black (color.h:42-48): a lobe drawn with probability 0 (a uniform
// draw that rounds to 1.0 picks transmission on a material without it)
// zeroes the throughput instead of producing 0 * inf.
__device__ __forceinline__ V3 cdiv(V3 a, float s) { return s == 0.0f ? v3(0.0f, 0.0f, 0.0f) : a * frcp(s); }
__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// ------------------------------------------------------------ camera
// camara.h:42-59 in FP32: the plane point minus the eye, normalised.
__device__ __forceinline__ V3 camera_dir(const tfg_camera& c, int x, int y, int w, int h, float ox, float oy) {
    const float u = ((float)x + ox) / (float)w;
    const float v = ((float)y + oy) / (float)h;
    const float px = (2.0f * u - 1.0f) * (c.plane_w * 0.5f);
    const float py = (1.0f - 2.0f * v) * (c.plane_h * 0.5f);
    const V3 d = ld3(c.front) * c.focal + ld3(c.left) * px + ld3(c.up) * py;
    return normalize(d);
}

// ------------------------------------------------------------ frames
// render.h:396-413
struct Frame {
    V3 t, b, n;
};
__device__ __forceinline__ Frame make_frame(V3 n) {
    V3 t = fabsf(n.x) > fabsf(n.y) ? v3(n.z, 0.0f, -n.x) : v3(0.0f, -n.z, n.y);
    t = normalize(t);
    V3 b = normalize(cross(n, t));
    return Frame{t, b, n};
}
__device__ __forceinline__ V3 to_world(const Frame& f, V3 l) { return f.t * l.x + f.b * l.y + f.n * l.z; }
__device__ __forceinline__ V3 to_local(const Frame& f, V3 w) { return v3(dot(w, f.t), dot(w, f.b), dot(w, f.n)); }

// render.h:379-393: phi = 2 pi r1, theta = acos(sqrt(1 - r2)).
__device__ __forceinline__ V3 cosine_sample(float r1, float r2) {
    const float ct = fsqrt(fmaxf(0.0f, 1.0f - r2));
    const float st = fsqrt(fmaxf(0.0f, r2));
    float sp, cp;
    sincos_rev(r1, sp, cp);   // phi = 2 pi r1
    return v3(st * cp, st * sp, ct);
}

// ------------------------------------------------------------ GGX
__device__ __forceinline__ float ggx_alpha(float roughness) { return fmaxf(roughness * roughness, 1e-3f); }
__device__ __forceinline__ float ggx_D(float a2, float nh) {
    const float d = nh * nh * (a2 - 1.0f) + 1.0f;
    return fdiv(a2, kPi * d * d);
}
__device__ __forceinline__ float ggx_G1(float a2, float nv) {
    nv = fmaxf(nv, 0.0f);
    return fdiv(2.0f * nv, nv + fsqrt(a2 + (1.0f - a2) * nv * nv));
}
__device__ __forceinline__ V3 schlick(V3 f0, float c) {
    const float m = 1.0f - fminf(fmaxf(c, 0.0f), 1.0f);
    const float m5 = m * m * m * m * m;
    return f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * m5;
}
// BRDF value f(wo, wi) for local directions (z = normal), both above.
__device__ __forceinline__ V3 ggx_eval(V3 ks, float alpha, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    const V3 h = normalize(wo + wi);
    const float a2 = alpha * alpha;
    const float D = ggx_D(a2, h.z);
    const float G = ggx_G1(a2, wo.z) * ggx_G1(a2, wi.z);
    const V3 F = schlick(ks, dot(wi, h));
    return F * fdiv(D * G, 4.0f * wo.z * wi.z);
}
// VNDF sample of a microfacet normal (Heitz 2018, listing 1).
__device__ __forceinline__ V3 ggx_sample_vndf(V3 wo, float alpha, float u1, float u2) {
    const V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) * frcp(fsqrt(lensq)) : v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(vh, t1);
    const float r = fsqrt(u1);
    float sp, cp;
    sincos_rev(u2, sp, cp);   // phi = 2 pi u2
    const float p1 = r * cp;
    const float s = 0.5f * (1.0f + vh.z);
    const float p2 = (1.0f - s) * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1)) + s * (r * sp);
    const V3 nh = t1 * p1 + t2 * p2 + vh * fsqrt(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)));
}

// ------------------------------------------------------------ analytic
// primitiva.h:137-187 in FP32: returns t (> 1e-6) or a negative value.
__device__ __forceinline__ float analytic_t(const tfg_analytic& q, V3 o, V3 d) {
    if (q.type == TFG_AN_SPHERE) {
        const V3 oc = o - ld3(q.a);
        const float a = dot(d, d);
        const float b = 2.0f * dot(oc, d);
        const float c = dot(oc, oc) - q.s * q.s;
        const float disc = b * b - 4.0f * a * c;
        if (disc < 0.0f) return -1.0f;
        const float sq = sqrtf(disc);
        const float t1 = (-b - sq) / (2.0f * a);
        const float t2 = (-b + sq) / (2.0f * a);
        if (t1 > 1e-6f && t1 < t2) return t1;
        if (t2 > 1e-6f) return t2;
        return -1.0f;
    }
    if (q.type == TFG_AN_PLANE) {
        const V3 n = ld3(q.a);
        const float denom = dot(n, d);
        if (fabsf(denom) > 1e-6f) {
            const float t = dot(n * q.s - o, n) / denom;
            return t > 1e-6f ? t : -1.0f;
        }
        return -1.0f;
    }
    // Moller-Trumbore, primitiva.h:170-187
    const V3 v0 = ld3(q.a), e1 = ld3(q.b) - v0, e2 = ld3(q.c) - v0;
    const V3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (fabsf(det) < 1e-6f) return -1.0f;
    const float inv = 1.0f / det;
    const V3 tv = o - v0;
    const float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const V3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    const float t = dot(e2, qv) * inv;
    return t > 1e-6f ? t : -1.0f;
}

// Surface normal of an analytic hit: sphere outward (primitiva.h:190-192),
// plane as stored (primitiva.h:194-196), triangle geometric normal turned
// toward the incoming ray.
__device__ __forceinline__ V3 analytic_normal(const tfg_analytic& q, V3 p, V3 d) {
    if (q.type == TFG_AN_SPHERE) return normalize(p - ld3(q.a));
    if (q.type == TFG_AN_PLANE) return ld3(q.a);
    V3 n = normalize(cross(ld3(q.b) - ld3(q.a), ld3(q.c) - ld3(q.a)));
    if (dot(n, d) > 0.0f) n = -n;
    return n;
}

// ------------------------------------------------------------ materials
struct Mat {
    V3 kd, ks, kt, le;
    float rough, ior;
};
__device__ __forceinline__ Mat load_mat(const tfg_material* __restrict__ m) {
    const float4* p = reinterpret_cast<const float4*>(m);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    return Mat{v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), v3(d.x, d.y, d.z), a.w, b.w};
}
// Lobe classification.  "diffuse": max(kd) > eps (render.h:80); "delta":
// a specular lobe with roughness 0 (the reference's mirror) or a
// transmission lobe, which NEE cannot evaluate and the cache skips.
__device__ __forceinline__ bool mat_diffuse(const Mat& m) { return vmax(m.kd) > kEps; }
__device__ __forceinline__ bool mat_glossy(const Mat& m) { return vmax(m.ks) > kEps && m.rough > 0.0f; }
__device__ __forceinline__ bool mat_delta(const Mat& m) {
    return (vmax(m.ks) > kEps && m.rough <= 0.0f) || vmax(m.kt) > kEps;
}

// Shading normal: barycentric vertex normals when the record has them
// (escenario.h:123-142), else the geometric normal turned toward the
// incoming ray.  ng must be the unnormalised geometric normal.
__device__ __forceinline__ V3 shading_normal(const tfg_tri_shading& s, V3 ng, float u, float v, V3 dir) {
    if (s.flags & 1u) {
        const float b0 = 1.0f - u - v;
        const V3 n = ld3(s.n0) * b0 + ld3(s.n1) * u + ld3(s.n2) * v;
        const float l = fsqrt(dot(n, n));
        if (l > 1e-6f) return n * frcp(l);
    }
    V3 n = normalize(ng);
    if (dot(n, dir) > 0.0f) n = -n;
    return n;
}

// Direction and distance to a point light, and whether it lies above the
// surface (the only test that drops a light before its shadow ray).
__device__ __forceinline__ bool light_facing(const tfg_point_light& L, V3 ps, V3 n, V3& dir, float& dist, float& c) {
    V3 d = ld3(L.pos) - ps;
    dist = fsqrt(dot(d, d));
    dir = d * frcp(dist);
    c = dot(n, dir);
    return c > 0.0f;
}

// Point-light contribution at shadow point ps (escenario.h:144-184) for one
// light, without the visibility term; returns false if the light is below
// the surface.  wo in world space, for the GGX term.
__device__ __forceinline__ bool light_term(const tfg_point_light& L, V3 ps, V3 n, V3 wo, const Mat& m, V3& contrib,
                                           V3& dir, float& dist) {
    float c;
    if (!light_facing(L, ps, n, dir, dist, c)) return false;
    V3 f = m.kd * (1.0f / kPi);   // constant
    if (mat_glossy(m)) {
        const Frame fr = make_frame(n);
        f = f + ggx_eval(m.ks, ggx_alpha(m.rough), to_local(fr, wo), to_local(fr, dir));
    }
    const float dc = fmaxf(dist, 1.0f);
    contrib = ld3(L.intensity) * frcp(dc * dc) * f * c;
    return true;
}

}  // namespace tfg
