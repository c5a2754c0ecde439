// oracle.cpp - CPU restatement of the reference hot path (test
// infrastructure; see oracle.h).  Compiled with -ffp-contract=off so every
// FP32 operation rounds exactly where it is written, like the HIP kernels
// built with the same flag.
//
// Pinning of each part (tests/ hold the checks):
//   RNG               closed-form uint32 recurrences cross-checked by an
//                     independent NumPy restatement and by hand-computed
//                     known answers (tests/test_oracle_rng_camera.py).
//   camera            invariants: unit directions, centre pixel -> front,
//                     corner rays span the 45 deg field of view.
//   watertight test   closed forms: axis-aligned hits at known t/u/v,
//                     parallel rays miss, shared-edge rays hit exactly one
//                     side's set, agreement with FP64 Moller-Trumbore.
//   CWBVH traversal   equals brute force over all triangles (bit-exact ids,
//                     t, u, v) on random meshes; any-hit equals
//                     "closest-hit t < tmax"; orc_cwbvh_validate proves
//                     decoded boxes contain their triangles.
//   BVH2 + MT         equals FP64-vector brute force MT; agrees with the
//                     watertight path except on edge/tie rays.
//   conditioning      orc_render_pixels_perturbed at delta 0 equals the
//   probe             integrator bit for bit; well-conditioned pixels of a
//                     closed-form scene stay within tolerance under it; a
//                     planted systematic error is not classed as
//                     ill-conditioned (tests/test_oracle_shading_cpu.py).
//   shading           cosine draw (exact map, moments 2/3 and 1/2); GGX D
//                     normalised, Smith G1 = its tan form, visible normals
//                     integrate to one, the visible-normal draw matches its
//                     density's moments, E[G1(wi)] = the BRDF's albedo,
//                     reciprocity; analytic sphere / plane distances at
//                     known points (tests/test_oracle_shading_cpu.py).
//   integrator        scenes with closed-form pixels
//                     (tests/closed_form_scenes.py): furnace rho * Le,
//                     mixed lobes kd / pd * Le or ks / ps * Le, point light
//                     kd / pi * I * cos / max(d, 1)^2 with a shadow, glass
//                     lens decided by Snell's law in angle form.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------- rng.h
// rng.h:12-17
inline uint32_t pcg32(uint32_t& s) {
    s = s * 747796405u + 2891336453u;
    const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (word >> 22u) ^ word;
}
// rng.h:20-22 (u32 -> f32 conversion rounds to nearest, so 1.0f is reachable)
inline float pcg32_float(uint32_t& s) { return (float)pcg32(s) * (1.0f / 4294967296.0f); }
// rng.h:25-35
inline uint32_t seed_of(int px, int py, int frame) {
    uint32_t s = (uint32_t)(py << 16) | (uint32_t)(px & 0xFFFF);
    s ^= (uint32_t)frame * 719393u;
    s = (s ^ 61u) ^ (s >> 16u);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15u);
    return s == 0u ? 1u : s;
}

// --------------------------------------------------- ray precomputation
struct RayPre {
    float O[3], D[3], rD[3];
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    uint32_t octinv;
};

inline float safe_rcp(float d) {
    if (std::fabs(d) < 1e-30f) d = std::copysign(1e-30f, d);
    return 1.0f / d;
}

RayPre make_ray(const tfg_ray& r) {
    RayPre p;
    p.O[0] = r.ox; p.O[1] = r.oy; p.O[2] = r.oz;
    p.D[0] = r.dx; p.D[1] = r.dy; p.D[2] = r.dz;
    for (int a = 0; a < 3; ++a) p.rD[a] = safe_rcp(p.D[a]);
    p.tmin = r.tmin;
    p.tmax = r.tmax;
    const float ax = std::fabs(p.D[0]), ay = std::fabs(p.D[1]), az = std::fabs(p.D[2]);
    p.kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    p.kx = (p.kz + 1) % 3;
    p.ky = (p.kx + 1) % 3;
    if (p.D[p.kz] < 0.0f) std::swap(p.kx, p.ky);
    // shear constants from the dominant axis's reciprocal (the same values
    // for every triangle the ray meets, which is what watertightness needs)
    p.Sz = p.rD[p.kz];
    p.Sx = p.D[p.kx] * p.Sz;
    p.Sy = p.D[p.ky] * p.Sz;
    p.octinv = 7u - ((p.rD[0] < 0.0f ? 4u : 0u) | (p.rD[1] < 0.0f ? 2u : 0u) | (p.rD[2] < 0.0f ? 1u : 0u));
    return p;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), FP32 with
// the paper's FP64 recomputation of zero edge functions.  Returns t, and the
// barycentric weights of v1 (u) and v2 (v) as the reference's (u, v).
bool watertight(const float* v0, const float* v1, const float* v2, const RayPre& r, float& t, float& u,
                float& v) {
    float A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = v0[a] - r.O[a];
        B[a] = v1[a] - r.O[a];
        C[a] = v2[a] - r.O[a];
    }
    const float Ax = A[r.kx] - r.Sx * A[r.kz];
    const float Ay = A[r.ky] - r.Sy * A[r.kz];
    const float Bx = B[r.kx] - r.Sx * B[r.kz];
    const float By = B[r.ky] - r.Sy * B[r.kz];
    const float Cx = C[r.kx] - r.Sx * C[r.kz];
    const float Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A[r.kz];
    const float Bz = r.Sz * B[r.kz];
    const float Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    u = V / det;
    v = W / det;
    return true;
}

// Affine unit-triangle test, traverse_cwbvh.cuh:189-209.
bool affine(const float* T, const RayPre& r, float tcur, float& t, float& u, float& v) {
    const float* T0 = T;
    const float* T1 = T + 4;
    const float* T2 = T + 8;
    const float transS = T2[0] * r.O[0] + T2[1] * r.O[1] + T2[2] * r.O[2] + T2[3];
    const float transD = T2[0] * r.D[0] + T2[1] * r.D[1] + T2[2] * r.D[2];
    const float d = -transS / transD;
    if (!(d > r.tmin) || d > tcur) return false;
    const float Ix = r.O[0] + d * r.D[0];
    const float Iy = r.O[1] + d * r.D[1];
    const float Iz = r.O[2] + d * r.D[2];
    const float uu = T0[0] * Ix + T0[1] * Iy + T0[2] * Iz + T0[3];
    const float vv = T1[0] * Ix + T1[1] * Iy + T1[2] * Iz + T1[3];
    if (uu >= 0.0f && vv >= 0.0f && uu + vv < 1.0f) {
        t = d;
        u = uu;
        v = vv;
        return true;
    }
    return false;
}

constexpr float kBoxSlack = 1.00000095367431640625f;   // 1 + 2^-20

inline uint32_t byte_of(uint32_t w, int j) { return (w >> (8 * j)) & 0xFFu; }

// Decodes one 80 B node (traverse_cwbvh.cuh:75-178) and returns the 32-bit
// hit mask: bits 24..31 = internal children (slot ^ octinv), bits 0..23 =
// triangles of hit leaf children.
uint32_t node_hitmask(const float* n, const RayPre& r, float tcur) {
    const uint32_t w = f2u(n[3]);
    const int ex = (int8_t)(w & 0xFF), ey = (int8_t)((w >> 8) & 0xFF), ez = (int8_t)((w >> 16) & 0xFF);
    const float idirx = u2f((uint32_t)(ex + 127) << 23) * r.rD[0];
    const float idiry = u2f((uint32_t)(ey + 127) << 23) * r.rD[1];
    const float idirz = u2f((uint32_t)(ez + 127) << 23) * r.rD[2];
    const float origx = (n[0] - r.O[0]) * r.rD[0];
    const float origy = (n[1] - r.O[1]) * r.rD[1];
    const float origz = (n[2] - r.O[2]) * r.rD[2];
    uint32_t hitmask = 0;
    for (int i = 0; i < 8; ++i) {
        const int h = i >> 2, j = i & 3;
        const uint32_t meta = byte_of(f2u(n[6 + h]), j);
        const float qlx = (float)byte_of(f2u(n[8 + h]), j);
        const float qly = (float)byte_of(f2u(n[10 + h]), j);
        const float qlz = (float)byte_of(f2u(n[12 + h]), j);
        const float qhx = (float)byte_of(f2u(n[14 + h]), j);
        const float qhy = (float)byte_of(f2u(n[16 + h]), j);
        const float qhz = (float)byte_of(f2u(n[18 + h]), j);
        const float lox = r.rD[0] < 0.0f ? qhx : qlx, hix = r.rD[0] < 0.0f ? qlx : qhx;
        const float loy = r.rD[1] < 0.0f ? qhy : qly, hiy = r.rD[1] < 0.0f ? qly : qhy;
        const float loz = r.rD[2] < 0.0f ? qhz : qlz, hiz = r.rD[2] < 0.0f ? qlz : qhz;
        const float tminx = std::fma(lox, idirx, origx), tmaxx = std::fma(hix, idirx, origx);
        const float tminy = std::fma(loy, idiry, origy), tmaxy = std::fma(hiy, idiry, origy);
        const float tminz = std::fma(loz, idirz, origz), tmaxz = std::fma(hiz, idirz, origz);
        const float tnear = std::fmax(std::fmax(std::fmax(tminx, tminy), tminz), r.tmin);
        const float tfar = std::fmin(std::fmin(std::fmin(tmaxx, tmaxy), tmaxz), tcur);
        if (!(tnear <= tfar * kBoxSlack)) continue;
        const uint32_t is_inner = (meta & (meta << 1)) & 0x10u;
        const uint32_t bit_index = (is_inner ? (meta ^ r.octinv) : meta) & 0x1Fu;
        const uint32_t child_bits = (meta >> 5) & 7u;
        hitmask |= child_bits << bit_index;
    }
    return hitmask;
}

inline int highest_bit(uint32_t x) { return 31 - __builtin_clz(x); }

// traverse_cwbvh.cuh:51-257 control flow: node groups popped highest bit
// first, leaf triangles of the decoded node tested before descending.
// any_hit: return at the first accepted triangle.
template <bool kAnyHit>
bool traverse(const float* nodes, const float* tris, uint32_t format, const RayPre& r, tfg_hit& hit,
              uint32_t* counts) {
    uint32_t stack_x[64], stack_y[64];
    int sp = 0;
    uint32_t ng_x = 0, ng_y = 0x80000000u;
    float tcur = r.tmax;
    uint32_t best_addr = 0xFFFFFFFFu, best_prim = 0xFFFFFFFFu;
    float bu = 0.f, bv = 0.f;
    uint32_t n_nodes = 0, n_tris = 0;
    bool found = false;
    for (;;) {
        uint32_t tg_x = 0, tg_y = 0;
        if (ng_y > 0x00FFFFFFu) {
            const uint32_t imask = ng_y;
            const int bit = highest_bit(ng_y & 0xFF000000u);
            ng_y &= ~(1u << bit);
            if (ng_y > 0x00FFFFFFu) {
                stack_x[sp] = ng_x;
                stack_y[sp] = ng_y;
                ++sp;
            }
            const uint32_t slot = ((uint32_t)bit - 24u) ^ r.octinv;
            const uint32_t rel = (uint32_t)__builtin_popcount(imask & ~(0xFFFFFFFFu << slot));
            const uint32_t ni = ng_x + rel;
            const float* n = nodes + 20 * (size_t)ni;
            ++n_nodes;
            const uint32_t hm = node_hitmask(n, r, tcur);
            ng_x = f2u(n[4]);
            ng_y = (hm & 0xFF000000u) | (f2u(n[3]) >> 24);
            tg_x = f2u(n[5]);
            tg_y = hm & 0x00FFFFFFu;
        }
        while (tg_y != 0) {
            const int b = highest_bit(tg_y);
            tg_y &= ~(1u << b);
            const uint32_t addr = tg_x + 4u * (uint32_t)b;
            const float* T = tris + 4 * (size_t)addr;
            ++n_tris;
            float t, u, v;
            bool ok;
            if (format == 0) {
                ok = watertight(T, T + 4, T + 8, r, t, u, v) && t > r.tmin && t <= tcur;
            } else {
                ok = affine(T, r, tcur, t, u, v);
            }
            if (!ok) continue;
            const uint32_t prim = f2u(T[15]);
            if (kAnyHit) {
                if (t < r.tmax) {
                    if (counts) { counts[0] = n_nodes; counts[1] = n_tris; }
                    return true;
                }
                continue;
            }
            if (t < tcur || prim < best_prim) {
                tcur = t;
                bu = u;
                bv = v;
                best_addr = addr;
                best_prim = prim;
                found = true;
            }
        }
        if (ng_y <= 0x00FFFFFFu) {
            if (sp > 0) {
                --sp;
                ng_x = stack_x[sp];
                ng_y = stack_y[sp];
            } else {
                break;
            }
        }
    }
    (void)best_addr;
    if (counts) { counts[0] = n_nodes; counts[1] = n_tris; }
    if (kAnyHit) return false;
    hit.t = tcur;
    hit.u = found ? bu : 0.f;
    hit.v = found ? bv : 0.f;
    hit.prim = found ? best_prim : 0xFFFFFFFFu;
    return found;
}

template <typename F>
void parallel_for(uint32_t n, int nthreads, F f) {
    if (nthreads <= 1 || n < 1024) {
        for (uint32_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    const uint32_t chunk = 256;
    std::atomic<uint32_t> next{0};
    for (int k = 0; k < nthreads; ++k) {
        th.emplace_back([&] {
            for (;;) {
                const uint32_t b = next.fetch_add(chunk);
                if (b >= n) break;
                const uint32_t e = std::min(n, b + chunk);
                for (uint32_t i = b; i < e; ++i) f(i);
            }
        });
    }
    for (auto& t : th) t.join();
}

// ------------------------------------------------ reference BVH2 + MT
struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Bvh2 {
    struct Node { double mn[3], mx[3]; int l = -1, r = -1, first = 0, end = 0; bool leaf = false; };
    std::vector<Node> nodes;
    std::vector<V3> v0, v1, v2;     // in BVH order
    std::vector<uint32_t> prim;     // original index per BVH slot
};

// primitiva.h:164-183 (vectors FP64, det/u/v/t stored in FP32).
bool moller_trumbore(V3 v0, V3 v1, V3 v2, V3 o, V3 d, float& t) {
    const V3 v0v1 = sub(v1, v0), v0v2 = sub(v2, v0);
    const V3 pvec = cross(d, v0v2);
    const float det = (float)dot(v0v1, pvec);
    if (std::fabs(det) < 1e-6f) return false;
    const float inv = 1.0f / det;
    const V3 tvec = sub(o, v0);
    const float u = (float)(dot(tvec, pvec) * inv);
    if (u < 0.0f || u > 1.0f) return false;
    const V3 qvec = cross(tvec, v0v1);
    const float v = (float)(dot(d, qvec) * inv);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = (float)(dot(v0v2, qvec) * inv);
    return t > 1e-6f;
}

// bvh.cpp:56-100
int bvh2_rec(Bvh2& b, const std::vector<V3>& a0, const std::vector<V3>& a1, const std::vector<V3>& a2,
             std::vector<int>& idx, int begin, int end) {
    const int me = (int)b.nodes.size();
    b.nodes.push_back({});
    Bvh2::Node nd;
    for (int k = 0; k < 3; ++k) { nd.mn[k] = 1e30; nd.mx[k] = -1e30; }
    for (int i = begin; i < end; ++i) {
        const int t = idx[i];
        const V3 vs[3] = {a0[t], a1[t], a2[t]};
        for (const V3& p : vs)
            for (int k = 0; k < 3; ++k) {
                nd.mn[k] = std::fmin(nd.mn[k], comp(p, k));
                nd.mx[k] = std::fmax(nd.mx[k], comp(p, k));
            }
    }
    const int cnt = end - begin;
    if (cnt <= 4) {
        nd.leaf = true;
        nd.first = begin;
        nd.end = end;
        b.nodes[me] = nd;
        return me;
    }
    // boundingBox.h:31-36
    const double tx = nd.mx[0] - nd.mn[0], ty = nd.mx[1] - nd.mn[1], tz = nd.mx[2] - nd.mn[2];
    const int axis = (tx > ty && tx > tz) ? 0 : (ty > tz ? 1 : 2);
    auto centroid = [&](int t) {
        const double lo = std::fmin(std::fmin(comp(a0[t], axis), comp(a1[t], axis)), comp(a2[t], axis));
        const double hi = std::fmax(std::fmax(comp(a0[t], axis), comp(a1[t], axis)), comp(a2[t], axis));
        return (lo + hi) * 0.5;
    };
    // bvh.cpp:83-88 sorts with std::sort; a stable sort fixes the order of
    // equal centroids so the oracle is deterministic.
    std::stable_sort(idx.begin() + begin, idx.begin() + end, [&](int x, int y) { return centroid(x) < centroid(y); });
    const int mid = begin + cnt / 2;
    const int l = bvh2_rec(b, a0, a1, a2, idx, begin, mid);
    const int r = bvh2_rec(b, a0, a1, a2, idx, mid, end);
    nd.l = l;
    nd.r = r;
    b.nodes[me] = nd;
    return me;
}

// boundingBox.h:50-82, except that a flat box (tMax == tMin on an axis)
// is kept: the reference's `tMax <= tMin` rejects axis-aligned leaves.
bool bvh2_box(const Bvh2::Node& n, V3 o, V3 d) {
    float tMin = -1e30f, tMax = 1e30f;
    for (int i = 0; i < 3; ++i) {
        const float invD = 1.0f / (float)comp(d, i);
        float t0 = (float)((n.mn[i] - comp(o, i)) * invD);
        float t1 = (float)((n.mx[i] - comp(o, i)) * invD);
        if (invD < 0.0f) std::swap(t0, t1);
        if (t0 > tMin) tMin = t0;
        if (t1 < tMax) tMax = t1;
        if (tMax < tMin) return false;
    }
    return true;
}

}  // namespace

// ====================================================== exported API
extern "C" {

uint32_t orc_pcg32(uint32_t* state) { return pcg32(*state); }
float orc_pcg32_float(uint32_t* state) { return pcg32_float(*state); }
uint32_t orc_seed(int px, int py, int frame) { return seed_of(px, py, frame); }

void orc_stream(int px, int py, int frame, int n, uint32_t* out_u32, float* out_f32) {
    uint32_t s = seed_of(px, py, frame);
    uint32_t s2 = s;
    for (int i = 0; i < n; ++i) {
        if (out_u32) out_u32[i] = pcg32(s);
        if (out_f32) out_f32[i] = pcg32_float(s2);
    }
}

static void norm3(double v[3]) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

// camara.h:29-38: front, up normalised; left = front x up; up = left x front.
void orc_camera_make(const double pos[3], const double front[3], const double up[3], double fov_deg, int width,
                     int height, tfg_camara* out) {
    double f[3] = {front[0], front[1], front[2]}, u[3] = {up[0], up[1], up[2]};
    norm3(f);
    norm3(u);
    double l[3] = {f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]};
    norm3(l);
    double u2[3] = {l[1] * f[2] - l[2] * f[1], l[2] * f[0] - l[0] * f[2], l[0] * f[1] - l[1] * f[0]};
    norm3(u2);
    // RenderKernel.cu:70-74 (float arithmetic on a double M_PI product)
    const float focal = 1.0f;
    const float fov_rad = (float)(fov_deg * M_PI / 180.0);
    const float alto = (float)(2.0f * focal * std::tan(fov_rad / 2.0f));
    const float ancho = alto * ((float)width / (float)height);
    out->posicion = {pos[0], pos[1], pos[2]};
    out->frente = {f[0], f[1], f[2]};
    out->arriba = {u2[0], u2[1], u2[2]};
    out->izquierda = {l[0], l[1], l[2]};
    out->ancho_plano = ancho;
    out->alto_plano = alto;
    out->distancia_focal = focal;
}

// RenderKernel.cu:68-90: FOV 45, focal 1, pos (0,-0.75,4.25), front (0,0,-1).
void orc_camera_default(int width, int height, tfg_camara* out) {
    const double pos[3] = {0.0, -0.75, 4.25}, front[3] = {0.0, 0.0, -1.0}, up[3] = {0.0, 1.0, 0.0};
    orc_camera_make(pos, front, up, 45.0, width, height, out);
}

// camara.h:42-59
void orc_camera_ray(const tfg_camara* c, int x, int y, int w, int h, float ox, float oy, double origin[3],
                    double dir[3]) {
    const float u = ((float)x + ox) / (float)w;
    const float v = ((float)y + oy) / (float)h;
    const float px = (2.0f * u - 1.0f) * (c->ancho_plano * 0.5f);
    const float py = (1.0f - 2.0f * v) * (c->alto_plano * 0.5f);
    const double f = c->distancia_focal;
    double p[3] = {c->posicion.x + c->frente.x * f + c->izquierda.x * (double)px + c->arriba.x * (double)py,
                   c->posicion.y + c->frente.y * f + c->izquierda.y * (double)px + c->arriba.y * (double)py,
                   c->posicion.z + c->frente.z * f + c->izquierda.z * (double)px + c->arriba.z * (double)py};
    double d[3] = {p[0] - c->posicion.x, p[1] - c->posicion.y, p[2] - c->posicion.z};
    norm3(d);
    origin[0] = c->posicion.x;
    origin[1] = c->posicion.y;
    origin[2] = c->posicion.z;
    dir[0] = d[0];
    dir[1] = d[1];
    dir[2] = d[2];
}

void orc_cwbvh_trace(const float* nodes, const float* tris, uint32_t format, const tfg_ray* rays, uint32_t n,
                     tfg_hit* hits, uint32_t* counts, int nthreads) {
    parallel_for(n, nthreads, [&](uint32_t i) {
        const RayPre r = make_ray(rays[i]);
        traverse<false>(nodes, tris, format, r, hits[i], counts ? counts + 2 * (size_t)i : nullptr);
    });
}

void orc_cwbvh_occluded(const float* nodes, const float* tris, uint32_t format, const tfg_ray* rays, uint32_t n,
                        uint32_t* occluded, uint32_t* counts, int nthreads) {
    parallel_for(n, nthreads, [&](uint32_t i) {
        const RayPre r = make_ray(rays[i]);
        tfg_hit h;
        occluded[i] = traverse<true>(nodes, tris, format, r, h, counts ? counts + 2 * (size_t)i : nullptr) ? 1u : 0u;
    });
}

void orc_brute_trace(const float* verts, uint32_t n_tris, const tfg_ray* rays, uint32_t n, tfg_hit* hits) {
    for (uint32_t i = 0; i < n; ++i) {
        const RayPre r = make_ray(rays[i]);
        float best = r.tmax, bu = 0.f, bv = 0.f;
        uint32_t bp = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < n_tris; ++k) {
            const float* v = verts + 9 * (size_t)k;
            float t, u, vv;
            if (!watertight(v, v + 3, v + 6, r, t, u, vv)) continue;
            if (!(t > r.tmin) || t > best) continue;
            if (t < best || k < bp) { best = t; bu = u; bv = vv; bp = k; }
        }
        hits[i].t = best;
        hits[i].u = bp == 0xFFFFFFFFu ? 0.f : bu;
        hits[i].v = bp == 0xFFFFFFFFu ? 0.f : bv;
        hits[i].prim = bp;
    }
}

int orc_tri_watertight(const float v[9], const float o[3], const float d[3], float tmin, float tmax, float* t,
                       float* u, float* v_out) {
    tfg_ray ray = {o[0], o[1], o[2], tmin, d[0], d[1], d[2], tmax};
    const RayPre r = make_ray(ray);
    float tt, uu, vv;
    if (!watertight(v, v + 3, v + 6, r, tt, uu, vv)) return 0;
    if (!(tt > tmin) || tt > tmax) return 0;
    *t = tt;
    *u = uu;
    *v_out = vv;
    return 1;
}

void* orc_bvh2_build(const tfg_primitiva* prims, uint32_t n) {
    auto* b = new Bvh2();
    std::vector<V3> a0, a1, a2;
    std::vector<uint32_t> orig;
    for (uint32_t i = 0; i < n; ++i) {
        if (prims[i].tipo != TFG_TRIANGULO) continue;
        const auto& t = prims[i].g.triangulo;
        a0.push_back({t.v0.x, t.v0.y, t.v0.z});
        a1.push_back({t.v1.x, t.v1.y, t.v1.z});
        a2.push_back({t.v2.x, t.v2.y, t.v2.z});
        orig.push_back(i);
    }
    std::vector<int> idx(a0.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    if (!idx.empty()) bvh2_rec(*b, a0, a1, a2, idx, 0, (int)idx.size());
    for (int t : idx) {
        b->v0.push_back(a0[t]);
        b->v1.push_back(a1[t]);
        b->v2.push_back(a2[t]);
        b->prim.push_back(orig[t]);
    }
    return b;
}

void orc_bvh2_free(void* bvh) { delete static_cast<Bvh2*>(bvh); }
uint32_t orc_bvh2_num_nodes(const void* bvh) { return (uint32_t) static_cast<const Bvh2*>(bvh)->nodes.size(); }

// escenario.h:187-237 stack traversal; closest hit with the smaller
// original index winning exact ties (the reference keeps the first found).
void orc_bvh2_trace(const void* bvh, const tfg_ray* rays, uint32_t n, tfg_hit* hits, int nthreads) {
    const Bvh2& b = *static_cast<const Bvh2*>(bvh);
    parallel_for(n, nthreads, [&](uint32_t i) {
        const tfg_ray& ry = rays[i];
        V3 o = {ry.ox, ry.oy, ry.oz};
        V3 d = {ry.dx, ry.dy, ry.dz};
        const double l = std::sqrt(dot(d, d));
        if (l > 0.0) d = {d.x / l, d.y / l, d.z / l};
        float best = ry.tmax;
        uint32_t bp = 0xFFFFFFFFu;
        const float tlo = std::max(1e-6f, ry.tmin);
        if (!b.nodes.empty()) {
            int stack[128];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const Bvh2::Node& nd = b.nodes[stack[--sp]];
                if (!bvh2_box(nd, o, d)) continue;
                if (nd.leaf) {
                    for (int k = nd.first; k < nd.end; ++k) {
                        float t;
                        if (!moller_trumbore(b.v0[k], b.v1[k], b.v2[k], o, d, t)) continue;
                        if (!(t > tlo) || t > best) continue;
                        if (t < best || b.prim[k] < bp) { best = t; bp = b.prim[k]; }
                    }
                } else {
                    stack[sp++] = nd.l;
                    stack[sp++] = nd.r;
                }
            }
        }
        hits[i].t = best;
        hits[i].u = 0.f;
        hits[i].v = 0.f;
        hits[i].prim = bp;
    });
}

namespace {
struct Validator {
    const float* nodes;
    uint32_t num_nodes;
    const float* tris;
    uint32_t num_tris;
    uint32_t format;
    const float* verts;
    std::vector<uint8_t> seen;
    uint32_t depth = 0;
    int err = 0;

    // Returns the prim ids below node ni (appended to out).
    void visit(uint32_t ni, uint32_t d, std::vector<uint32_t>& out) {
        if (err) return;
        if (ni >= num_nodes) { err = 1; return; }
        if (d > 64) { err = 2; return; }
        depth = std::max(depth, d);
        const float* n = nodes + 20 * (size_t)ni;
        const uint32_t w = f2u(n[3]);
        const int e[3] = {(int8_t)(w & 0xFF), (int8_t)((w >> 8) & 0xFF), (int8_t)((w >> 16) & 0xFF)};
        const uint32_t imask = w >> 24;
        const uint32_t child_base = f2u(n[4]);
        const uint32_t tri_base = f2u(n[5]);
        if (tri_base % 4 != 0) { err = 3; return; }
        uint32_t rank = 0;
        for (int s = 0; s < 8; ++s) {
            const int h = s >> 2, j = s & 3;
            const uint32_t meta = byte_of(f2u(n[6 + h]), j);
            double lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                const double sc = std::ldexp(1.0, e[a]);
                lo[a] = (double)n[a] + sc * byte_of(f2u(n[8 + 2 * a + h]), j);
                hi[a] = (double)n[a] + sc * byte_of(f2u(n[14 + 2 * a + h]), j);
            }
            std::vector<uint32_t> below;
            if (imask & (1u << s)) {
                if (meta != ((1u << 5) | (24u + s))) { err = 4; return; }
                visit(child_base + rank, d + 1, below);
                ++rank;
            } else if (meta != 0) {
                const uint32_t off = meta & 0x1F, bits = meta >> 5;
                if (bits != 1 && bits != 3 && bits != 7) { err = 5; return; }
                const uint32_t cnt = bits == 1 ? 1 : (bits == 3 ? 2 : 3);
                if (off + cnt > 24) { err = 6; return; }
                for (uint32_t k = 0; k < cnt; ++k) {
                    const uint32_t rec = tri_base / 4 + off + k;
                    if (rec >= num_tris) { err = 7; return; }
                    if (seen[rec]) { err = 8; return; }
                    seen[rec] = 1;
                    below.push_back(f2u(tris[16 * (size_t)rec + 15]));
                }
            }
            if (err) return;
            for (uint32_t p : below) {
                const float* v = verts + 9 * (size_t)p;
                for (int k = 0; k < 3; ++k)
                    for (int a = 0; a < 3; ++a)
                        if ((double)v[3 * k + a] < lo[a] || (double)v[3 * k + a] > hi[a]) { err = 9; return; }
                out.push_back(p);
            }
        }
        if (rank != (uint32_t)__builtin_popcount(imask)) err = 10;
    }
};
}  // namespace

int orc_cwbvh_validate(const float* nodes, uint32_t num_nodes, const float* tris, uint32_t num_tris, uint32_t format,
                       const float* verts, uint32_t* depth) {
    Validator v{nodes, num_nodes, tris, num_tris, format, verts, std::vector<uint8_t>(num_tris, 0)};
    std::vector<uint32_t> all;
    v.visit(0, 1, all);
    if (depth) *depth = v.depth;
    if (v.err) return v.err;
    for (uint8_t s : v.seen)
        if (!s) return 11;
    std::vector<uint8_t> prim_seen(num_tris, 0);
    for (uint32_t p : all) {
        if (p >= num_tris || prim_seen[p]) return 12;
        prim_seen[p] = 1;
    }
    (void)format;
    return 0;
}

}  // extern "C"

// =================================================== integrator (render.h)
namespace {

struct D3 { double x, y, z; };
inline D3 d3(double x, double y, double z) { return {x, y, z}; }
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 dnorm(D3 a) {   // vector3d.h:65-71
    const double l = std::sqrt(ddot(a, a));
    return l > 0.0 ? a * (1.0 / l) : a;
}
inline D3 dld(const float* p) { return {p[0], p[1], p[2]}; }

struct Col { float r, g, b; };   // color.h
inline Col operator+(Col a, Col b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
inline Col operator*(Col a, Col b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
inline Col operator-(Col a, Col b) { return {a.r - b.r, a.g - b.g, a.b - b.b}; }
inline Col operator*(Col a, float s) { return {a.r * s, a.g * s, a.b * s}; }
inline Col cdiv(Col a, float s) { return s == 0.0f ? Col{0, 0, 0} : Col{a.r / s, a.g / s, a.b / s}; }   // color.h:42-48
inline float cmax(Col c) { float m = c.r >= c.g ? c.r : c.g; return c.b > m ? c.b : m; }   // color.h:58-70
inline Col cld(const float* p) { return {p[0], p[1], p[2]}; }

const double kPiD = 3.14159265358979323846;
const float kEpsF = 1e-3f;

struct OMat {
    Col kd, ks, kt, le;
    float rough, ior;
};
OMat omat(const tfg_material& m) { return {cld(m.kd), cld(m.ks), cld(m.kt), cld(m.le), m.roughness, m.ior}; }
bool o_diffuse(const OMat& m) { return cmax(m.kd) > kEpsF; }
bool o_glossy(const OMat& m) { return cmax(m.ks) > kEpsF && m.rough > 0.0f; }

struct OFrame { D3 t, b, n; };
OFrame oframe(D3 n) {   // render.h:396-413
    D3 t = std::fabs(n.x) > std::fabs(n.y) ? d3(n.z, 0, -n.x) : d3(0, -n.z, n.y);
    t = dnorm(t);
    D3 b = dnorm(dcross(n, t));
    return {t, b, n};
}
D3 o_world(const OFrame& f, D3 l) { return f.t * l.x + f.b * l.y + f.n * l.z; }
D3 o_local(const OFrame& f, D3 w) { return {ddot(w, f.t), ddot(w, f.b), ddot(w, f.n)}; }

double ggx_alpha_d(float r) { return std::max((double)r * r, 1e-3); }
double ggx_D_d(double a2, double nh) { const double d = nh * nh * (a2 - 1.0) + 1.0; return a2 / (kPiD * d * d); }
double ggx_G1_d(double a2, double nv) { nv = std::max(nv, 0.0); return 2.0 * nv / (nv + std::sqrt(a2 + (1.0 - a2) * nv * nv)); }
Col schlick_d(Col f0, double c) {
    const double m = 1.0 - std::min(std::max(c, 0.0), 1.0);
    const float m5 = (float)(m * m * m * m * m);
    return {f0.r + (1.0f - f0.r) * m5, f0.g + (1.0f - f0.g) * m5, f0.b + (1.0f - f0.b) * m5};
}
Col ggx_eval_d(Col ks, double alpha, D3 wo, D3 wi) {
    if (wo.z <= 0.0 || wi.z <= 0.0) return {0, 0, 0};
    const D3 h = dnorm(wo + wi);
    const double a2 = alpha * alpha;
    const double v = ggx_D_d(a2, h.z) * ggx_G1_d(a2, wo.z) * ggx_G1_d(a2, wi.z) / (4.0 * wo.z * wi.z);
    return schlick_d(ks, ddot(wi, h)) * (float)v;
}
D3 ggx_vndf_d(D3 wo, double alpha, double u1, double u2) {   // Heitz 2018
    const D3 vh = dnorm(d3(alpha * wo.x, alpha * wo.y, wo.z));
    const double lensq = vh.x * vh.x + vh.y * vh.y;
    const D3 t1 = lensq > 0.0 ? d3(-vh.y, vh.x, 0.0) * (1.0 / std::sqrt(lensq)) : d3(1, 0, 0);
    const D3 t2 = dcross(vh, t1);
    const double r = std::sqrt(u1), phi = 2.0 * kPiD * u2;
    const double p1 = r * std::cos(phi);
    const double s = 0.5 * (1.0 + vh.z);
    const double p2 = (1.0 - s) * std::sqrt(std::max(0.0, 1.0 - p1 * p1)) + s * r * std::sin(phi);
    const D3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::max(0.0, 1.0 - p1 * p1 - p2 * p2));
    return dnorm(d3(alpha * nh.x, alpha * nh.y, std::max(0.0, nh.z)));
}

// render.h:383-390 cosine-weighted hemisphere draw about +z: phi = 2 pi r1,
// theta = acos(sqrt(1 - r2)), so cos(theta) = sqrt(1 - r2) has pdf 2 cos.
D3 cosine_local(double r1, double r2) {
    const double phi = 2.0 * M_PI * r1, th = std::acos(std::sqrt(1.0 - r2));
    return {std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), std::cos(th)};
}

// primitiva.h:137-187: vectors in FP64, the scalars the reference keeps in
// float rounded to float.  Returns t > 1e-6 or -1.
float o_analytic_t(const tfg_analytic& q, D3 o, D3 d) {
    if (q.type == TFG_AN_SPHERE) {
        const D3 oc = o - dld(q.a);
        const float a = (float)ddot(d, d);
        const float b = (float)(2.0f * ddot(oc, d));
        const float c = (float)(ddot(oc, oc) - (double)(q.s * q.s));
        const float disc = b * b - 4.0f * a * c;
        if (disc < 0) return -1.0f;
        const float sq = std::sqrt(disc);
        const float t1 = (-b - sq) / (2.0f * a), t2 = (-b + sq) / (2.0f * a);
        if (t1 > 1e-6f && t1 < t2) return t1;
        if (t2 > 1e-6f) return t2;
        return -1.0f;
    }
    if (q.type == TFG_AN_PLANE) {
        const D3 n = dld(q.a);
        const float denom = (float)ddot(n, d);
        if (std::fabs(denom) > 1e-6f) {
            const float t = (float)(ddot(n * (double)q.s - o, n) / denom);
            return t > 1e-6f ? t : -1.0f;
        }
        return -1.0f;
    }
    const V3 v0 = {q.a[0], q.a[1], q.a[2]}, v1 = {q.b[0], q.b[1], q.b[2]}, v2 = {q.c[0], q.c[1], q.c[2]};
    float t;
    if (!moller_trumbore(v0, v1, v2, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, t)) return -1.0f;
    return t;
}

bool o_closest(const orc_scene* sc, D3 o, D3 d, tfg_hit& h) {
    tfg_ray r = {(float)o.x, (float)o.y, (float)o.z, 0.0f, (float)d.x, (float)d.y, (float)d.z,
                 std::numeric_limits<float>::infinity()};
    const RayPre rp = make_ray(r);
    bool hit = traverse<false>(sc->nodes, sc->tris, 0, rp, h, nullptr);
    for (uint32_t k = 0; k < sc->n_analytic; ++k) {   // escenario.h:323-334
        const float ta = o_analytic_t(sc->analytic[k], o, d);
        if (ta > 0.0f && ta < h.t) {
            h.t = ta;
            h.u = h.v = 0.0f;
            h.prim = 0x80000000u | k;
            hit = true;
        }
    }
    return hit;
}
bool o_occluded(const orc_scene* sc, D3 o, D3 d, float tmax) {
    for (uint32_t k = 0; k < sc->n_analytic; ++k) {
        const float ta = o_analytic_t(sc->analytic[k], o, d);
        if (ta > 0.0f && ta < tmax) return true;
    }
    tfg_ray r = {(float)o.x, (float)o.y, (float)o.z, 0.0f, (float)d.x, (float)d.y, (float)d.z, tmax};
    const RayPre rp = make_ray(r);
    tfg_hit h;
    return traverse<true>(sc->nodes, sc->tris, 0, rp, h, nullptr);
}

// Shading normal and material of a hit (escenario.h:123-142 for mesh
// triangles, primitiva.h:190-220 for analytic primitives).
void o_surface(const orc_scene* sc, const tfg_hit& h, D3 p, D3 d, D3& n, uint32_t& mat) {
    if (h.prim & 0x80000000u) {
        const tfg_analytic& q = sc->analytic[h.prim & 0x7FFFFFFFu];
        mat = q.material;
        if (q.type == TFG_AN_SPHERE) n = dnorm(p - dld(q.a));
        else if (q.type == TFG_AN_PLANE) n = dld(q.a);
        else {
            n = dnorm(dcross(dld(q.b) - dld(q.a), dld(q.c) - dld(q.a)));
            if (ddot(n, d) > 0.0) n = -n;
        }
        return;
    }
    const tfg_tri_shading& sh = sc->shading[h.prim];
    mat = sh.material;
    if (sh.flags & 1u) {   // primitiva.h:84-101
        const float b0 = 1.0f - h.u - h.v;
        const D3 ns = dld(sh.n0) * (double)b0 + dld(sh.n1) * (double)h.u + dld(sh.n2) * (double)h.v;
        const float len = (float)std::sqrt(ddot(ns, ns));
        if (len > 1e-6f) { n = ns * (double)(1.0f / len); return; }
    }
    n = dnorm(dld(sh.ng));
    if (ddot(n, d) > 0.0) n = -n;
}

// escenario.h:144-184 with the mesh occluding (the reference's CWBVH mode
// never tests shadow rays against the mesh, escenario.h:168,241-244).
Col o_direct(const orc_scene* sc, D3 ps, D3 n, D3 wo, const OMat& m, uint64_t& nshadow) {
    Col L = {0, 0, 0};
    for (uint32_t i = 0; i < sc->n_lights; ++i) {
        const tfg_point_light& lt = sc->lights[i];
        D3 dl = dld(lt.pos) - ps;
        float dist = (float)std::sqrt(ddot(dl, dl));
        dl = dnorm(dl);
        const float c = (float)ddot(n, dl);
        if (c <= 0.0f) continue;
        ++nshadow;
        if (o_occluded(sc, ps + dl * kEpsF, dl, dist)) continue;
        Col f = cdiv(m.kd, (float)kPiD);
        if (o_glossy(m)) {
            const OFrame fr = oframe(n);
            f = f + ggx_eval_d(m.ks, ggx_alpha_d(m.rough), o_local(fr, wo), o_local(fr, dl));
        }
        dist = std::max(dist, 1.0f);
        const Col Li = cdiv(cld(lt.intensity), dist * dist);
        L = L + Li * f * c;
    }
    return L;
}

// Conditioning probe (test infrastructure): with delta > 0, before every
// closest-hit query the ray origin moves by delta * (1 + |o|) and the
// direction by delta (then renormalised), each along a vector of U[-1, 1]
// components from a SplitMix64 stream (Steele, Lea & Flood 2014) that is
// separate from the path's PCG32 draws.  A path whose result changes under
// such a displacement is ill-conditioned: rounding at the level of FP32
// arithmetic can change its discrete decisions (triangle, occlusion,
// Russian roulette).
struct Perturb {
    double delta = 0.0;
    uint64_t state = 0;
};
double splitmix_unit(uint64_t& s) {   // U[-1, 1)
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

Col o_path(const orc_scene* sc, int W, int H, int x, int y, int frame, int max_depth, int rr_depth,
           uint64_t& next, uint64_t& nshadow, Perturb* pert = nullptr) {
    uint32_t s = seed_of(x, y, frame);
    const float ox = pcg32_float(s);
    const float oy = pcg32_float(s);
    const tfg_camera& c = *sc->cam;
    // camara.h:42-59: plane point minus eye, normalised (in FP64)
    const float u = ((float)x + ox) / (float)W;
    const float v = ((float)y + oy) / (float)H;
    const float px = (2.0f * u - 1.0f) * (c.plane_w * 0.5f);
    const float py = (1.0f - 2.0f * v) * (c.plane_h * 0.5f);
    D3 o = dld(c.pos);
    D3 d = dnorm(dld(c.front) * (double)c.focal + dld(c.left) * (double)px + dld(c.up) * (double)py);
    Col color = {0, 0, 0}, camino = {1, 1, 1};
    double ior = 1.0;
    for (int depth = 0; depth < max_depth; ++depth) {
        if (pert && pert->delta > 0.0) {
            const double so = pert->delta * (1.0 + std::sqrt(ddot(o, o)));
            const double ux = splitmix_unit(pert->state), uy = splitmix_unit(pert->state), uz = splitmix_unit(pert->state);
            o = o + d3(ux, uy, uz) * so;
            const double vx = splitmix_unit(pert->state), vy = splitmix_unit(pert->state), vz = splitmix_unit(pert->state);
            d = dnorm(d + d3(vx, vy, vz) * pert->delta);
        }
        tfg_hit h;
        ++next;
        if (!o_closest(sc, o, d, h)) break;   // black background, render.h:227-239
        const D3 p = o + d * (double)h.t;
        D3 n;
        uint32_t mat_id;
        o_surface(sc, h, p, d, n, mat_id);
        const OMat m = omat(sc->materials[mat_id]);
        if (cmax(m.le) > 0.0f) { color = color + camino * m.le; break; }
        const D3 wo = -d;
        if (o_diffuse(m) || o_glossy(m)) color = color + camino * o_direct(sc, p + n * kEpsF, n, wo, m, nshadow);
        const double mkd = cmax(m.kd), mks = cmax(m.ks), mkt = cmax(m.kt);
        const double sum = mkd + mks + mkt;
        if (!(sum > 0.0)) break;
        const double pd = mkd / sum, ps = mks / sum, pt = 1.0 - pd - ps;
        const double r = pcg32_float(s);
        D3 wi;
        if (r < pd) {
            camino = cdiv(camino * m.kd, (float)pd);
            const double r1 = pcg32_float(s), r2 = pcg32_float(s);
            const D3 wl = cosine_local(r1, r2);
            if (wl.z <= kEpsF) break;
            wi = o_world(oframe(n), wl);
        } else if (r < pd + ps) {
            if (o_glossy(m)) {
                const double r1 = pcg32_float(s), r2 = pcg32_float(s);
                const OFrame fr = oframe(n);
                const D3 wol = o_local(fr, wo);
                if (wol.z <= 0.0) break;
                const double alpha = ggx_alpha_d(m.rough);
                const D3 hm = ggx_vndf_d(wol, alpha, r1, r2);
                const double oh = ddot(wol, hm);
                const D3 wil = hm * (2.0 * oh) - wol;
                if (wil.z <= 0.0) break;
                camino = camino * schlick_d(m.ks, oh) * (float)(ggx_G1_d(alpha * alpha, wil.z) / ps);
                wi = dnorm(o_world(fr, wil));
            } else {
                wi = dnorm(d - n * (2.0 * ddot(d, n)));
                camino = cdiv(camino * m.ks, (float)ps);
            }
        } else {   // render.h:342-376
            D3 N = n;
            double cosi = std::max(-1.0, std::min(1.0, ddot(d, N)));
            double etat;
            if (cosi > 0.0) { N = -N; etat = 1.0; } else { cosi = -cosi; etat = m.ior; }
            const double eta = ior / etat;
            const double k = 1.0 - eta * eta * (1.0 - cosi * cosi);
            if (k < 0.0) wi = dnorm(d - N * (2.0 * ddot(d, N)));
            else { ior = etat; wi = dnorm(d * eta + N * (eta * cosi - std::sqrt(k))); }
            camino = cdiv(camino * m.kt, (float)pt);
        }
        if (depth > rr_depth) {   // render.h:209-224
            double prob = std::max((double)camino.r, std::max((double)camino.g, (double)camino.b));
            if (prob > 0.95) prob = 0.95;
            if (pcg32_float(s) > prob) break;
            camino = cdiv(camino, (float)prob);
        }
        o = p + wi * kEpsF;
        d = wi;
    }
    return color;
}

}  // namespace

extern "C" {

// Shading functions of the integrator, exported for the closed-form tests.
void orc_cosine_local(double r1, double r2, double out[3]) {
    const D3 w = cosine_local(r1, r2);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}
double orc_ggx_D(double alpha, double nh) { return ggx_D_d(alpha * alpha, nh); }
double orc_ggx_G1(double alpha, double nv) { return ggx_G1_d(alpha * alpha, nv); }
void orc_ggx_eval(const float ks[3], double alpha, const double wo[3], const double wi[3], float out[3]) {
    const Col f = ggx_eval_d({ks[0], ks[1], ks[2]}, alpha, d3(wo[0], wo[1], wo[2]), d3(wi[0], wi[1], wi[2]));
    out[0] = f.r; out[1] = f.g; out[2] = f.b;
}
void orc_ggx_vndf(const double wo[3], double alpha, double u1, double u2, double out[3]) {
    const D3 h = ggx_vndf_d(d3(wo[0], wo[1], wo[2]), alpha, u1, u2);
    out[0] = h.x; out[1] = h.y; out[2] = h.z;
}
float orc_analytic_t(const tfg_analytic* q, const double o[3], const double d[3]) {
    return o_analytic_t(*q, d3(o[0], o[1], o[2]), d3(d[0], d[1], d[2]));
}

void orc_render_rows(const orc_scene* sc, int width, int height, int y0, int y1, int frame0, int spp, int max_depth,
                     int rr_depth, double* accum, uint64_t* rays_out, int nthreads) {
    const uint32_t n = (uint32_t)width * (uint32_t)(y1 - y0);
    std::atomic<uint64_t> ext{0}, shd{0};
    parallel_for(n, nthreads, [&](uint32_t i) {
        const int x = (int)(i % (uint32_t)width), y = y0 + (int)(i / (uint32_t)width);
        uint64_t e = 0, sh = 0;
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < spp; ++k) {
            const Col c = o_path(sc, width, height, x, y, frame0 + k, max_depth, rr_depth, e, sh);
            acc[0] += c.r;
            acc[1] += c.g;
            acc[2] += c.b;
        }
        accum[3 * (size_t)i] += acc[0];
        accum[3 * (size_t)i + 1] += acc[1];
        accum[3 * (size_t)i + 2] += acc[2];
        ext += e;
        shd += sh;
    });
    if (rays_out) { rays_out[0] = ext.load(); rays_out[1] = shd.load(); }
}

void orc_render(const orc_scene* sc, int width, int height, int frame0, int spp, int max_depth, int rr_depth,
                double* accum, uint64_t* rays_out, int nthreads) {
    orc_render_rows(sc, width, height, 0, height, frame0, spp, max_depth, rr_depth, accum, rays_out, nthreads);
}

void orc_render_pixel(const orc_scene* sc, int width, int height, int x, int y, int frame, int max_depth,
                      int rr_depth, float* out_rgb) {
    uint64_t e = 0, sh = 0;
    const Col c = o_path(sc, width, height, x, y, frame, max_depth, rr_depth, e, sh);
    out_rgb[0] = c.r;
    out_rgb[1] = c.g;
    out_rgb[2] = c.b;
}

// The paths of pixels xy[2i], xy[2i + 1] of one frame under the conditioning
// probe of magnitude delta (stream seed: pseed and i).
void orc_render_pixels_perturbed(const orc_scene* sc, int width, int height, const int32_t* xy, uint32_t n, int frame,
                                 int max_depth, int rr_depth, double delta, uint32_t pseed, float* out_rgb,
                                 int nthreads) {
    parallel_for(n, nthreads, [&](uint32_t i) {
        uint64_t e = 0, sh = 0;
        Perturb p;
        p.delta = delta;
        p.state = ((uint64_t)pseed << 32) ^ (uint64_t)i;
        const Col c = o_path(sc, width, height, xy[2 * i], xy[2 * i + 1], frame, max_depth, rr_depth, e, sh, &p);
        out_rgb[3 * (size_t)i] = c.r;
        out_rgb[3 * (size_t)i + 1] = c.g;
        out_rgb[3 * (size_t)i + 2] = c.b;
    });
}

}  // extern "C"

// ======================================== NRC capture (render.h:33-262 cache branch)
namespace {

struct NrcOut {
    float pixel[3];
    uint32_t flags;        // bit0 training, bit1 head, bit2 query, bit3 tail set
    float head[16];        // head (training) or query vertex
    float tail[16];
    float lsuf[3], tsuf[3], norm[3];
};

void pack_vtx(float* dst, D3 p, D3 dm, D3 nm, const OMat& m) {
    dst[0] = (float)p.x; dst[1] = (float)p.y; dst[2] = (float)p.z; dst[3] = m.rough;
    dst[4] = (float)dm.x; dst[5] = (float)dm.y; dst[6] = (float)dm.z; dst[7] = m.kd.r;
    dst[8] = (float)nm.x; dst[9] = (float)nm.y; dst[10] = (float)nm.z; dst[11] = m.kd.g;
    dst[12] = m.kd.b; dst[13] = m.ks.r; dst[14] = m.ks.g; dst[15] = m.ks.b;
}

void o_path_nrc(const orc_scene* sc, int W, int H, int x, int y, int frame, int max_depth, int rr_depth, int mode,
                float train_frac, NrcOut& R) {
    std::memset(&R, 0, sizeof(R));
    const bool infer = mode & 1, train = mode & 2;
    uint32_t s = seed_of(x, y, frame);
    const float ox = pcg32_float(s);
    const float oy = pcg32_float(s);
    const tfg_camera& c = *sc->cam;
    const float u = ((float)x + ox) / (float)W;
    const float v = ((float)y + oy) / (float)H;
    const float px = (2.0f * u - 1.0f) * (c.plane_w * 0.5f);
    const float py = (1.0f - 2.0f * v) * (c.plane_h * 0.5f);
    D3 o = dld(c.pos);
    D3 d = dnorm(dld(c.front) * (double)c.focal + dld(c.left) * (double)px + dld(c.up) * (double)py);
    // render.h:443 and :65
    const bool training = train ? (infer ? (pcg32_float(s) < train_frac) : true) : false;
    const bool unbiased = training && (pcg32_float(s) < 0.0625f);
    Col color = {0, 0, 0}, camino = {1, 1, 1}, pre = {0, 0, 0}, tthr = {0, 0, 0};
    double ior = 1.0;
    float a0 = 0.f, acur = 0.f, asuf = 0.f, apre = 0.f, pdf_last = 1.0f;
    bool head = false, query = false;
    if (training) R.flags |= 1u;
    const float kPiF = 3.14159265358979323846f;
    for (int depth = 0; depth < max_depth; ++depth) {
        tfg_hit h;
        if (!o_closest(sc, o, d, h)) break;
        const D3 p = o + d * (double)h.t;
        D3 n;
        uint32_t mat_id;
        o_surface(sc, h, p, d, n, mat_id);
        const OMat m = omat(sc->materials[mat_id]);
        const float cos_theta = (float)std::fabs(ddot(n, -d));
        const bool diffuse = o_diffuse(m), glossy = o_glossy(m);
        const bool delta = (cmax(m.ks) > kEpsF && m.rough <= 0.0f) || cmax(m.kt) > kEpsF;
        // render.h:86-102
        const float dist_sq = h.t * h.t;
        if (depth == 0) { a0 = dist_sq / (4.0f * kPiF * std::max(cos_theta, 1e-4f)); acur = 0.f; }
        else {
            const float spread = dist_sq / (std::max(pdf_last, 1e-4f) * std::max(cos_theta, 1e-4f));
            if (!head) acur += spread; else asuf += spread;
        }
        const bool foot_ok = acur >= a0 * 0.01f;
        if (cmax(m.le) > 0.0f) {   // render.h:104-123
            color = color + camino * m.le;
            if (head) {
                OMat z = m; z.kd = {0, 0, 0}; z.ks = {0, 0, 0};
                pack_vtx(R.tail, p, -d, n, z);
                R.flags |= 8u;
            }
            break;
        }
        const D3 wo = -d;
        uint64_t nsh = 0;
        if (diffuse || glossy) {
            const Col nee = camino * o_direct(sc, p + n * kEpsF, n, wo, m, nsh);
            color = color + nee;
        }
        const D3 dm = {wo.x * 0.5 + 0.5, wo.y * 0.5 + 0.5, wo.z * 0.5 + 0.5};
        const D3 nm = {n.x * 0.5 + 0.5, n.y * 0.5 + 0.5, n.z * 0.5 + 0.5};
        if (head) {   // render.h:134-161
            if (!unbiased && asuf >= apre) {
                pack_vtx(R.tail, p, dm, nm, m);
                const Col rt = m.kd + m.ks + m.kt;
                const Col ts = camino * rt;
                R.tsuf[0] = ts.r; R.tsuf[1] = ts.g; R.tsuf[2] = ts.b;
                R.flags |= 8u;
                break;
            }
        }
        if (diffuse && !delta && foot_ok) {   // render.h:166-202
            Col refl = m.kd + m.ks + m.kt;
            refl = {std::max(refl.r, 1e-3f), std::max(refl.g, 1e-3f), std::max(refl.b, 1e-3f)};
            if (!training && infer) {
                pack_vtx(R.head, p, dm, nm, m);
                tthr = camino * refl;
                query = true;
                break;
            } else if (training && !head) {
                pack_vtx(R.head, p, dm, nm, m);
                tthr = camino * refl;
                pre = color;
                apre = acur;
                head = true;
            }
        }
        const double mkd = cmax(m.kd), mks = cmax(m.ks), mkt = cmax(m.kt);
        const double sum = mkd + mks + mkt;
        if (!(sum > 0.0)) break;
        const double pd = mkd / sum, ps = mks / sum, pt = 1.0 - pd - ps;
        const double r = pcg32_float(s);
        D3 wi;
        bool dead = false;
        if (r < pd) {
            camino = cdiv(camino * m.kd, (float)pd);
            const double r1 = pcg32_float(s), r2 = pcg32_float(s);
            const D3 wl = cosine_local(r1, r2);
            if (wl.z <= kEpsF) dead = true;
            else {
                wi = o_world(oframe(n), wl);
                pdf_last = (float)(std::max(std::fabs(ddot(wi, n)), 1e-4) / kPiD);
            }
        } else if (r < pd + ps) {
            if (glossy) {
                const double r1 = pcg32_float(s), r2 = pcg32_float(s);
                const OFrame fr = oframe(n);
                const D3 wol = o_local(fr, wo);
                if (wol.z <= 0.0) dead = true;
                else {
                    const double alpha = ggx_alpha_d(m.rough);
                    const D3 hm = ggx_vndf_d(wol, alpha, r1, r2);
                    const double oh = ddot(wol, hm);
                    const D3 wil = hm * (2.0 * oh) - wol;
                    if (wil.z <= 0.0) dead = true;
                    else {
                        const double a2 = alpha * alpha;
                        camino = camino * schlick_d(m.ks, oh) * (float)(ggx_G1_d(a2, wil.z) / ps);
                        wi = dnorm(o_world(fr, wil));
                        pdf_last = (float)(ggx_D_d(a2, hm.z) * ggx_G1_d(a2, wol.z) / (4.0 * wol.z));
                    }
                }
            } else {
                wi = dnorm(d - n * (2.0 * ddot(d, n)));
                camino = cdiv(camino * m.ks, (float)ps);
                pdf_last = 1e6f;
            }
        } else {
            D3 N = n;
            double cosi = std::max(-1.0, std::min(1.0, ddot(d, N)));
            double etat;
            if (cosi > 0.0) { N = -N; etat = 1.0; } else { cosi = -cosi; etat = m.ior; }
            const double eta = ior / etat;
            const double k = 1.0 - eta * eta * (1.0 - cosi * cosi);
            if (k < 0.0) wi = dnorm(d - N * (2.0 * ddot(d, N)));
            else { ior = etat; wi = dnorm(d * eta + N * (eta * cosi - std::sqrt(k))); }
            camino = cdiv(camino * m.kt, (float)pt);
            pdf_last = 1e6f;
        }
        if (dead) break;
        if (depth > rr_depth) {
            double prob = std::max((double)camino.r, std::max((double)camino.g, (double)camino.b));
            if (prob > 0.95) prob = 0.95;
            if (pcg32_float(s) > prob) break;
            camino = cdiv(camino, (float)prob);
        }
        o = p + wi * kEpsF;
        d = wi;
    }
    Col pix = color;
    if (training && head) {
        pix = pre;
        R.flags |= 2u;
        const Col ls = color - pre;
        R.lsuf[0] = ls.r; R.lsuf[1] = ls.g; R.lsuf[2] = ls.b;
        R.norm[0] = tthr.r; R.norm[1] = tthr.g; R.norm[2] = tthr.b;
    }
    if (query) R.flags |= 4u;
    if (query || (training && head)) { R.norm[0] = tthr.r; R.norm[1] = tthr.g; R.norm[2] = tthr.b; }
    R.pixel[0] = pix.r; R.pixel[1] = pix.g; R.pixel[2] = pix.b;
}

}  // namespace

extern "C" {
// Cache-mode paths of every pixel of one frame; out: W*H records of
// 64 floats = pixel(3), flags(1, as bits), head(16), tail(16), lsuf(3),
// tsuf(3), norm(3), pad.
void orc_render_nrc(const orc_scene* sc, int width, int height, int frame, int max_depth, int rr_depth, int mode,
                    float train_frac, float* out, int nthreads) {
    const uint32_t n = (uint32_t)width * (uint32_t)height;
    parallel_for(n, nthreads, [&](uint32_t i) {
        NrcOut R;
        o_path_nrc(sc, width, height, (int)(i % (uint32_t)width), (int)(i / (uint32_t)width), frame, max_depth,
                   rr_depth, mode, train_frac, R);
        float* o = out + 64 * (size_t)i;
        std::memset(o, 0, 64 * sizeof(float));
        o[0] = R.pixel[0]; o[1] = R.pixel[1]; o[2] = R.pixel[2];
        std::memcpy(&o[3], &R.flags, 4);
        std::memcpy(&o[4], R.head, 16 * 4);
        std::memcpy(&o[20], R.tail, 16 * 4);
        std::memcpy(&o[36], R.lsuf, 12);
        std::memcpy(&o[39], R.tsuf, 12);
        std::memcpy(&o[42], R.norm, 12);
    });
}
}  // extern "C"
