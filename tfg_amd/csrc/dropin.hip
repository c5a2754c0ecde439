// dropin.hip - launchKernelRenderTiny (RenderKernel.h:55-61) on the
// reference's own buffers: Camara, Primitiva (analytic list and the
// per-primIdx mesh copy), LuzPuntual, TinyBVHD_GPU, ImagenGPU, DatosMLP,
// Color and RegistroEntrenamiento.
//
// The frame itself is the wavefront tracer of pt.hip.  This unit converts
// the reference's scene records into the tracer's compact layout (cached
// per scene buffer; tfg_dropin_invalidate() forces a reconversion after the
// caller rewrites them in place) and writes the frame's outputs back in
// the reference formats:
//   image        ImagenGPU.datos[y*W+x]                  (render.h:449)
//   queries      buffer_inference[pix], buffer_throughput[pix] (0 if none)
//                                                         (render.h:452-458)
//   records      buffer_training[atomicAdd(counter) % cap] (render.h:461-464)
#include <hip/hip_runtime.h>

#include <mutex>

#include "host_common.h"
#include "pt_internal.h"
#include "wave_ops.h"

namespace tfg {
uint32_t known_tree_depth(const void* d_nodes);
uint32_t known_tree_format(const void* d_nodes);
int pt_render_export(tfg_pt* pt, int32_t frame, tfg_color* d_image, hipStream_t stream);
const PtParams& pt_params(const tfg_pt* pt);
}  // namespace tfg

using namespace tfg;

namespace {

struct DropIn {
    std::mutex mu;
    tfg_pt* pt = nullptr;
    uint32_t W = 0, H = 0, max_depth = 0, rr_depth = 0;
    // converted scene and the reference buffers it was converted from
    const void* k_bvh = nullptr;
    const void* k_mesh = nullptr;
    const void* k_prims = nullptr;
    const void* k_lights = nullptr;
    int k_n_mesh = -1, k_n_prims = -1, k_n_lights = -1;
    size_t cap_mesh = 0;
    tfg_tri_shading* shading = nullptr;
    tfg_material* mats = nullptr;   // n_mesh + n_prims entries
    tfg_point_light* lights = nullptr;
    tfg_analytic* analytic = nullptr;
    float* verts = nullptr;          // affine (tinybvh-layout) trees: 9 floats per primIdx
    size_t cap_verts = 0;
    tfg_pt_scene scene{};
};
DropIn g_drop;

__device__ __forceinline__ tfg_material mat_of(const tfg_primitiva& p) {
    tfg_material m{};
    m.kd[0] = p.difuso.r; m.kd[1] = p.difuso.g; m.kd[2] = p.difuso.b;
    m.ks[0] = p.especular.r; m.ks[1] = p.especular.g; m.ks[2] = p.especular.b;
    m.kt[0] = p.transmision.r; m.kt[1] = p.transmision.g; m.kt[2] = p.transmision.b;
    m.le[0] = p.emision.r; m.le[1] = p.emision.g; m.le[2] = p.emision.b;
    m.roughness = 0.0f;   // Primitiva has no roughness: mirror specular lobe
    m.ior = p.indice_refraccion;
    return m;
}

// Mesh Primitiva (indexed by primIdx, tinybvh_wrapper.cpp:77-79) -> shading
// record + material; smooth normals when the vertex normals differ
// (escenario.h:126-134).  verts (nullable): the FP32 vertices, rounded from
// the FP64 ones as tinybvh_wrapper.cpp:38-40 rounds them for the build.
__global__ void k_conv_mesh(const tfg_primitiva* __restrict__ prims, uint32_t n, tfg_tri_shading* sh, tfg_material* mats,
                            float* verts) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const tfg_primitiva p = prims[i];
        const auto& t = p.g.triangulo;
        tfg_tri_shading s{};
        s.n0[0] = (float)t.v0_normal.x; s.n0[1] = (float)t.v0_normal.y; s.n0[2] = (float)t.v0_normal.z;
        s.n1[0] = (float)t.v1_normal.x; s.n1[1] = (float)t.v1_normal.y; s.n1[2] = (float)t.v1_normal.z;
        s.n2[0] = (float)t.v2_normal.x; s.n2[1] = (float)t.v2_normal.y; s.n2[2] = (float)t.v2_normal.z;
        const bool smooth = t.v0_normal.x != t.v1_normal.x || t.v0_normal.y != t.v1_normal.y ||
                            t.v0_normal.z != t.v1_normal.z || t.v0_normal.x != t.v2_normal.x ||
                            t.v0_normal.y != t.v2_normal.y || t.v0_normal.z != t.v2_normal.z;
        s.flags = smooth ? 1u : 0u;
        s.material = i;
        const double e1x = t.v1.x - t.v0.x, e1y = t.v1.y - t.v0.y, e1z = t.v1.z - t.v0.z;
        const double e2x = t.v2.x - t.v0.x, e2y = t.v2.y - t.v0.y, e2z = t.v2.z - t.v0.z;
        s.ng[0] = (float)(e1y * e2z - e1z * e2y);
        s.ng[1] = (float)(e1z * e2x - e1x * e2z);
        s.ng[2] = (float)(e1x * e2y - e1y * e2x);
        sh[i] = s;
        mats[i] = mat_of(p);
        if (verts) {
            float* v = verts + 9 * (size_t)i;
            v[0] = (float)t.v0.x; v[1] = (float)t.v0.y; v[2] = (float)t.v0.z;
            v[3] = (float)t.v1.x; v[4] = (float)t.v1.y; v[5] = (float)t.v1.z;
            v[6] = (float)t.v2.x; v[7] = (float)t.v2.y; v[8] = (float)t.v2.z;
        }
    }
}

__global__ void k_conv_prims(const tfg_primitiva* __restrict__ prims, uint32_t n, uint32_t mat_base, tfg_analytic* an,
                             tfg_material* mats) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const tfg_primitiva p = prims[i];
    tfg_analytic q{};
    q.type = (uint32_t)p.tipo;
    q.material = mat_base + i;
    if (p.tipo == TFG_ESFERA) {
        q.a[0] = (float)p.g.esfera.centro.x; q.a[1] = (float)p.g.esfera.centro.y; q.a[2] = (float)p.g.esfera.centro.z;
        q.s = p.g.esfera.radio;
    } else if (p.tipo == TFG_PLANO) {
        q.a[0] = (float)p.g.plano.normal.x; q.a[1] = (float)p.g.plano.normal.y; q.a[2] = (float)p.g.plano.normal.z;
        q.s = p.g.plano.distancia;
    } else {
        const auto& t = p.g.triangulo;
        q.a[0] = (float)t.v0.x; q.a[1] = (float)t.v0.y; q.a[2] = (float)t.v0.z;
        q.b[0] = (float)t.v1.x; q.b[1] = (float)t.v1.y; q.b[2] = (float)t.v1.z;
        q.c[0] = (float)t.v2.x; q.c[1] = (float)t.v2.y; q.c[2] = (float)t.v2.z;
    }
    an[i] = q;
    mats[mat_base + i] = mat_of(p);
}

__global__ void k_conv_lights(const tfg_luz_puntual* __restrict__ l, uint32_t n, tfg_point_light* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tfg_point_light p{};
    p.pos[0] = (float)l[i].posicion.x; p.pos[1] = (float)l[i].posicion.y; p.pos[2] = (float)l[i].posicion.z;
    p.intensity[0] = l[i].intensidad.r; p.intensity[1] = l[i].intensidad.g; p.intensity[2] = l[i].intensidad.b;
    out[i] = p;
}

// Camara (FP64 vectors) -> tfg_camera.
__global__ void k_conv_camera(const tfg_camara* __restrict__ c, tfg_camera* out) {
    if (threadIdx.x != 0) return;
    const tfg_camara k = *c;
    tfg_camera o{};
    o.pos[0] = (float)k.posicion.x; o.pos[1] = (float)k.posicion.y; o.pos[2] = (float)k.posicion.z;
    o.front[0] = (float)k.frente.x; o.front[1] = (float)k.frente.y; o.front[2] = (float)k.frente.z;
    o.up[0] = (float)k.arriba.x; o.up[1] = (float)k.arriba.y; o.up[2] = (float)k.arriba.z;
    o.left[0] = (float)k.izquierda.x; o.left[1] = (float)k.izquierda.y; o.left[2] = (float)k.izquierda.z;
    o.focal = k.distancia_focal;
    o.plane_w = k.ancho_plano;
    o.plane_h = k.alto_plano;
    *out = o;
}

__device__ __forceinline__ tfg_vec3d dv(float x, float y, float z) { return tfg_vec3d{x, y, z}; }

// Vertex record (4 float4) -> DatosGeometricos.
__device__ __forceinline__ tfg_datos_geometricos geom_of(const float4* v) {
    const float4 a = v[0], b = v[1], c = v[2], d = v[3];
    tfg_datos_geometricos g;
    g.posicion = dv(a.x, a.y, a.z);
    g.direccion = dv(b.x, b.y, b.z);
    g.normal = dv(c.x, c.y, c.z);
    g.difuso = tfg_color{b.w, c.w, d.x};
    g.especular = tfg_color{d.y, d.z, d.w};
    return g;
}

__global__ __launch_bounds__(256) void k_export(PtParams P, tfg_datos_mlp* inf, tfg_color* thr_out,
                                                tfg_registro_entrenamiento* recs, uint32_t* counter, uint32_t cap) {
    for (uint32_t b = blockIdx.x * blockDim.x; b < P.n_paths; b += gridDim.x * blockDim.x) {
        const uint32_t pid = b + threadIdx.x;
        bool want_rec = false;
        uint32_t fl = 0;
        if (pid < P.n_paths && P.mode) {
            fl = P.flags[pid];
            const bool head = (fl & PF_TRAIN) && (fl & PF_HEAD);
            const bool need = head || (fl & PF_QUERY);
            if (need) {
                const tfg_datos_geometricos g = geom_of(P.head + (size_t)kVertF4 * pid);
                tfg_datos_mlp d = inf[pid];
                d.posicion = g.posicion;
                d.direccion = g.direccion;
                d.normal = g.normal;
                d.difuso = g.difuso;
                d.especular = g.especular;
                inf[pid] = d;
                const float4 t = P.tthr[pid];
                thr_out[pid] = tfg_color{t.x, t.y, t.z};
            } else {
                thr_out[pid] = tfg_color{0.f, 0.f, 0.f};
            }
            want_rec = head && counter != nullptr;
        } else if (pid < P.n_paths) {
            thr_out[pid] = tfg_color{0.f, 0.f, 0.f};
        }
        const uint32_t slot = wave_append(want_rec, counter ? counter : P.counters + CNT_TOTAL);
        if (want_rec) {
            tfg_registro_entrenamiento r;
            r.head = geom_of(P.head + (size_t)kVertF4 * pid);
            if (fl & PF_TAIL) {
                r.tail = geom_of(P.tail + (size_t)kVertF4 * pid);
                const float4 ts = P.tsuf[pid];
                r.throughput_sufijo = tfg_color{ts.x, ts.y, ts.z};
            } else {
                r.tail = tfg_datos_geometricos{};
                r.throughput_sufijo = tfg_color{0.f, 0.f, 0.f};
            }
            const float4 pr = P.pre[pid], rd = P.rad[pid], tt = P.tthr[pid];
            r.luz_acumulada_sufijo = tfg_color{rd.x - pr.x, rd.y - pr.y, rd.z - pr.z};
            r.factor_normalizacion = tfg_color{tt.x, tt.y, tt.z};
            r.valido = 1;
            recs[slot % cap] = r;
        }
    }
}

int convert_scene(DropIn& g, const tfg_primitiva* prims, int n_prims, const tfg_luz_puntual* luces, int n_luces,
                  tfg_cwbvh_gpu bvh, const tfg_primitiva* mesh, int n_mesh, hipStream_t s) {
    if (g.k_bvh == bvh.d_nodos && g.k_mesh == mesh && g.k_prims == prims && g.k_lights == luces &&
        g.k_n_mesh == n_mesh && g.k_n_prims == n_prims && g.k_n_lights == n_luces)
        return TFG_OK;
    const size_t need = (size_t)n_mesh;
    if (need > g.cap_mesh) {
        if (g.shading) (void)hipFree(g.shading);
        if (g.mats) (void)hipFree(g.mats);
        g.shading = nullptr;
        g.mats = nullptr;
        g.cap_mesh = 0;
        TFG_HIP_CHECK(hipMalloc(&g.shading, sizeof(tfg_tri_shading) * need));
        TFG_HIP_CHECK(hipMalloc(&g.mats, sizeof(tfg_material) * (need + kMaxAnalytic)));
        g.cap_mesh = need;
    }
    // a tree this library did not upload is taken as tinybvh's layout (affine
    // records, tinybvh_wrapper.cpp:3): the tracer traces it with the
    // watertight test on the Primitiva copy's vertices
    const bool affine = known_tree_format(bvh.d_nodos) != TFG_TRI_WATERTIGHT;
    if (affine && need > g.cap_verts) {
        if (g.verts) (void)hipFree(g.verts);
        g.verts = nullptr;
        g.cap_verts = 0;
        TFG_HIP_CHECK(hipMalloc(&g.verts, 9 * sizeof(float) * need));
        g.cap_verts = need;
    }
    if (!g.lights) TFG_HIP_CHECK(hipMalloc(&g.lights, sizeof(tfg_point_light) * 64));
    if (!g.analytic) TFG_HIP_CHECK(hipMalloc(&g.analytic, sizeof(tfg_analytic) * kMaxAnalytic));
    if (n_mesh)
        hipLaunchKernelGGL(k_conv_mesh, dim3(1024), dim3(256), 0, s, mesh, (uint32_t)n_mesh, g.shading, g.mats,
                           affine ? g.verts : nullptr);
    if (n_prims) hipLaunchKernelGGL(k_conv_prims, dim3(1), dim3(64), 0, s, prims, (uint32_t)n_prims, (uint32_t)n_mesh,
                                    g.analytic, g.mats);
    if (n_luces) hipLaunchKernelGGL(k_conv_lights, dim3(1), dim3(64), 0, s, luces, (uint32_t)n_luces, g.lights);
    TFG_HIP_CHECK(hipGetLastError());
    g.scene.bvh = bvh;
    g.scene.d_shading = g.shading;
    g.scene.d_materials = g.mats;
    g.scene.d_lights = g.lights;
    g.scene.n_lights = (uint32_t)n_luces;
    g.scene.d_analytic = g.analytic;
    g.scene.n_analytic = (uint32_t)n_prims;
    g.scene.tri_format = affine ? TFG_TRI_AFFINE : TFG_TRI_WATERTIGHT;
    g.scene.d_verts = affine ? g.verts : nullptr;
    g.scene.n_verts = affine ? (uint32_t)n_mesh : 0u;
    const int rc = tfg_pt_set_scene(g.pt, &g.scene);
    if (rc != TFG_OK) return rc;
    g.k_bvh = bvh.d_nodos;
    g.k_mesh = mesh;
    g.k_prims = prims;
    g.k_lights = luces;
    g.k_n_mesh = n_mesh;
    g.k_n_prims = n_prims;
    g.k_n_lights = n_luces;
    return TFG_OK;
}

}  // namespace

extern "C" {

int tfg_dropin_configure(uint32_t max_depth, uint32_t rr_depth) {
    std::lock_guard<std::mutex> lk(g_drop.mu);
    if (max_depth == 0 || max_depth > kMaxDepth) { set_error("tfg_dropin_configure: max_depth must be 1..%u", kMaxDepth); return TFG_E_INVALID; }
    g_drop.max_depth = max_depth;
    g_drop.rr_depth = rr_depth;
    if (g_drop.pt) {
        tfg_pt_destroy(g_drop.pt);
        g_drop.pt = nullptr;
        g_drop.k_bvh = nullptr;
    }
    return TFG_OK;
}

int tfg_dropin_invalidate(void) {
    std::lock_guard<std::mutex> lk(g_drop.mu);
    g_drop.k_bvh = nullptr;
    g_drop.k_n_mesh = -1;
    return TFG_OK;
}

int tfg_launch_render_tiny(const tfg_camara* d_camara, const tfg_primitiva* d_prims, int n_prims,
                           const tfg_luz_puntual* d_luces, int n_luces, const tfg_primitiva* d_prims_malla,
                           int n_prims_malla, int width, int height, int spp, int frame, tfg_cwbvh_gpu bvh,
                           const tfg_primitiva* d_prims_bvh, int n_prims_bvh, tfg_imagen_gpu imagen,
                           unsigned int* d_dev_counter, tfg_registro_entrenamiento* d_recs,
                           unsigned int* d_counter_train, int max_cap_train, tfg_datos_mlp* d_inference,
                           tfg_color* d_throughput, int usar_red_inferencia, int entrenar_red, tfg_stream stream) {
    (void)d_prims_malla;   // unused by kernelRender_tiny (escenario.h:316-366)
    (void)n_prims_malla;
    (void)spp;             // stored but unused by the reference (render.h:30-31)
    (void)d_dev_counter;   // unused by the reference
    std::lock_guard<std::mutex> lk(g_drop.mu);
    if (!d_camara || width <= 0 || height <= 0 || !imagen.datos || imagen.anchura != width || imagen.altura != height ||
        !d_throughput || (n_prims && !d_prims) || (n_luces && !d_luces) || n_prims < 0 || n_luces < 0 || n_prims_bvh < 0) {
        set_error("tfg_launch_render_tiny: bad argument");
        return TFG_E_INVALID;
    }
    if (n_prims > (int)kMaxAnalytic || n_luces > 64) {
        set_error("tfg_launch_render_tiny: at most %u analytic primitives and 64 lights", kMaxAnalytic);
        return TFG_E_LIMIT;
    }
    if (usar_red_inferencia && !d_inference) { set_error("tfg_launch_render_tiny: buffer_inference is NULL"); return TFG_E_INVALID; }
    if (entrenar_red && d_counter_train && (!d_recs || max_cap_train <= 0)) {
        set_error("tfg_launch_render_tiny: training buffers missing");
        return TFG_E_INVALID;
    }
    if (!bvh.d_nodos || n_prims_bvh == 0 || !d_prims_bvh) {
        set_error("tfg_launch_render_tiny: a CWBVH mesh is required (TinyBVH::construirBVH)");
        return TFG_E_INVALID;
    }
    DropIn& g = g_drop;
    if (!g.max_depth) {
        g.max_depth = 20;   // render.h:63
        g.rr_depth = 10;    // render.h:210
    }
    if (!g.pt || g.W != (uint32_t)width || g.H != (uint32_t)height) {
        if (g.pt) tfg_pt_destroy(g.pt);
        g.pt = nullptr;
        tfg_pt_config cfg{(uint32_t)width, (uint32_t)height, g.max_depth, g.rr_depth, 3u, 0.05f, 0u};
        const int rc = tfg_pt_create(&cfg, &g.pt);
        if (rc != TFG_OK) return rc;
        g.W = (uint32_t)width;
        g.H = (uint32_t)height;
        g.k_bvh = nullptr;
    }
    hipStream_t s = stream;
    int rc = convert_scene(g, d_prims, n_prims, d_luces, n_luces, bvh, d_prims_bvh, n_prims_bvh, s);
    if (rc != TFG_OK) return rc;
    const uint32_t mode = (usar_red_inferencia ? MODE_INFER : 0u) | (entrenar_red ? MODE_TRAIN : 0u);
    rc = tfg_pt_set_mode(g.pt, mode);
    if (rc != TFG_OK) return rc;
    const PtParams& P = pt_params(g.pt);
    hipLaunchKernelGGL(k_conv_camera, dim3(1), dim3(64), 0, s, d_camara, const_cast<tfg_camera*>(P.cam));
    TFG_HIP_CHECK(hipGetLastError());
    rc = pt_render_export(g.pt, frame, imagen.datos, s);
    if (rc != TFG_OK) return rc;
    hipLaunchKernelGGL(k_export, dim3(1024), dim3(256), 0, s, pt_params(g.pt), d_inference, d_throughput, d_recs,
                       entrenar_red ? d_counter_train : nullptr, (uint32_t)(max_cap_train > 0 ? max_cap_train : 1));
    TFG_HIP_CHECK(hipGetLastError());
    return TFG_OK;
}

}  // extern "C"
