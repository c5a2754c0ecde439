// A reference-style caller of the reference-named shims (include/tfg_ref):
// the scene, camera and buffers are the reference's own structs, and the
// frame is the sequence main.cpp:1016-1056 runs, through the reference's
// function names (TinyBVH::construirBVH, launchInicializarCamara,
// launchKernelRenderTiny, ColorMLP::inference, launchKernelComposite,
// launchKernelPrepararInferenciaTail, launchKernelCalcularTargets,
// ColorMLP::train_step, save_model / load_model).  Built by tests/cpp/Makefile
// with hipcc and linked against libtfg_hip.so; run by tests/test_shims_gpu.py,
// which checks every output file against the oracle.
//
//   shim_run <in_dir> <out_dir>
//   in:  params.txt (W H frame_plain frame_cache), prims.bin (analytic
//        Primitiva), mesh.bin (triangle Primitiva), luces.bin (LuzPuntual)
//   out: raw arrays named after the buffers of main.cpp, and scalars.txt
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// the reference's structs (vector3d.h, color.h, primitiva.h, luzPuntual.h,
// camara.h, imagen_gpu.h, tinybvh_wrapper.h:15-21, mlp_types.h)
struct Vector3d { double x, y, z; };
struct Color { float r, g, b; };
enum TipoPrimitiva { ESFERA, PLANO, TRIANGULO };
struct Primitiva {
    TipoPrimitiva tipo;
    Color emision, difuso, especular, transmision;
    float indice_refraccion;
    union {
        struct { Vector3d centro; float radio; } esfera;
        struct { Vector3d normal; float distancia; } plano;
        struct { Vector3d v0, v1, v2, v0_normal, v1_normal, v2_normal; } triangulo;
    };
};
struct LuzPuntual { Vector3d posicion; Color intensidad; };
struct Camara { Vector3d posicion, frente, arriba, izquierda; float ancho_plano, alto_plano, distancia_focal; };
struct ImagenGPU { int anchura, altura; Color* datos; };
struct TinyBVHD_GPU { float4* d_nodos; float4* d_tris; unsigned int num_nodos, num_tris; };
struct DatosMLP { Vector3d posicion, direccion, normal; Color difuso, especular, color; };
struct SceneBounds { Vector3d min, max; };
struct DatosGeometricos { Vector3d posicion, direccion, normal; Color difuso, especular; };
struct RegistroEntrenamiento {
    DatosGeometricos head, tail;
    Color luz_acumulada_sufijo, throughput_sufijo, factor_normalizacion;
    bool valido;
};

#include "../../include/tfg_ref/NrcKernel.h"
#include "../../include/tfg_ref/RenderKernel.h"
#include "../../include/tfg_ref/mlp.h"
#include "../../include/tfg_ref/tinybvh_wrapper.h"

namespace {

template <typename T>
std::vector<T> read_all(const std::string& path) {
    std::vector<T> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return v;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)n / sizeof(T));
    if (!v.empty() && std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) v.clear();
    std::fclose(f);
    return v;
}

template <typename T>
bool dump(const std::string& path, const T* d_ptr, size_t count) {
    std::vector<T> h(count);
    if (count && hipMemcpy(h.data(), d_ptr, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return false;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(h.data(), sizeof(T), count, f) == count;
    std::fclose(f);
    return ok;
}

template <typename T>
T* upload(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, (h.empty() ? 1 : h.size()) * sizeof(T)) != hipSuccess) return nullptr;
    if (!h.empty()) (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return d;
}

bool check(const char* what) {
    const hipError_t e = hipDeviceSynchronize();
    const std::string msg = tfg_last_error();
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        return false;
    }
    std::printf("%s done%s%s\n", what, msg.empty() ? "" : "; last library message: ", msg.c_str());
    std::fflush(stdout);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <in_dir> <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string in = argv[1], out = std::string(argv[2]) + "/";
    int W = 0, H = 0, frame_plain = 0, frame_cache = 0;
    {
        FILE* f = std::fopen((in + "/params.txt").c_str(), "r");
        if (!f || std::fscanf(f, "%d %d %d %d", &W, &H, &frame_plain, &frame_cache) != 4) {
            std::fprintf(stderr, "bad params.txt\n");
            return 2;
        }
        std::fclose(f);
    }
    const std::vector<Primitiva> prims = read_all<Primitiva>(in + "/prims.bin");
    const std::vector<Primitiva> mesh = read_all<Primitiva>(in + "/mesh.bin");
    const std::vector<LuzPuntual> luces = read_all<LuzPuntual>(in + "/luces.bin");
    if (mesh.empty() || luces.empty()) {
        std::fprintf(stderr, "missing scene files\n");
        return 2;
    }
    const int n = W * H;

    // init() (main.cpp:870-930): BVH, camera, scene and frame buffers
    TinyBVH bvh_modelo_tiny;
    bvh_modelo_tiny.construirBVH(mesh);
    bvh_modelo_tiny.obtenerInfo();
    Camara* d_camara = nullptr;
    (void)hipMalloc(&d_camara, sizeof(Camara));
    launchInicializarCamara(d_camara, W, H);
    Primitiva* d_primitivas = upload(prims);
    LuzPuntual* d_luces = upload(luces);
    Color *d_imagen = nullptr, *d_pred = nullptr, *d_thr = nullptr, *d_tail_pred = nullptr;
    DatosMLP *d_inf = nullptr, *d_tail_in = nullptr, *d_train = nullptr;
    RegistroEntrenamiento* d_regs = nullptr;
    unsigned int* d_counter = nullptr;
    (void)hipMalloc(&d_imagen, n * sizeof(Color));
    (void)hipMalloc(&d_pred, n * sizeof(Color));
    (void)hipMalloc(&d_thr, n * sizeof(Color));
    (void)hipMalloc(&d_tail_pred, n * sizeof(Color));
    (void)hipMalloc(&d_inf, n * sizeof(DatosMLP));
    (void)hipMalloc(&d_tail_in, n * sizeof(DatosMLP));
    (void)hipMalloc(&d_train, n * sizeof(DatosMLP));
    (void)hipMalloc(&d_regs, n * sizeof(RegistroEntrenamiento));
    (void)hipMalloc(&d_counter, sizeof(unsigned int));
    ImagenGPU imagen{W, H, d_imagen};
    const dim3 block(16, 16), grid((W + 15) / 16, (H + 15) / 16);
    SceneBounds bounds{{-6.0, -4.0, -10.0}, {6.0, 6.0, 2.0}};   // main.cpp:870-872
    ColorMLP mlp(15, 3, 16384, nullptr);
    mlp.setBounds(bounds.min, bounds.max);
    mlp.save_model(out + "model_before.bin");
    if (!check("init")) return 1;

    // a plain frame (both flags false)
    (void)hipMemset(d_imagen, 0, n * sizeof(Color));
    launchKernelRenderTiny(grid, block, d_camara, d_primitivas, (int)prims.size(), d_luces, (int)luces.size(), nullptr,
                           0, W, H, 1, frame_plain, bvh_modelo_tiny.obtenerDatosGPU(), bvh_modelo_tiny.getPrimitivasGPU(),
                           (int)bvh_modelo_tiny.getNumPrimitivas(), imagen, nullptr, nullptr, nullptr, 0, d_inf, d_thr,
                           false, false);
    if (!check("plain frame") || !dump(out + "img_plain.bin", d_imagen, n)) return 1;

    // the NRC frame of main.cpp:1000-1056
    (void)hipMemset(d_counter, 0, sizeof(unsigned int));
    (void)hipMemset(d_imagen, 0, n * sizeof(Color));
    launchKernelRenderTiny(grid, block, d_camara, d_primitivas, (int)prims.size(), d_luces, (int)luces.size(), nullptr,
                           0, W, H, 1, frame_cache, bvh_modelo_tiny.obtenerDatosGPU(), bvh_modelo_tiny.getPrimitivasGPU(),
                           (int)bvh_modelo_tiny.getNumPrimitivas(), imagen, nullptr, d_regs, d_counter, n, d_inf, d_thr,
                           true, true);
    if (!check("cache frame") || !dump(out + "img_prefix.bin", d_imagen, n) || !dump(out + "thr.bin", d_thr, n) ||
        !dump(out + "inference.bin", d_inf, n))
        return 1;
    mlp.inference(d_inf, d_pred, n);
    launchKernelComposite(grid, block, d_imagen, d_pred, d_thr, W, H);
    if (!check("inference + composite") || !dump(out + "pred.bin", d_pred, n) ||
        !dump(out + "img_composite.bin", d_imagen, n))
        return 1;
    unsigned int n_train = 0;
    (void)hipMemcpy(&n_train, d_counter, sizeof(unsigned int), hipMemcpyDeviceToHost);
    const unsigned int counter = n_train;
    n_train = n_train < 65536u ? n_train : 65536u;
    n_train = n_train < (unsigned)n ? n_train : (unsigned)n;
    n_train = n_train / 4 * 4;
    float loss = 0.0f;
    if (!dump(out + "registros.bin", d_regs, n_train)) return 1;
    if (n_train > 0) {   // main.cpp:1043 trains above 1024 records; the test scene is smaller
        const int threads = 256, blocks = ((int)n_train + threads - 1) / threads;
        launchKernelPrepararInferenciaTail(dim3(blocks), dim3(threads), d_regs, d_tail_in, (int)n_train);
        mlp.inference(d_tail_in, d_tail_pred, n_train);
        launchKernelCalcularTargets(dim3(blocks), dim3(threads), d_regs, d_tail_pred, d_train, (int)n_train);
        if (!check("tail inference + targets") || !dump(out + "tail_in.bin", d_tail_in, n_train) ||
            !dump(out + "tail_pred.bin", d_tail_pred, n_train) || !dump(out + "train_final.bin", d_train, n_train))
            return 1;
        loss = mlp.train_step(d_train, n_train);
        if (!check("train step")) return 1;
    }
    mlp.save_model(out + "model_after.bin");
    ColorMLP reloaded(15, 3, 16384, nullptr);
    reloaded.setBounds(bounds.min, bounds.max);
    const bool loaded = reloaded.load_model(out + "model_after.bin");
    reloaded.inference(d_inf, d_pred, n);
    if (!check("reload + inference") || !dump(out + "pred_reloaded.bin", d_pred, n)) return 1;
    FILE* f = std::fopen((out + "scalars.txt").c_str(), "w");
    if (!f) return 1;
    std::fprintf(f, "%u %u %.9g %d %u %u\n", counter, n_train, loss, loaded ? 1 : 0, bvh_modelo_tiny.getNumNodos(),
                 bvh_modelo_tiny.getNumPrimitivas());
    std::fclose(f);
    std::printf("shim_run: %d x %d, %u records, %u trained, loss %.6g\n", W, H, counter, n_train, loss);
    return 0;
}
