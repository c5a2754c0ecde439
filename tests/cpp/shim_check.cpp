// Compile check of include/tfg_ref/*.h against stand-in types with the
// reference's field names and layouts (vector3d.h, color.h, primitiva.h,
// camara.h, luzPuntual.h, imagen_gpu.h, tinybvh_wrapper.h:15-21,
// mlp_types.h).  Built by tests/test_shims_cpu.py with -fsyntax-only.
#include <hip/hip_runtime.h>

#include <cstdint>

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

void drive(Camara* cam, Primitiva* p, LuzPuntual* l, ImagenGPU img, RegistroEntrenamiento* r, DatosMLP* d, Color* c,
           unsigned* cnt, int* idx, float* in, float* tgt, SceneBounds b, std::vector<Primitiva>& host) {
    TinyBVH bvh;
    bvh.construirBVH(host);
    launchInicializarCamara(cam, 64, 64);
    launchKernelRenderTiny(dim3(4, 4), dim3(16, 16), cam, p, 0, l, 1, nullptr, 0, 64, 64, 1, 0, bvh.obtenerDatosGPU(),
                           bvh.getPrimitivasGPU(), bvh.getNumPrimitivas(), img, nullptr, r, cnt, 4096, d, c, true, true);
    launchKernelComposite(dim3(4, 4), dim3(16, 16), c, c, c, 64, 64);
    launchKernelPrepararInferenciaTail(dim3(1), dim3(256), r, d, 16);
    launchKernelCalcularTargets(dim3(1), dim3(256), r, c, d, 16);
    launchGenerarIndicesAleatorios(1, 256, nullptr, idx, 16, 16, 0u);
    launchPrepararDatosEntrenamiento(1, 256, nullptr, d, idx, 16, 15, 3, b, in, tgt);
    launchPrepararDatosInferencia(1, 256, nullptr, d, 16u, 15u, in, b);
    launchGuardarSalidaInferencia(1, 256, nullptr, tgt, c, 16);
    ColorMLP mlp(15, 3, 16384);
    mlp.setBounds(Vector3d{-6, -4, -10}, Vector3d{6, 6, 2});
    (void)mlp.train_step(d, 4096);
    mlp.inference(d, c, 4096);
    mlp.save_model("m.bin");
    (void)mlp.load_model("m.bin");
}
