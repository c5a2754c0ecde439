// tinybvh_wrapper.h - drop-in replacement of the reference's TinyBVH class
// (tinybvh_wrapper.h:23-50, tinybvh_wrapper.cpp:30-98) for callers built
// with hipcc; include after the caller's Primitiva and TinyBVHD_GPU.  The
// CWBVH is built by tfg_cwbvh_build (8-wide, 80 B nodes, 64 B triangle
// records with primIdx in the last word) instead of tinybvh, and a copy of
// the triangle Primitiva records in primIdx order is uploaded next to it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../tfg.h"

class TinyBVH {
public:
    TinyBVH() = default;
    ~TinyBVH() { liberar(); }
    TinyBVH(const TinyBVH&) = delete;
    TinyBVH& operator=(const TinyBVH&) = delete;
    TinyBVH(TinyBVH&&) = delete;
    TinyBVH& operator=(TinyBVH&&) = delete;

    // tinybvh_wrapper.cpp:30-83: triangles in input order, others skipped.
    void construirBVH(const std::vector<Primitiva>& primitivas_host) {
        std::vector<float> verts;
        std::vector<Primitiva> tris;
        verts.reserve(primitivas_host.size() * 9);
        for (const Primitiva& p : primitivas_host) {
            if (p.tipo != TRIANGULO) continue;
            const auto& t = p.triangulo;
            const float v[9] = {(float)t.v0.x, (float)t.v0.y, (float)t.v0.z, (float)t.v1.x, (float)t.v1.y,
                                (float)t.v1.z, (float)t.v2.x, (float)t.v2.y, (float)t.v2.z};
            verts.insert(verts.end(), v, v + 9);
            tris.push_back(p);
        }
        liberar();
        if (tris.empty()) return;
        tfg_cwbvh_host* h = nullptr;
        if (tfg_cwbvh_build(verts.data(), (uint32_t)tris.size(), TFG_TRI_WATERTIGHT, &h) != TFG_OK) return;
        uint32_t depth = 0;
        tfg_cwbvh_host_info(h, &num_nodos_, nullptr, &depth, nullptr);
        const int rc = tfg_cwbvh_upload(h, &gpu_);
        tfg_cwbvh_host_free(h);
        if (rc != TFG_OK) return;
        if (hipMalloc(&prims_, tris.size() * sizeof(Primitiva)) != hipSuccess) return;
        if (hipMemcpy(prims_, tris.data(), tris.size() * sizeof(Primitiva), hipMemcpyHostToDevice) != hipSuccess) return;
        num_prims_ = (unsigned)tris.size();
    }
    void obtenerInfo() {
        std::printf("CWBVH nodes (80 bytes each): %u\n", gpu_.num_nodos);
        std::printf("Referenced primitives: %u\n", gpu_.num_tris);
    }
    TinyBVHD_GPU obtenerDatosGPU() const {
        TinyBVHD_GPU d;
        __builtin_memcpy(&d, &gpu_, sizeof(d));
        return d;
    }
    const TinyBVHD_GPU* getNodosGPU() const { return reinterpret_cast<const TinyBVHD_GPU*>(&gpu_); }
    const Primitiva* getPrimitivasGPU() const { return prims_; }
    unsigned int getNumNodos() const { return num_nodos_; }
    unsigned int getNumPrimitivas() const { return num_prims_; }

private:
    void liberar() {
        if (gpu_.d_nodos) tfg_cwbvh_gpu_free(&gpu_);
        if (prims_) (void)hipFree(prims_);
        prims_ = nullptr;
        num_prims_ = 0;
        num_nodos_ = 0;
    }
    tfg_cwbvh_gpu gpu_{};
    Primitiva* prims_ = nullptr;
    unsigned int num_prims_ = 0;
    unsigned int num_nodos_ = 0;
};
