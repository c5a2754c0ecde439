// mlp.h - drop-in replacement of the reference's ColorMLP (mlp.h:19-70,
// mlp.cpp:16-276) on top of the library's radiance cache; include after the
// caller's DatosMLP, Color and Vector3d.  The tiny-cuda-nn JSON
// configuration argument is accepted and ignored (the network is fixed: 64
// encoded inputs, 5 x 64 ReLU, 3 outputs, EMA(0.99) over Adam(1e-2)); the
// weights file is this library's own format (tfg_nrc_get_weights order).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../tfg.h"

class ColorMLP {
public:
    // config_override stands for the reference's tcnn::json argument
    // (mlp.cpp:16-73: HashGrid + OneBlob + Identity by default, any tcnn
    // model by JSON).  This shim does not read it: the network is always
    // the library's (12-frequency position encoding + one-blob direction,
    // normal and roughness + albedos, 5 x 64 ReLU, BF16 operands, RGB out;
    // RelativeL2Luminance, EMA 0.99 over Adam 1e-2), and a non-null
    // configuration is reported on stderr once.  n_in must be 15 and n_out
    // 3 (DatosMLP's 15 inputs, Color's 3 outputs); any other pair leaves the
    // object without a network, whose calls then do nothing.
    ColorMLP(uint32_t n_in, uint32_t n_out, uint32_t batch, const void* config_override = nullptr) {
        if (config_override)
            std::fprintf(stderr, "ColorMLP: the tiny-cuda-nn configuration is not used; the library's fixed network "
                                 "is trained instead\n");
        if (n_in != 15 || n_out != 3) return;   // DatosMLP carries 15 inputs, Color 3 outputs
        tfg_nrc_default_config(&cfg_);
        cfg_.batch = (batch + 63) / 64 * 64;
        cfg_.steps_per_frame = 1;
        cfg_.min_records = 99;                  // mlp.cpp:96
        tfg_nrc_create(&cfg_, &nrc_);
    }
    ~ColorMLP() { tfg_nrc_destroy(nrc_); }
    ColorMLP(const ColorMLP&) = delete;
    ColorMLP& operator=(const ColorMLP&) = delete;

    float train_step(DatosMLP* buffer_samples_gpu, uint32_t n_total_samples_disponibles) {
        float loss = 0.0f;
        if (nrc_) tfg_nrc_train_datos(nrc_, reinterpret_cast<const tfg_datos_mlp*>(buffer_samples_gpu),
                                      n_total_samples_disponibles, &loss, nullptr);
        return loss;
    }
    void inference(DatosMLP* buffer_samples_gpu, Color* output_gpu, uint32_t n_samples) {
        if (nrc_) tfg_nrc_infer_datos(nrc_, reinterpret_cast<const tfg_datos_mlp*>(buffer_samples_gpu), n_samples,
                                      reinterpret_cast<tfg_color*>(output_gpu), nullptr);
    }
    // mlp.cpp:89-93; the cache is rebuilt with the new bounds (weights kept).
    void setBounds(const Vector3d& mn, const Vector3d& mx) {
        if (!nrc_) return;
        std::vector<float> m(TFG_NRC_PARAMS), e(TFG_NRC_PARAMS);
        tfg_nrc_get_weights(nrc_, m.data(), e.data());
        cfg_.bounds_min[0] = (float)mn.x; cfg_.bounds_min[1] = (float)mn.y; cfg_.bounds_min[2] = (float)mn.z;
        cfg_.bounds_max[0] = (float)mx.x; cfg_.bounds_max[1] = (float)mx.y; cfg_.bounds_max[2] = (float)mx.z;
        tfg_nrc_destroy(nrc_);
        nrc_ = nullptr;
        if (tfg_nrc_create(&cfg_, &nrc_) == TFG_OK) tfg_nrc_set_weights(nrc_, m.data(), e.data());
    }
    void save_model(const std::string& filename) {
        std::vector<float> m(TFG_NRC_PARAMS), e(TFG_NRC_PARAMS);
        if (!nrc_ || tfg_nrc_get_weights(nrc_, m.data(), e.data()) != TFG_OK) return;
        FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) return;
        const uint32_t hdr[2] = {TFG_NRC_PARAMS, cfg_.precision};
        std::fwrite("TFGNRC01", 1, 8, f);
        std::fwrite(hdr, sizeof(hdr), 1, f);
        std::fwrite(m.data(), sizeof(float), m.size(), f);
        std::fwrite(e.data(), sizeof(float), e.size(), f);
        std::fclose(f);
    }
    bool load_model(const std::string& filename) {
        FILE* f = std::fopen(filename.c_str(), "rb");
        if (!f || !nrc_) return false;
        char magic[8];
        uint32_t hdr[2];
        std::vector<float> m(TFG_NRC_PARAMS), e(TFG_NRC_PARAMS);
        const bool ok = std::fread(magic, 1, 8, f) == 8 && std::string(magic, 8) == "TFGNRC01" &&
                        std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == TFG_NRC_PARAMS &&
                        std::fread(m.data(), sizeof(float), m.size(), f) == m.size() &&
                        std::fread(e.data(), sizeof(float), e.size(), f) == e.size();
        std::fclose(f);
        return ok && tfg_nrc_set_weights(nrc_, m.data(), e.data()) == TFG_OK;
    }
    tfg_nrc* handle() const { return nrc_; }

private:
    tfg_nrc_config cfg_{};
    tfg_nrc* nrc_ = nullptr;
};
