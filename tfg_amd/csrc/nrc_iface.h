// nrc_iface.h - what the path tracer needs from the radiance cache.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/tfg.h"


// Device buffers of a path tracer frame that the cache reads.
struct tfg_nrc_frame {
    const float4* records = nullptr;   // kRecF4 float4 per training record
    const float4* queries = nullptr;   // kQueryF4 float4 per query
    const uint32_t* counters = nullptr;   // CNT_RECORDS, CNT_QUERIES
    uint32_t n_paths = 0;
    uint32_t mode = 0;
};

namespace tfg {
// Enqueues, on stream s, the cache work of one frame: inference of every
// query composited into image, then (when training) tail inference,
// targets and the optimiser steps.  Graph-capturable.
int nrc_frame_enqueue(tfg_nrc* nrc, const tfg_nrc_frame& f, tfg_color* image, hipStream_t s);
// Deferred training: the samples the frame's optimiser steps will draw, with
// their targets (the head vertex and target of each, in sample order,
// nrc_gather_f4 float4 in all), and the record count are prepared on s;
// nrc_train_gathered_enqueue later runs those steps on them, with the same
// results as training on the frame's records at the end of the frame.
size_t nrc_gather_f4(const tfg_nrc* nrc);
int nrc_gather_enqueue(tfg_nrc* nrc, const tfg_nrc_frame& f, float4* out, uint32_t* out_n, hipStream_t s);
int nrc_train_gathered_enqueue(tfg_nrc* nrc, const float4* gathered, const uint32_t* n_rec, hipStream_t s);
}  // namespace tfg
