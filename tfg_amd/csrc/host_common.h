// host_common.h - error reporting shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/tfg.h"

namespace tfg {

void set_error(const char* fmt, ...);

inline int hip_fail(hipError_t e, const char* what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return TFG_E_HIP;
}

}  // namespace tfg

#define TFG_HIP_CHECK(call)                                  \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return tfg::hip_fail(_e, #call); \
    } while (0)
