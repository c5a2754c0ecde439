// nrc.hip - radiance cache (placeholder until the MLP kernels land).
#include "host_common.h"
#include "nrc_iface.h"

struct tfg_nrc {
    int unused;
};

namespace tfg {
int nrc_frame_enqueue(tfg_nrc*, const tfg_nrc_frame&, tfg_color*, hipStream_t) {
    set_error("radiance cache not available");
    return TFG_E_STATE;
}
}  // namespace tfg
