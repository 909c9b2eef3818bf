// mixed_kernels.hip -- segment-fused mixed-precision path (placeholder until
// the fused kernels land).
#include "hom_internal.h"

namespace hom {

int mixed_run(hom_plan *p, const int64_t *, const int64_t *, const uint8_t *, int64_t, int) {
    return set_error(p, HOM_ERR_STATE, "The mixed-precision path is not built yet.");
}

void mixed_free(hom_plan *) {}

}  // namespace hom
