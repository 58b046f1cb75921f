// Fused structured-box kernels (placeholder until the fused path lands).
#include <hip/hip_runtime.h>

#include "case.h"
#include "fused.h"

namespace mplb {

bool fused_supported(const mplb_case*, std::string* why) {
    if (why) *why = "not built";
    return false;
}
int fused_create(mplb_case*) { return MPLB_E_STATE; }
void fused_destroy(mplb_case*) {}
int fused_after_init(mplb_case*) { return 0; }
int fused_step(mplb_case*) { return MPLB_E_STATE; }
int fused_sync_to_generic(mplb_case*) { return 0; }
int fused_after_write(mplb_case*) { return 0; }
int fused_set_field(mplb_case*, const char*, const double*, size_t) { return 1; }
int fused_get_field(mplb_case*, const char*, double*, size_t) { return 1; }

}  // namespace mplb
