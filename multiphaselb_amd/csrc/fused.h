// Fused structured-box path (kernels_fused.hip): one pull-stream + macroscopic +
// collide kernel per time step on boxes whose neighbour table is the duplicate-node
// periodic wrap, FP32 arithmetic with FP32 or FP16-deviation population storage.
#pragma once

#include <string>

struct mplb_case;

namespace mplb {

bool fused_supported(const mplb_case* c, std::string* why);
int fused_create(mplb_case* c);
void fused_destroy(mplb_case* c);
int fused_after_init(mplb_case* c);
int fused_step(mplb_case* c);
// make the generic device arrays (rho, T, U, f, g ...) describe the current time level
int fused_sync_to_generic(mplb_case* c);
int fused_after_write(mplb_case* c);
// return 1 when the request should fall through to the generic field copy
int fused_set_field(mplb_case* c, const char* name, const double* host, size_t n);
int fused_get_field(mplb_case* c, const char* name, double* host, size_t n);

}  // namespace mplb
