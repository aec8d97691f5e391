// Placeholder for the adjoint entry point (removed once implemented).
#include <hip/hip_runtime.h>
#include "../../include/smg.h"
extern "C" {
smg_status smg_pe_backward(smg_ctx*, int, const float*, int, double*, double*, smg_stats*) { return SMG_ERR_STATE; }
}
