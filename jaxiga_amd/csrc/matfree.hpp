// Internal launcher of the matrix-free operator, shared by the Krylov drivers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jaxiga_hip.h"

namespace iga {

// y = K x over the free dofs (x, y FP64) with a workspace filled by
// iga_mf_setup (its precision and layout are looked up by address); a nonzero
// *skip makes every launch a no-op (used by the Krylov loops once they have
// converged). With `parts`, the x.y partials (one per block of the last
// kernel, at most kRedBlocks) are stored there and their count in *nparts.
int mf_apply_launch(const iga_patch_t* patch, const void* work, int64_t work_bytes, int n_free,
                    const int32_t* freed, const int32_t* full_to_free, const double* x, double* y,
                    const int* skip, double* parts, int* nparts, hipStream_t st);

}  // namespace iga
