// Allocation-free mixed-precision solve (its buffers in a fixed device array).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace omh {

// Largest n x nrhs the fixed buffer holds.
constexpr int64_t NOALLOC_MAX_ELEMS = int64_t(1) << 22;

// mixed_refine (dense residual) with its buffers in the fixed array; no
// device allocation.  Throws std::invalid_argument when n x r exceeds
// NOALLOC_MAX_ELEMS.
int mixed_refine_noalloc(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                         const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                         double* final_rel_update, bool* converged, hipStream_t st);

}  // namespace omh
