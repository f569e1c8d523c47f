// Pivoted fallback factorisation: FP64 LU with partial pivoting.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace omh {

// Workspace bytes of lu_factor / lu_solve for an n x n matrix.
size_t lu_workspace_bytes(int64_t n);

// P A = L U in place (A n x n, leading dimension lda); ipiv (device, n
// entries): row j was interchanged with row ipiv[j] at step j (0-based, as
// LAPACK's DGETRF).  Returns 0, or 1 + the first column whose pivot is
// exactly zero (the factor is then unusable).
int lu_factor(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* work, hipStream_t st);

// B <- A^-1 B (n x nrhs, leading dimension ldb) with lu_factor's output.
void lu_solve(const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, const void* work, double* B, int64_t ldb,
              int64_t nrhs, hipStream_t st);
void lu_solve_chunk(const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, const void* work, double* B,
                    int64_t ldb, int64_t r, hipStream_t st);

}  // namespace omh
