// Dense MFMA linear algebra used by the HeadMat solve and the gains.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace omh {

constexpr int FACT_NB = 256;   // outer block of the factorisation

// C = alpha op(A) op(B) + beta C (column-major).  lower: square C, only
// entries with row >= col are computed (only ta=false, tb=true).
// prof: optional name under which the interior launch of a lower (SYRK)
// product is timed (ProfScope name + "_t<tile>").
template <typename T>
void gemm(bool ta, bool tb, bool lower, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
          const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st, const char* prof = nullptr);

struct FactorInfo {
    int info = 0;              // 0, or 1 + index of the first (near-)zero pivot
    const void* dinv = nullptr;
    const void* signs = nullptr;
};

// Bytes of workspace sym_factor / sym_solve need for an n x n matrix of
// element size elem.
size_t factor_workspace_bytes(int64_t n, size_t elem);

// In-place pivot-free signed Cholesky H = L S L^T of the lower triangle of A
// (n x n, leading dimension lda).  The strict upper triangle is not touched.
template <typename T>
FactorInfo sym_factor(T* A, int64_t n, int64_t lda, void* work, hipStream_t st);

// W <- H^-1 W for r right-hand sides with the factor of sym_factor.
template <typename T>
void sym_solve(const T* L, int64_t n, int64_t lda, const void* work, T* W, int64_t ldw, int64_t r, hipStream_t st);

void convert_f64_to_f32(const double* src, int64_t lds, float* dst, int64_t ldd, int64_t m, int64_t n, bool lower,
                        hipStream_t st);
void copy_f64(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n, hipStream_t st);
double frob_norm_f64(const double* a, int64_t lda, int64_t m, int64_t n, hipStream_t st);
double frob_norm_f32(const float* a, int64_t lda, int64_t m, int64_t n, hipStream_t st);

// Iterative refinement: X = H^-1 B with the FP32 factor (L32, work32) and
// FP64 residuals R = B - H X.  Stops when ||dX||_F / ||X||_F <= tol or after
// max_iter corrections; returns the number of corrections applied.
int mixed_refine(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                 const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                 double* final_rel_update, hipStream_t st);

// out(e, j) = sum_k vals(e,k) Hm(cols(e,k), j)  (3 non-zeros per row).
void sparse_rows(const int64_t* d_cols, const double* d_vals, int64_t ne, const double* Hm, int64_t ldh, int64_t n,
                 double* out, int64_t ldo, hipStream_t st);

}  // namespace omh
