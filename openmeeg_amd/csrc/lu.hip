// Pivoted fallback factorisation: LU with partial (row) pivoting in FP64,
// P A = L U, for matrices on which the pivot-free signed Cholesky breaks
// down (an unknown order in which a leading block is singular or
// indefinite).  The reference solves every symmetric HeadMat with the
// Bunch-Kaufman factorisation DSPTRF/DSPTRS (symmatrix.cpp:133-152,
// symmatrix.h:259-297); this path solves any non-singular matrix, at twice
// the flops of a symmetric factorisation, and runs only when the signed
// Cholesky reports a vanishing pivot.
//
// Right-looking, blocks of LU_NB = 64 columns:
//   panel   column by column: k_lu_colmax (per-block partial maxima of
//           |A(j:n, j)|), k_lu_pivot (one workgroup: the first row of the
//           largest magnitude, as LAPACK's IDAMAX; the row swap inside the
//           panel), k_lu_rank1 (the multipliers of column j and the rank-1
//           update of the panel's later columns, one thread per row);
//   swaps   k_lu_laswp: the panel's interchanges on every other column;
//   U12     L11^-1 A12 (unit lower 64 x 64 inverse, k_lu_trinv, then a GEMM);
//   A22    -= L21 U12 (MFMA GEMM).
// The solve applies the interchanges, then blocked forward (unit L) and
// backward (U) substitution with the stored 64 x 64 diagonal-block
// inverses and GEMMs for the off-diagonal blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "common.hpp"
#include "linalg.hpp"
#include "lu.hpp"

namespace omh {

namespace {

constexpr int LU_NB = 64;
constexpr int MAX_PARTS = 256;   // partial maxima per column search

__global__ void k_lu_zero(int32_t* p) { p[0] = 0; }

__global__ void __launch_bounds__(256) k_lu_colmax(const double* __restrict__ A, int64_t lda, int64_t j, int64_t n,
                                                   double* __restrict__ pval, int64_t* __restrict__ pidx) {
    __shared__ double sv[256];
    __shared__ int64_t si[256];
    double best = -1.0;
    int64_t bi = n;
    const double* col = A + j * lda;
    for (int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v = fabs(col[i]);
        if (v > best) {   // rows visited in increasing order: the first maximum is kept
            best = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double a = sv[threadIdx.x], b = sv[threadIdx.x + s];
            const int64_t ia = si[threadIdx.x], ib = si[threadIdx.x + s];
            if (b > a || (b == a && ib < ia)) {
                sv[threadIdx.x] = b;
                si[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pval[blockIdx.x] = sv[0];
        pidx[blockIdx.x] = si[0];
    }
}

// One workgroup: the pivot of column j from nparts partial maxima; rows j
// and p are swapped in the panel columns [k0, k0 + nb).  A zero column sets
// *info = j + 1 (first one only) and leaves the rows in place.
__global__ void __launch_bounds__(256) k_lu_pivot(double* __restrict__ A, int64_t lda, int64_t j, int64_t k0, int nb,
                                                  const double* __restrict__ pval, const int64_t* __restrict__ pidx,
                                                  int nparts, int32_t* __restrict__ ipiv, int32_t* __restrict__ info) {
    __shared__ int64_t p_sh;
    if (threadIdx.x == 0) {
        double best = -1.0;
        int64_t bi = j;
        for (int q = 0; q < nparts; ++q)
            if (pval[q] > best || (pval[q] == best && pidx[q] < bi)) {
                best = pval[q];
                bi = pidx[q];
            }
        if (!(best > 0.0)) {   // exactly zero (or NaN) column: singular
            if (*info == 0) *info = (int32_t)(j + 1);
            bi = j;
        }
        ipiv[j] = (int32_t)bi;
        p_sh = bi;
    }
    __syncthreads();
    const int64_t p = p_sh;
    if (p != j)
        for (int c = threadIdx.x; c < nb; c += blockDim.x) {
            double* cp = A + (k0 + c) * lda;
            const double t = cp[j];
            cp[j] = cp[p];
            cp[p] = t;
        }
}

// Rows i in (j, n): l = A(i, j) / A(j, j) stored in place, then
// A(i, c) -= l A(j, c) for the panel columns c in (j, kend).
__global__ void __launch_bounds__(256) k_lu_rank1(double* __restrict__ A, int64_t lda, int64_t j, int64_t kend,
                                                  int64_t n) {
    const int64_t i = j + 1 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double piv = A[j + j * lda];
    const double l = piv != 0.0 ? A[i + j * lda] / piv : 0.0;
    A[i + j * lda] = l;
    for (int64_t c = j + 1; c < kend; ++c) A[i + c * lda] -= l * A[j + c * lda];
}

// The interchanges ipiv[k0 .. k0 + nb) on the columns [c0, c1) of A.
__global__ void __launch_bounds__(256) k_lu_laswp(double* __restrict__ A, int64_t lda, int64_t k0, int nb,
                                                  const int32_t* __restrict__ ipiv, int64_t c0, int64_t c1) {
    const int64_t c = c0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= c1) return;
    double* cp = A + c * lda;
    for (int q = 0; q < nb; ++q) {
        const int64_t j = k0 + q, p = ipiv[j];
        if (p != j) {
            const double t = cp[j];
            cp[j] = cp[p];
            cp[p] = t;
        }
    }
}

// Inverse of the nb x nb diagonal block at (k0, k0): unit lower L11 into
// Linv and upper U11 into Uinv (both nb x nb, leading dimension LU_NB).
// One workgroup of LU_NB threads; thread c solves column c.
__global__ void __launch_bounds__(LU_NB) k_lu_trinv(const double* __restrict__ A, int64_t lda, int64_t k0, int nb,
                                                    double* __restrict__ Linv, double* __restrict__ Uinv) {
    __shared__ double B[LU_NB][LU_NB + 1];
    const int c = threadIdx.x;
    for (int r = 0; r < LU_NB; ++r) B[r][c] = (r < nb && c < nb) ? A[(k0 + r) + (k0 + c) * lda] : 0.0;
    __syncthreads();
    if (c >= nb) {
        for (int r = 0; r < LU_NB; ++r) Linv[r + c * LU_NB] = Uinv[r + c * LU_NB] = r == c ? 1.0 : 0.0;
        return;
    }
    // column c of L^-1: x_r = e_c(r) - sum_{s<r} L(r,s) x_s
    double x[LU_NB];
#pragma unroll
    for (int r = 0; r < LU_NB; ++r) {
        double v = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int s = 0; s < r; ++s) v -= B[r][s] * x[s];
        x[r] = r < nb ? v : 0.0;
    }
#pragma unroll
    for (int r = 0; r < LU_NB; ++r) Linv[r + c * LU_NB] = x[r];
    // column c of U^-1: x_r = (e_c(r) - sum_{s>r} U(r,s) x_s) / U(r,r)
#pragma unroll
    for (int r = LU_NB - 1; r >= 0; --r) {
        double v = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int s = r + 1; s < LU_NB; ++s) v -= B[r][s] * x[s];
        x[r] = r < nb ? v / B[r][r] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < LU_NB; ++r) Uinv[r + c * LU_NB] = x[r];
}

// The interchanges ipiv[0 .. n) on the rows of B (n x r), forward.
__global__ void __launch_bounds__(256) k_lu_permute_rhs(double* __restrict__ B, int64_t ldb, int64_t n, int64_t r,
                                                        const int32_t* __restrict__ ipiv) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= r) return;
    double* cp = B + c * ldb;
    for (int64_t j = 0; j < n; ++j) {
        const int64_t p = ipiv[j];
        if (p != j) {
            const double t = cp[j];
            cp[j] = cp[p];
            cp[p] = t;
        }
    }
}

struct LuWork {
    double* Linv;     // nblk blocks of LU_NB x LU_NB
    double* Uinv;     // nblk blocks
    double* T;        // LU_NB x max(n, nrhs) temporary
    double* pval;     // MAX_PARTS
    int64_t* pidx;    // MAX_PARTS
};

LuWork carve(void* work, int64_t n) {
    const int64_t nblk = cdiv(n, LU_NB);
    char* p = (char*)work;
    LuWork w;
    w.Linv = (double*)p;
    p += sizeof(double) * (size_t)nblk * LU_NB * LU_NB;
    w.Uinv = (double*)p;
    p += sizeof(double) * (size_t)nblk * LU_NB * LU_NB;
    w.pval = (double*)p;
    p += sizeof(double) * MAX_PARTS;
    w.pidx = (int64_t*)p;
    p += sizeof(int64_t) * MAX_PARTS;
    w.T = (double*)p;
    return w;
}

}  // namespace

size_t lu_workspace_bytes(int64_t n) {
    const int64_t nblk = cdiv(n, LU_NB);
    return sizeof(double) * (size_t)nblk * LU_NB * LU_NB * 2 + sizeof(double) * MAX_PARTS +
           sizeof(int64_t) * MAX_PARTS + sizeof(double) * (size_t)LU_NB * (size_t)n;
}

int lu_factor(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* work, hipStream_t st) {
    if (n <= 0) return 0;
    LuWork w = carve(work, n);
    int32_t* d_info = (int32_t*)scratch(SCR_SMALL, 64, st);
    k_lu_zero<<<1, 1, 0, st>>>(d_info);
    OMH_LAUNCH_CHECK();
    for (int64_t k0 = 0; k0 < n; k0 += LU_NB) {
        const int nb = (int)std::min<int64_t>(LU_NB, n - k0);
        const int64_t kend = k0 + nb;
        for (int64_t j = k0; j < kend; ++j) {
            const int64_t rows = n - j;
            const int parts = (int)std::min<int64_t>(MAX_PARTS, cdiv(rows, 256));
            k_lu_colmax<<<parts, 256, 0, st>>>(A, lda, j, n, w.pval, w.pidx);
            k_lu_pivot<<<1, 256, 0, st>>>(A, lda, j, k0, nb, w.pval, w.pidx, parts, ipiv, d_info);
            if (rows > 1) k_lu_rank1<<<(unsigned)cdiv(rows - 1, 256), 256, 0, st>>>(A, lda, j, kend, n);
            OMH_LAUNCH_CHECK();
        }
        // the panel's interchanges on the columns left and right of it
        if (k0 > 0) k_lu_laswp<<<(unsigned)cdiv(k0, 256), 256, 0, st>>>(A, lda, k0, nb, ipiv, 0, k0);
        if (kend < n) k_lu_laswp<<<(unsigned)cdiv(n - kend, 256), 256, 0, st>>>(A, lda, k0, nb, ipiv, kend, n);
        const int64_t b = k0 / LU_NB;
        double* Li = w.Linv + (size_t)b * LU_NB * LU_NB;
        double* Ui = w.Uinv + (size_t)b * LU_NB * LU_NB;
        k_lu_trinv<<<1, LU_NB, 0, st>>>(A, lda, k0, nb, Li, Ui);
        OMH_LAUNCH_CHECK();
        if (kend < n) {
            const int64_t nc = n - kend;
            // U12 = L11^-1 A12 (through a copy of A12)
            copy_f64(A + k0 + kend * lda, lda, w.T, LU_NB, nb, nc, st);
            gemm<double>(false, false, false, nb, nc, nb, 1.0, Li, LU_NB, w.T, LU_NB, 0.0, A + k0 + kend * lda, lda, st);
            // A22 -= L21 U12
            gemm<double>(false, false, false, n - kend, nc, nb, -1.0, A + kend + k0 * lda, lda, A + k0 + kend * lda,
                         lda, 1.0, A + kend + kend * lda, lda, st);
        }
    }
    int32_t info = 0;
    OMH_HIP(hipMemcpyAsync(&info, d_info, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    return info;
}

void lu_solve(const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, const void* work, double* B, int64_t ldb,
              int64_t nrhs, hipStream_t st) {
    if (n <= 0 || nrhs <= 0) return;
    // the workspace's temporary holds LU_NB x n: right-hand sides in chunks of n
    for (int64_t c0 = 0; c0 < nrhs; c0 += n)
        lu_solve_chunk(LU, n, lda, ipiv, work, B + c0 * ldb, ldb, std::min<int64_t>(n, nrhs - c0), st);
}

void lu_solve_chunk(const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, const void* work, double* B,
                    int64_t ldb, int64_t r, hipStream_t st) {
    LuWork w = carve(const_cast<void*>(work), n);
    k_lu_permute_rhs<<<(unsigned)cdiv(r, 256), 256, 0, st>>>(B, ldb, n, r, ipiv);
    OMH_LAUNCH_CHECK();
    const int64_t nblk = cdiv(n, LU_NB);
    // forward: Y = L^-1 P B, block by block
    for (int64_t b = 0; b < nblk; ++b) {
        const int64_t k0 = b * LU_NB, nb = std::min<int64_t>(LU_NB, n - k0), kend = k0 + nb;
        copy_f64(B + k0, ldb, w.T, LU_NB, nb, r, st);
        gemm<double>(false, false, false, nb, r, nb, 1.0, w.Linv + (size_t)b * LU_NB * LU_NB, LU_NB, w.T, LU_NB, 0.0,
                     B + k0, ldb, st);
        if (kend < n)
            gemm<double>(false, false, false, n - kend, r, nb, -1.0, LU + kend + k0 * lda, lda, B + k0, ldb, 1.0,
                         B + kend, ldb, st);
    }
    // backward: X = U^-1 Y
    for (int64_t b = nblk - 1; b >= 0; --b) {
        const int64_t k0 = b * LU_NB, nb = std::min<int64_t>(LU_NB, n - k0);
        copy_f64(B + k0, ldb, w.T, LU_NB, nb, r, st);
        gemm<double>(false, false, false, nb, r, nb, 1.0, w.Uinv + (size_t)b * LU_NB * LU_NB, LU_NB, w.T, LU_NB, 0.0,
                     B + k0, ldb, st);
        if (k0 > 0)
            gemm<double>(false, false, false, k0, r, nb, -1.0, LU + k0 * lda, lda, B + k0, ldb, 1.0, B, ldb, st);
    }
}

}  // namespace omh
