// Dense linear algebra on MFMA for the HeadMat solve and the gain.
//
// The reference factors the packed HeadMat with Bunch-Kaufman DSPTRF and
// inverts/solves with DSPTRI/DSPTRS (symmatrix.h:259-297,
// symmatrix.cpp:133-152), then forms the gain with sparse x dense and DGEMM
// products (gain.h:52-89, sparse_matrix.cpp:36-51, matrix.h:333-350).
//
// Here the matrix is held in the triangles-first unknown order, where
// H = L S L^T exists without pivoting (S = diag(+-1): triangle unknowns +1,
// vertex unknowns -1).  The factorisation is right-looking and blocked by
// NB = 256:
//   1. k_diag_factor: one workgroup factors the NB x NB diagonal block
//      (inner blocks of 32 in LDS), writes the signs and the explicit
//      inverse of the diagonal L block;
//   2. panel X = A21 * Linv^T (MFMA GEMM), L21 = X * S1;
//   3. trailing update A22 -= X * L21^T on the lower triangle (MFMA GEMM).
// Triangular solves with many right-hand sides use the stored diagonal
// inverses and MFMA GEMM updates.
//
// MFMA: v_mfma_f64_16x16x4f64 (FP64) and v_mfma_f32_16x16x4f32 (FP32).
// Operand lane layout (one 16x16x4 block): A(i,k) in lane i + 16k,
// B(k,j) in lane j + 16k; D(i,j) in lane j + 16*(i%4), element i/4 for FP64
// and in lane j + 16*(i/4), element i%4 for FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "linalg.hpp"

namespace omh {

namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;   // LDS destination of direct global-to-LDS loads
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
// row(lane, r): row inside the 16x16 block of accumulator element r of a lane
// (the column is lane % 16).  Measured on gfx950 by omh_probe_mfma_layout.
template <> struct Acc<double> {
    typedef double4_t type;
    static __device__ __forceinline__ type mfma(double a, double b, type c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Acc<float> {
    typedef float4_t type;
    static __device__ __forceinline__ type mfma(float a, float b, type c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ __forceinline__ void upper_tile(int64_t t, int64_t& lo, int64_t& hi) {
    int64_t h = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (h * (h + 1) / 2 > t) --h;
    while ((h + 1) * (h + 2) / 2 <= t) ++h;
    hi = h;
    lo = t - h * (h + 1) / 2;
}

// k-slice depth of the GEMM tiles (FP32 may use deeper slices: fewer
// barriers per MFMA, at twice the LDS per workgroup).
#ifndef OMH_GEMM_BK_F32
#define OMH_GEMM_BK_F32 16
#endif
template <typename T> struct KSlice { static constexpr int v = sizeof(T) == 4 ? OMH_GEMM_BK_F32 : 16; };
constexpr int LDS_PAD = 16;   // row stride BM+16: the four 16-lane row groups of a fragment read hit disjoint banks

template <typename T> struct Vec;
template <> struct Vec<float> { typedef float4 type; static constexpr int n = 4; };
template <> struct Vec<double> { typedef double2 type; static constexpr int n = 2; };

// Loads an R x BK operand tile into registers.  MK: element (r, kk) at
// base[r + kk*ld] (contiguous along r); else base[kk + r*ld] (contiguous
// along kk).  load_full: the tile is in range and 16-byte aligned (vector
// loads, no guards); load_guarded: element-wise guarded loads.
template <typename T, bool MK, int R>
struct TileLoader {
    static constexpr int BK = KSlice<T>::v;
    static constexpr int V = Vec<T>::n;
    static constexpr int NV = R * BK / V / 256;   // vectors per thread
    T reg[NV][V];

    __device__ __forceinline__ static void coords(int v, int& r, int& kk) {
        if (MK) { kk = v / (R / V); r = (v % (R / V)) * V; }
        else { r = v / (BK / V); kk = (v % (BK / V)) * V; }
    }

    __device__ __forceinline__ void load_full(const T* __restrict__ base, int64_t ld, int64_t r0, int64_t k0, int tid) {
#pragma unroll
        for (int p = 0; p < NV; ++p) {
            int r, kk;
            coords(tid + 256 * p, r, kk);
            const T* ptr = MK ? base + (r0 + r) + (k0 + kk) * ld : base + (k0 + kk) + (r0 + r) * ld;
            const typename Vec<T>::type x = *reinterpret_cast<const typename Vec<T>::type*>(ptr);
            const T* xs = reinterpret_cast<const T*>(&x);
#pragma unroll
            for (int q = 0; q < V; ++q) reg[p][q] = xs[q];
        }
    }

    __device__ __forceinline__ void load_guarded(const T* __restrict__ base, int64_t ld, int64_t r0, int64_t k0,
                                                 int64_t rmax, int64_t kmax, int tid) {
#pragma unroll
        for (int p = 0; p < NV; ++p) {
            int r, kk;
            coords(tid + 256 * p, r, kk);
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int64_t rr = r0 + r + (MK ? q : 0), kq = k0 + kk + (MK ? 0 : q);
                reg[p][q] = (rr < rmax && kq < kmax) ? (MK ? base[rr + kq * ld] : base[kq + rr * ld]) : T(0);
            }
        }
    }

    // LDS tile S[kk][r] (row stride R + LDS_PAD)
    __device__ __forceinline__ void store(T* __restrict__ S, int tid) const {
#pragma unroll
        for (int p = 0; p < NV; ++p) {
            int r, kk;
            coords(tid + 256 * p, r, kk);
#pragma unroll
            for (int q = 0; q < V; ++q) {
                if (MK) S[kk * (R + LDS_PAD) + r + q] = reg[p][q];
                else S[(kk + q) * (R + LDS_PAD) + r] = reg[p][q];
            }
        }
    }
};

// L2-aware orders.  Tile rows are grouped in bands of BAND rows; inside a band
// the tiles run column by column, so a run of consecutive tiles covers a
// compact block of rows and columns and shares its operand panels.  (The
// lower-triangle band B holds 8B*h tiles left of its diagonal block and the
// h(h+1)/2 of that block, h = min(8, ntm - 8B); full bands before B hold
// 32B^2 + 4B tiles in all.)
constexpr int BAND = 8;

__device__ __forceinline__ void lower_band_tile(int64_t t, int64_t ntm, int64_t& tm, int64_t& tn) {
    const int64_t nfull = ntm / BAND;
    int64_t B = (int64_t)((sqrt(16.0 + 128.0 * (double)t) - 4.0) / 64.0);
    while (B > 0 && 32 * B * B + 4 * B > t) --B;
    while (32 * (B + 1) * (B + 1) + 4 * (B + 1) <= t) ++B;
    if (B > nfull) B = nfull;
    const int64_t h = min((int64_t)BAND, ntm - BAND * B);
    int64_t o = t - (32 * B * B + 4 * B);
    const int64_t left = BAND * B * h;
    if (o < left) {
        tn = o / h;
        tm = BAND * B + o % h;
        return;
    }
    o -= left;
    int64_t c = 0, start = 0;
    while (c < h - 1 && start + (h - c) <= o) {
        start += h - c;
        ++c;
    }
    tn = BAND * B + c;
    tm = BAND * B + c + (o - start);
}

__device__ __forceinline__ void full_band_tile(int64_t t, int64_t ntm, int64_t ntn, int64_t& tm, int64_t& tn) {
    const int64_t B = t / (BAND * ntn);
    const int64_t h = min((int64_t)BAND, ntm - BAND * B);
    const int64_t o = t - B * BAND * ntn;
    tn = o / h;
    tm = BAND * B + o % h;
}

// Workgroups are dispatched to the 8 XCDs round-robin; workgroup b takes tile
// (b % 8) * per + b / 8 so that each XCD (and its L2) works through one
// contiguous part of the tile order.  Tiles past ntiles return at once.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t per) { return (b % 8) * per + b / 8; }

__device__ __forceinline__ void map_tile(const TileMap& map, int64_t t, int64_t& tm, int64_t& tn) {
    switch (map.kind) {
        case 1: upper_tile(t, tn, tm); break;   // tn <= tm
        case 2: tm = map.tm0; tn = t; break;
        case 3: tn = map.tn0; tm = t; break;
        case 4: lower_band_tile(t, map.ntm, tm, tn); break;
        case 5: full_band_tile(t, map.ntm, map.tn0, tm, tn); break;   // tn0 holds the tile columns
        case 6: {   // lower part of the first tn0 tile columns, row by row
            const int64_t nc = map.tn0, top = nc * (nc + 1) / 2;
            if (t < top) {
                upper_tile(t, tn, tm);
            } else {
                tm = nc + (t - top) / nc;
                tn = (t - top) % nc;
            }
            break;
        }
        default: tn = t / map.ntm; tm = t - tn * map.ntm; break;
    }
}

// Tile rows of a launch that skips structurally zero rows: compressed tile
// index x in [c[k], c[k + 1]) is tile row s[k] + x - c[k] (n intervals in
// ascending order; n = 0: the identity).
constexpr int MAX_REMAP = 8;
struct TileRemap {
    int n = 0;
    int64_t c[MAX_REMAP + 1] = {};
    int64_t s[MAX_REMAP] = {};
};

__device__ __forceinline__ int64_t remap_tile(const TileRemap& r, int64_t x) {
    int k = 0;
    while (k + 1 < r.n && x >= r.c[k + 1]) ++k;
    return r.s[k] + (x - r.c[k]);
}

// C = alpha * op(A) * op(B) + beta * C on a (32 WB) x (32 WB) tile per
// workgroup of 256 threads (2 x 2 waves, each 16 WB x 16 WB = WB x WB MFMA
// blocks), k-slices of BK (16; FP32 configurable) double-buffered in LDS.
// TA: op(A)(i,kk) = A[kk + i*lda] (A stored k x m); else A[i + kk*lda].
// TB: op(B)(kk,j) = B[j + kk*ldb] (B stored n x k); else B[kk + j*ldb].
// LOWER: only entries with row >= col are written (m == n).
// EDGE = false: every tile of the launch lies inside m x n and the operands
// are 16-byte aligned, so the k loop uses unguarded vector loads (a partial
// last k-slice is done after the loop).  EDGE = true: guarded loads.
template <typename T, bool TA, bool TB, bool LOWER, int WB, bool EDGE>
__device__ __forceinline__ void gemm_tile(int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A,
                                          int64_t lda, const T* __restrict__ B, int64_t ldb, T beta,
                                          T* __restrict__ C, int64_t ldc, int64_t tm, int64_t tn) {
    constexpr int BM = 32 * WB, BN = 32 * WB;
    constexpr int BK = KSlice<T>::v;
    __shared__ T As[2][BK * (BM + LDS_PAD)];
    __shared__ T Bs[2][BK * (BN + LDS_PAD)];
    const int64_t row0 = tm * BM, col0 = tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * (16 * WB), wn = (wave & 1) * (16 * WB);

    typename Acc<T>::type acc[WB][WB];
#pragma unroll
    for (int a = 0; a < WB; ++a)
#pragma unroll
        for (int b = 0; b < WB; ++b) acc[a][b] = typename Acc<T>::type{0, 0, 0, 0};

    TileLoader<T, !TA, BM> la;
    TileLoader<T, TB, BN> lb;

    auto compute = [&](const T* Ab, const T* Bb) {
#pragma unroll
        for (int k4 = 0; k4 < BK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            T fa[WB], fb[WB];
#pragma unroll
            for (int a = 0; a < WB; ++a) fa[a] = Ab[kr * (BM + LDS_PAD) + wm + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < WB; ++b) fb[b] = Bb[kr * (BN + LDS_PAD) + wn + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < WB; ++a)
#pragma unroll
                for (int b = 0; b < WB; ++b) acc[a][b] = Acc<T>::mfma(fa[a], fb[b], acc[a][b]);
        }
    };

    if (!EDGE) {
        const int64_t nk = k / BK;   // full k-slices
        if (nk > 0) {
            la.load_full(A, lda, row0, 0, tid);
            lb.load_full(B, ldb, col0, 0, tid);
            la.store(As[0], tid);
            lb.store(Bs[0], tid);
            __syncthreads();
            for (int64_t kt = 0; kt < nk; ++kt) {
                const int buf = (int)(kt & 1);
                if (kt + 1 < nk) {
                    la.load_full(A, lda, row0, (kt + 1) * BK, tid);
                    lb.load_full(B, ldb, col0, (kt + 1) * BK, tid);
                }
                compute(As[buf], Bs[buf]);
                if (kt + 1 < nk) {
                    la.store(As[buf ^ 1], tid);
                    lb.store(Bs[buf ^ 1], tid);
                }
                __syncthreads();
            }
        }
        if (k % BK) {   // partial last k-slice
            la.load_guarded(A, lda, row0, nk * BK, m, k, tid);
            lb.load_guarded(B, ldb, col0, nk * BK, n, k, tid);
            la.store(As[0], tid);
            lb.store(Bs[0], tid);
            __syncthreads();
            compute(As[0], Bs[0]);
        }
    } else {
        // the same double-buffered pipeline with guarded loads
        const int64_t nk = (k + BK - 1) / BK;
        if (nk > 0) {
            la.load_guarded(A, lda, row0, 0, m, k, tid);
            lb.load_guarded(B, ldb, col0, 0, n, k, tid);
            la.store(As[0], tid);
            lb.store(Bs[0], tid);
            __syncthreads();
            for (int64_t kt = 0; kt < nk; ++kt) {
                const int buf = (int)(kt & 1);
                if (kt + 1 < nk) {
                    la.load_guarded(A, lda, row0, (kt + 1) * BK, m, k, tid);
                    lb.load_guarded(B, ldb, col0, (kt + 1) * BK, n, k, tid);
                }
                compute(As[buf], Bs[buf]);
                if (kt + 1 < nk) {
                    la.store(As[buf ^ 1], tid);
                    lb.store(Bs[buf ^ 1], tid);
                }
                __syncthreads();
            }
        }
    }
    // Epilogue.  FP32: a lane's 4 elements are 4 consecutive rows of one
    // column, written with one 16-byte access on interior tiles.
    const bool vec_c = !EDGE && (!LOWER || tm != tn) && (reinterpret_cast<uintptr_t>(C) % 16 == 0) && (ldc % 4 == 0);
#pragma unroll
    for (int a = 0; a < WB; ++a)
#pragma unroll
        for (int b = 0; b < WB; ++b) {
            if (sizeof(T) == 4 && vec_c) {
                const int64_t gi = row0 + wm + a * 16 + 4 * (lane >> 4);
                const int64_t gj = col0 + wn + b * 16 + (lane & 15);
                float4* cp = reinterpret_cast<float4*>(C + gi + gj * ldc);
                float4 v = make_float4(alpha * acc[a][b][0], alpha * acc[a][b][1], alpha * acc[a][b][2],
                                       alpha * acc[a][b][3]);
                if (beta != T(0)) {
                    const float4 c = *cp;
                    v.x += beta * c.x; v.y += beta * c.y; v.z += beta * c.z; v.w += beta * c.w;
                }
                *cp = v;
                continue;
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gi = row0 + wm + a * 16 + Acc<T>::row(lane, r);
                const int64_t gj = col0 + wn + b * 16 + (lane & 15);
                if (gi < m && gj < n && (!LOWER || gi >= gj)) {
                    T v = alpha * acc[a][b][r];
                    if (beta != T(0)) v += beta * C[gi + gj * ldc];
                    C[gi + gj * ldc] = v;
                }
            }
        }
}

template <typename T, bool TA, bool TB, bool LOWER, int WB, bool EDGE>
__global__ void __launch_bounds__(256) k_gemm(int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A,
                                              int64_t lda, const T* __restrict__ B, int64_t ldb, T beta,
                                              T* __restrict__ C, int64_t ldc, TileMap map) {
    int64_t t = blockIdx.x;
    if (map.kind >= 4) {
        t = xcd_tile(t, map.per);
        if (t >= map.ntiles) return;
    }
    int64_t tm, tn;
    map_tile(map, t, tm, tn);
    gemm_tile<T, TA, TB, LOWER, WB, EDGE>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn);
}

// Split-K product for outputs of few tiles with a long inner dimension:
// workgroup (x, y) computes tile x of W_y = op(A) op(B) restricted to the
// y-th chunk of kc inner indices (the last one shorter), W_y at W + y m n
// with leading dimension m.  k_splitk_sum then forms C = alpha sum_y W_y +
// beta C, summing the chunks in order (deterministic).
template <typename T, bool TA, bool TB, int WB, bool EDGE>
__global__ void __launch_bounds__(256) k_gemm_splitk(int64_t m, int64_t n, int64_t k, int64_t kc,
                                                     const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                     int64_t ldb, T* __restrict__ W, TileMap map) {
    int64_t tm, tn;
    map_tile(map, blockIdx.x, tm, tn);
    const int64_t k0 = (int64_t)blockIdx.y * kc;
    const int64_t ks = k - k0 < kc ? k - k0 : kc;
    gemm_tile<T, TA, TB, false, WB, EDGE>(m, n, ks, T(1), A + (TA ? k0 : k0 * lda), lda, B + (TB ? k0 * ldb : k0),
                                          ldb, T(0), W + (int64_t)blockIdx.y * m * n, m, tm, tn);
}

template <typename T>
__global__ void __launch_bounds__(256) k_splitk_sum(const T* __restrict__ W, int nsplit, int64_t m, int64_t n, T alpha,
                                                    T beta, T* __restrict__ C, int64_t ldc) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        T acc = W[t];
        for (int y = 1; y < nsplit; ++y) acc += W[t + y * total];
        const int64_t i = t % m, j = t / m;
        T v = alpha * acc;
        if (beta != T(0)) v += beta * C[i + j * ldc];
        C[i + j * ldc] = v;
    }
}

// FP64 C = alpha A B^T + beta C on 128 x 128 tiles (A m x k and B n x k,
// column-major: op(A)(i, kk) = A[i + kk lda], op(B)(kk, j) = B[j + kk ldb]),
// every tile inside m x n (interior launches; a partial last k-slice is
// zero-padded).  512
// threads as 2 x 4 waves of 64 x 32 (4 x 2 MFMA blocks of 16 x 16 x 4);
// k-slices of 16 in two LDS buffers ([kk][row], a k-row of 128 values is
// 1 KB), each k-row fetched by one direct global-to-LDS load one slice
// ahead; each wave's fragments are double-buffered in registers over the
// 4-wide k steps.  Every entry sums its products in the same order as
// gemm_tile (MFMAs over k in steps of 4): the results are bitwise those of
// the 64 x 64 tiles.  LOWER: entries with row < col are not written.
constexpr int DT = 128, DBK = 16, DPAD = 16;

// mrows: rows of C (and A) that exist; a last tile row past them reads its
// A rows clamped to arow_max (the last even row pair, so that the 16-byte
// loads stay inside A) and writes only rows below mrows.
template <bool LOWER, bool PARTIAL>
__global__ void __launch_bounds__(512) k_gemm_f64_dma(int64_t k, double alpha, const double* __restrict__ A,
                                                      int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                      double beta, double* __restrict__ C, int64_t ldc, TileMap map,
                                                      int64_t mrows, int64_t arow_max) {
    __shared__ double As0[DBK * (DT + DPAD)];
    __shared__ double As1[DBK * (DT + DPAD)];
    __shared__ double Bs0[DBK * (DT + DPAD)];
    __shared__ double Bs1[DBK * (DT + DPAD)];
    int64_t t = blockIdx.x;
    if (map.kind >= 4) {
        t = xcd_tile(t, map.per);
        if (t >= map.ntiles) return;
    }
    int64_t tm, tn;
    map_tile(map, t, tm, tn);
    const int64_t row0 = tm * DT, col0 = tn * DT;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = (wave >> 2) * 64, wc = (wave & 3) * 32;
    // DMA: wave w moves k-rows 2 w and 2 w + 1 of the A and B slices; past
    // k (a partial last slice) the rows are zero, as gemm_tile's guarded
    // loads make them
    auto dma = [&](int64_t k0, double* As, double* Bs) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int kk = 2 * wave + q;
            if (k0 + kk < k) {
                const int64_t ar = PARTIAL ? min(row0 + 2 * lane, arow_max) : row0 + 2 * lane;
                __builtin_amdgcn_global_load_lds((const void*)(A + ar + (k0 + kk) * lda),
                                                 (lds_void*)&As[kk * (DT + DPAD)], 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(B + col0 + (k0 + kk) * ldb + 2 * lane),
                                                 (lds_void*)&Bs[kk * (DT + DPAD)], 16, 0, 0);
            } else {
                As[kk * (DT + DPAD) + 2 * lane] = 0.0;
                As[kk * (DT + DPAD) + 2 * lane + 1] = 0.0;
                Bs[kk * (DT + DPAD) + 2 * lane] = 0.0;
                Bs[kk * (DT + DPAD) + 2 * lane + 1] = 0.0;
            }
        }
    };
    const int fi = lane & 15, fk = lane >> 4;
    // fragments of k step j (4 k values) of a slice: A(i, k) in lane i + 16 k
    double F[2][6];
    auto frags = [&](const double* As, const double* Bs, int j, double (&f)[6]) {
        const int kr = 4 * j + fk;
#pragma unroll
        for (int a = 0; a < 4; ++a) f[a] = As[kr * (DT + DPAD) + wr + 16 * a + fi];
#pragma unroll
        for (int b = 0; b < 2; ++b) f[4 + b] = Bs[kr * (DT + DPAD) + wc + 16 * b + fi];
    };
    double4_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
    const int64_t ns = (k + DBK - 1) / DBK;
    dma(0, As0, Bs0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    frags(As0, Bs0, 0, F[0]);
    for (int64_t s = 0; s < ns; s += 2) {
        // two slices per trip, so that buffers and fragment slots stay static
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (s + h >= ns) break;
            double* Ac = h ? As1 : As0;
            double* Bc = h ? Bs1 : Bs0;
            double* An = h ? As0 : As1;
            double* Bn = h ? Bs0 : Bs1;
            const bool more = s + h + 1 < ns;
            __builtin_amdgcn_sched_barrier(0);
            if (more) dma((s + h + 1) * DBK, An, Bn);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double (&f)[6] = F[j & 1];
                double (&g)[6] = F[(j + 1) & 1];
                if (j == 3) {
                    // the next slice has landed everywhere, and this one is
                    // read everywhere before its buffer is refilled
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                    if (more) frags(An, Bn, 0, g);
                } else {
                    frags(Ac, Bc, j + 1, g);
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[a], f[4 + b], acc[a][b], 0, 0, 0);
                // the next step's fragment reads between this step's MFMAs,
                // after the first two (whose wait covers this step's reads only)
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);   // MFMA
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // DS read
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // epilogue: FP64 accumulator element r of lane is row (lane / 16) + 4 r
    // of column lane % 16 of its block
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gi = row0 + wr + 16 * a + fk + 4 * r, gj = col0 + wc + 16 * b + fi;
                if ((!PARTIAL || gi < mrows) && (!LOWER || gi >= gj)) {
                    double v = alpha * acc[a][b][r];
                    if (beta != 0.0) v += beta * C[gi + gj * ldc];
                    C[gi + gj * ldc] = v;
                }
            }
}

// Grouped GEMM: the tiles of ng independent products (descriptor d[g], its
// tiles numbered from d[g].tile0) in one launch; each workgroup walks the
// tile list with stride gridDim.x (persistent when the grid is capped).
template <typename T, bool TA, bool TB, bool LOWER, int WB, bool EDGE>
__global__ void __launch_bounds__(256) k_gemm_grouped(const GemmDesc* __restrict__ d, int ng, int64_t total) {
    for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = ng - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (d[mid].tile0 <= t) lo = mid;
            else hi = mid - 1;
        }
        const GemmDesc& g = d[lo];
        int64_t tm, tn;
        map_tile(g.map, t - g.tile0, tm, tn);
        __syncthreads();   // the previous tile's LDS reads are done
        gemm_tile<T, TA, TB, LOWER, WB, EDGE>(g.m, g.n, g.k, (T)g.alpha, (const T*)g.A, g.lda, (const T*)g.B, g.ldb,
                                              (T)g.beta, (T*)g.C, g.ldc, tm, tn);
    }
}

// ---------------------------------------------------------------------------
// diagonal block: signed Cholesky A11 = L11 S1 L11^T and Linv = L11^-1.
// Both kernels are latency bound (one NB x NB block, about 11 MFLOP), so
// they are built to keep the serial chain short.
// k_diag_factor: one workgroup; inner panels of IB = 32 columns:
//   1. the panel (rows c0.., IB columns) is read into LDS, transposed;
//   2. one wave factors the IB x IB pivot block in registers (lane i holds
//      the symmetric row i, row j is broadcast with readlane, so the 32
//      pivot steps need no workgroup barrier) and inverts it by forward
//      substitution (lane j holds column j of the inverse Mi);
//   3. the rows below: L_r = A_r Mi^T S1 (one thread per row);
//   4. the rest of the block -= L S L^T (4 x 4 register tiles over the lower
//      triangle; A is read and written in global memory, L2 resident).
// The IB x IB inverses are stored as the diagonal IB-blocks of Linv.
// k_diag_inverse: one workgroup per IB-column block c of Linv; the column
// block is built in LDS by block forward substitution from those inverses:
// X_rc = -Mi_r sum_{c <= q < r} L_rq X_qc.
// ---------------------------------------------------------------------------
constexpr int IB = 32;
constexpr int DF_THREADS = 256;     // factor: one wave per SIMD, so that the
                                    // workgroup fits beside a running GEMM
                                    // workgroup on one CU
constexpr int DIAG_THREADS = 512;   // inverse
constexpr int PT_LD = FACT_NB + 4;   // row stride of the transposed panel (16-byte rows)
constexpr int DT_TILES = 4;          // trailing 4 x 4 tiles per thread and pass

__device__ __forceinline__ float lane_read(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_read(double v, int src) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// four consecutive LDS values (16-byte aligned for float, 32 for double)
__device__ __forceinline__ void lds4(const float* p, float (&v)[4]) {
    const float4 q = *(const float4*)p;
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void lds4(const double* p, double (&v)[4]) {
    const double2 q0 = *(const double2*)p, q1 = *(const double2*)(p + 2);
    v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
}

// Step 2 (wave 0).  Pt[l][i] (l, i < IB) holds the symmetric pivot block,
// padded with the identity beyond w.  Lanes 0..31 hold rows; lanes 32..63
// repeat them and store nothing.  L replaces the pivot block in Pt (row i
// in Pt[.][i]; entries above the diagonal are not meaningful) and is also
// kept row-major in Ls; Mi = L^-1 is stored row-major.
template <typename T>
__device__ __forceinline__ void pivot_block(T (*Pt)[PT_LD], int w, T (*Mi)[IB], T (*Ls)[IB + 4], T* __restrict__ sg,
                                            T* __restrict__ ig, int* __restrict__ info, int64_t kglob, T tiny) {
    const int lane = threadIdx.x & 63;
    int i = lane & (IB - 1);
    asm volatile("" : "+v"(i));   // keeps per-lane selects inside the panel loop
    T a[IB];
#pragma unroll
    for (int l = 0; l < IB; ++l) a[l] = Pt[l][i];
    // right-looking, one pivot per step: with a = column j after the
    // previous steps and d = a_j, A(i,l) -= a_i a_l / d and
    // L(i,j) = a_i / (sqrt|d| s_j), s_j = sign(d)
    // (one reciprocal and one square root per step: 1/(sqrt|d| s) = sqrt|d| / d)
    int bad = -1;
#pragma unroll
    for (int j = 0; j < IB; ++j) {
        const T d = lane_read(a[j], j);
        const T s = d >= T(0) ? T(1) : T(-1);
        const T id = T(1) / d;
        const T r = sqrt(fabs(d));
        const T g = r * id;
        const T f = a[j] * id;
#pragma unroll
        for (int l = j + 1; l < IB; ++l) a[l] -= f * lane_read(a[l], j);
        a[j] = (i == j) ? r : a[j] * g;
        if (lane == 0) ig[j] = s * g;   // 1 / L(j, j)
        if (j < w) {
            if (lane == 0) sg[j] = s;
            if (bad < 0 && !(fabs(d) > tiny)) bad = j;
        }
    }
    // L to LDS (row i of L in Ls[i][.] for the substitution, in Pt[.][i]
    // for the stores)
    if (lane < IB) {
#pragma unroll
        for (int l = 0; l < IB; ++l) {
            Ls[i][l] = a[l];
            Pt[l][i] = a[l];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Mi = L^-1 by forward substitution, lane i holds column i
    T x[IB];
#pragma unroll
    for (int r = 0; r < IB; ++r) {
        T acc0 = (r == i) ? T(1) : T(0), acc1 = T(0);
#pragma unroll
        for (int l = 0; l < r; ++l) {
            if (l & 1) acc1 -= Ls[r][l] * x[l];
            else acc0 -= Ls[r][l] * x[l];
        }
        x[r] = (acc0 + acc1) * ig[r];
    }
    if (lane < IB) {
#pragma unroll
        for (int r = 0; r < IB; ++r) Mi[r][i] = x[r];
    }
    if (lane == 0 && bad >= 0 && info[0] == 0) info[0] = (int)(kglob + bad + 1);
}

template <typename T>
__device__ __forceinline__ void diag_body(T* __restrict__ A11, int64_t lda, int nb, T* __restrict__ sgn,
                                          int* __restrict__ info, int64_t k0, T tiny, T* __restrict__ Linv,
                                          int64_t ldl) {
    __shared__ __align__(16) T Pt[IB][PT_LD];
    __shared__ T Mi[IB][IB];
    __shared__ __align__(16) T Ls[IB][IB + 4];
    __shared__ T sg[IB];
    __shared__ T ig[IB];
    const int tid = threadIdx.x;
    for (int c0 = 0; c0 < nb; c0 += IB) {
        const int w = min(IB, nb - c0);
        const int m = nb - c0;
        // 1. panel into LDS; the pivot block is mirrored (its part above the
        // diagonal in A is stale) and padded with the identity
        for (int e = tid; e < m * w; e += DF_THREADS) {
            const int r = e % m, l = e / m;
            if (r < l) continue;
            const T v = A11[(c0 + r) + (int64_t)(c0 + l) * lda];
            Pt[l][r] = v;
            if (r < w) Pt[r][l] = v;
        }
        if (w < IB) {
            for (int e = tid; e < IB * IB; e += DF_THREADS) {
                const int r = e % IB, l = e / IB;
                if (r >= w || l >= w) Pt[l][r] = (r == l) ? T(1) : T(0);
            }
        }
        __syncthreads();
        // 2. pivot block
        if (tid < 64) pivot_block<T>(Pt, w, Mi, Ls, sg, ig, info, k0 + c0, tiny);
        __syncthreads();
        for (int e = tid; e < IB * IB; e += DF_THREADS) {
            const int r = e % IB, j = e / IB;
            if (r < w && j < w) {
                Linv[(c0 + r) + (int64_t)(c0 + j) * ldl] = Mi[r][j];
                if (r >= j) A11[(c0 + r) + (int64_t)(c0 + j) * lda] = Pt[j][r];
            }
        }
        if (tid < w) sgn[k0 + c0 + tid] = sg[tid];
        if (m <= w) break;   // last inner panel (w == IB otherwise)
        // 3. rows below, one thread per row: L(r, j) = s_j sum_{l <= j}
        // A(r, l) Mi(j, l).  Pt keeps L s.
        const int mt = m - IB;
        for (int rl = tid; rl < mt; rl += DF_THREADS) {
            T row[IB];
#pragma unroll
            for (int l = 0; l < IB; ++l) row[l] = Pt[l][IB + rl];
#pragma unroll 2
            for (int j = 0; j < IB; ++j) {
                T acc0 = 0, acc1 = 0;   // Mi(j, l) = 0 for l > j
#pragma unroll
                for (int l = 0; l < IB; l += 2) {
                    acc0 += row[l] * Mi[j][l];
                    acc1 += row[l + 1] * Mi[j][l + 1];
                }
                const T acc = acc0 + acc1;
                A11[(c0 + IB + rl) + (int64_t)(c0 + j) * lda] = acc * sg[j];
                Pt[j][IB + rl] = acc;
            }
        }
        __syncthreads();
        // 4. trailing lower triangle, rows and columns IB.. of the panel
        const int nt = (mt + 3) >> 2;
        const int ntri = nt * (nt + 1) / 2;
        for (int t0 = 0; t0 < ntri; t0 += DF_THREADS * DT_TILES) {
            T acc[DT_TILES][4][4];
            int R[DT_TILES], C[DT_TILES];
#pragma unroll
            for (int u = 0; u < DT_TILES; ++u) {
                const int t = t0 + u * DF_THREADS + tid;
                int rr = -1, cc = 0;
                if (t < ntri) {
                    // t = q (q + 1) / 2 + p, p <= q: tile (nt-1-p, nt-1-q), so
                    // that consecutive lanes take consecutive row tiles
                    int q = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
                    while (q * (q + 1) / 2 > t) --q;
                    while ((q + 1) * (q + 2) / 2 <= t) ++q;
                    rr = nt - 1 - (t - q * (q + 1) / 2);
                    cc = nt - 1 - q;
                }
                R[u] = rr;
                C[u] = cc;
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int r = 4 * rr + a, c = 4 * cc + b;
                        acc[u][a][b] = (rr >= 0 && r < mt && r >= c)
                                           ? A11[(c0 + IB + r) + (int64_t)(c0 + IB + c) * lda]
                                           : T(0);
                    }
            }
#pragma unroll
            for (int u = 0; u < DT_TILES; ++u) {
                if (R[u] < 0) continue;
#pragma unroll 8
                for (int j = 0; j < IB; ++j) {
                    T x[4], y[4];
                    lds4(&Pt[j][IB + 4 * R[u]], x);
                    lds4(&Pt[j][IB + 4 * C[u]], y);
                    const T s = sg[j];
#pragma unroll
                    for (int b = 0; b < 4; ++b) y[b] *= s;
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) acc[u][a][b] -= x[a] * y[b];
                }
            }
#pragma unroll
            for (int u = 0; u < DT_TILES; ++u) {
                const int rr = R[u], cc = C[u];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int r = 4 * rr + a, c = 4 * cc + b;
                        if (rr >= 0 && r < mt && r >= c)
                            A11[(c0 + IB + r) + (int64_t)(c0 + IB + c) * lda] = acc[u][a][b];
                    }
            }
        }
        __syncthreads();
    }
}

// Column block cb (IB columns) of Linv for the factored nb x nb diagonal
// block (lower triangle of A); the diagonal IB-blocks of Linv are already
// stored.  Rows above the column block are zero.
template <typename T>
__device__ __forceinline__ void diag_inverse_body(const T* __restrict__ A, int64_t lda, int nb, T* __restrict__ Linv,
                                                  int64_t ldl, int cb) {
    __shared__ T Xs[FACT_NB][IB + 1];           // rows c0.. of the column block
    __shared__ T Lr[IB][FACT_NB - IB + 1];      // L(r0 + i, c0 + q), q < r0 - c0
    __shared__ T Mr[IB][IB + 1];                // Mi of block row r0
    __shared__ T Tt[IB][IB + 1];
    const int tid = threadIdx.x;
    const int c0 = cb * IB;
    if (c0 >= nb) return;
    const int wc = min(IB, nb - c0);
    for (int t = tid; t < c0 * wc; t += DIAG_THREADS) {
        const int i = t % c0, j = t / c0;
        Linv[i + (int64_t)(c0 + j) * ldl] = T(0);
    }
    for (int t = tid; t < IB * IB; t += DIAG_THREADS) {
        const int i = t % IB, j = t / IB;
        Xs[i][j] = (i < wc && j < wc) ? Linv[(c0 + i) + (int64_t)(c0 + j) * ldl] : T(0);
    }
    const int ti = tid / (IB / 2), tj = 2 * (tid % (IB / 2));   // 32 rows x 16 column pairs
    for (int r0 = c0 + IB; r0 < nb; r0 += IB) {
        const int rw = min(IB, nb - r0);
        const int K = r0 - c0;
        for (int t = tid; t < IB * K; t += DIAG_THREADS) {
            const int i = t % IB, q = t / IB;
            Lr[i][q] = (i < rw) ? A[(r0 + i) + (int64_t)(c0 + q) * lda] : T(0);
        }
        for (int t = tid; t < IB * IB; t += DIAG_THREADS) {
            const int i = t % IB, j = t / IB;
            Mr[i][j] = (i < rw && j < rw) ? Linv[(r0 + i) + (int64_t)(r0 + j) * ldl] : T(0);
        }
        __syncthreads();
        {
            T a0 = 0, a1 = 0;
            for (int q = 0; q < K; ++q) {
                const T l = Lr[ti][q];
                a0 += l * Xs[q][tj];
                a1 += l * Xs[q][tj + 1];
            }
            Tt[ti][tj] = a0;
            Tt[ti][tj + 1] = a1;
        }
        __syncthreads();
        {
            T a0 = 0, a1 = 0;
            for (int l = 0; l <= ti; ++l) {
                const T mm = Mr[ti][l];
                a0 += mm * Tt[l][tj];
                a1 += mm * Tt[l][tj + 1];
            }
            Xs[K + ti][tj] = -a0;
            Xs[K + ti][tj + 1] = -a1;
        }
        __syncthreads();
    }
    const int rows = nb - c0 - IB;
    for (int t = tid; t < rows * wc; t += DIAG_THREADS) {
        const int i = t % rows, j = t / rows;
        Linv[(c0 + IB + i) + (int64_t)(c0 + j) * ldl] = Xs[IB + i][j];
    }
}

// FP64 variant of diag_inverse_body with a third of the LDS (47.6 KB
// against 142 KB), so that its workgroups start on a CU beside one running
// FP64 trailing-update workgroup (73.7 KB) instead of waiting for an empty
// CU: each workgroup builds 16 columns of Linv (the columns of a forward
// substitution are independent), and the block row L(r0.., c0..r0) is staged
// 32 columns at a time through a buffer that then holds Mi of the block row.
// Every entry sums its products in the same order as diag_inverse_body, so
// the two give the same bits.  Workgroup cb: columns 16 cb .. 16 cb + 15.
constexpr int INV64_CW = 16;
__device__ __forceinline__ void diag_inverse_body_f64(const double* __restrict__ A, int64_t lda, int nb,
                                                      double* __restrict__ Linv, int64_t ldl, int cb) {
    __shared__ double Xs[FACT_NB][INV64_CW + 1];   // rows c0.. of the columns j0 .. j0 + 15
    __shared__ double Ls[IB][IB + 1];              // a 32-column chunk of the block row, then Mi of the row
    __shared__ double Tt[IB][INV64_CW + 1];
    const int tid = threadIdx.x;
    const int j0 = cb * INV64_CW;
    if (j0 >= nb) return;
    const int c0 = j0 / IB * IB;                   // the IB-block the columns lie in
    const int jo = j0 - c0;                        // their offset inside it
    const int wc = min(INV64_CW, nb - j0);
    for (int t = tid; t < c0 * wc; t += DIAG_THREADS) {
        const int i = t % c0, j = t / c0;
        Linv[i + (int64_t)(j0 + j) * ldl] = 0.0;
    }
    const int dw = min(IB, nb - c0);               // rows of the diagonal IB-block
    for (int t = tid; t < IB * INV64_CW; t += DIAG_THREADS) {
        const int i = t % IB, j = t / IB;
        Xs[i][j] = (i < dw && j < wc) ? Linv[(c0 + i) + (int64_t)(j0 + j) * ldl] : 0.0;
    }
    const int ti = tid / INV64_CW, tj = tid % INV64_CW;   // 32 rows x 16 columns
    for (int r0 = c0 + IB; r0 < nb; r0 += IB) {
        const int rw = min(IB, nb - r0);
        const int K = r0 - c0;
        double a = 0.0;
        for (int q0 = 0; q0 < K; q0 += IB) {
            __syncthreads();                       // Xs rows written, the last chunk read
            for (int t = tid; t < IB * IB; t += DIAG_THREADS) {
                const int i = t % IB, q = t / IB;
                Ls[i][q] = (i < rw) ? A[(r0 + i) + (int64_t)(c0 + q0 + q) * lda] : 0.0;
            }
            __syncthreads();
            for (int q = 0; q < IB; ++q) a += Ls[ti][q] * Xs[q0 + q][tj];
        }
        Tt[ti][tj] = a;
        __syncthreads();                           // Tt complete, Ls free
        for (int t = tid; t < IB * IB; t += DIAG_THREADS) {
            const int i = t % IB, j = t / IB;
            Ls[i][j] = (i < rw && j < rw) ? Linv[(r0 + i) + (int64_t)(r0 + j) * ldl] : 0.0;
        }
        __syncthreads();
        double b = 0.0;
        for (int l = 0; l <= ti; ++l) b += Ls[ti][l] * Tt[l][tj];
        Xs[K + ti][tj] = -b;
    }
    __syncthreads();
    const int rows = nb - c0 - IB;
    for (int t = tid; t < rows * wc; t += DIAG_THREADS) {
        const int i = t % rows, j = t / rows;
        Linv[(c0 + IB + i) + (int64_t)(j0 + j) * ldl] = Xs[IB + i][j];
    }
}

// workgroups of the inverse of one diagonal block
template <typename T>
constexpr int inv_blocks() { return sizeof(T) == 8 ? FACT_NB / INV64_CW : FACT_NB / IB; }

template <typename T>
__global__ void __launch_bounds__(DF_THREADS) k_diag_factor(T* __restrict__ A, int64_t lda, int nb,
                                                              T* __restrict__ sgn, int* __restrict__ info, int64_t k0,
                                                              T tiny, T* __restrict__ Linv, int64_t ldl) {
    diag_body<T>(A, lda, nb, sgn, info, k0, tiny, Linv, ldl);
}

template <typename T>
__global__ void __launch_bounds__(DIAG_THREADS) k_diag_inverse(const T* __restrict__ A, int64_t lda, int nb,
                                                               T* __restrict__ Linv, int64_t ldl) {
    if constexpr (sizeof(T) == 8)
        diag_inverse_body_f64((const double*)A, lda, nb, (double*)Linv, ldl, blockIdx.x);
    else
        diag_inverse_body<T>(A, lda, nb, Linv, ldl, blockIdx.x);
}

// The 32-column FP64 inverse (142 KB of LDS), for comparison runs
// (OMH_DIAG_INV64_WIDE=1).
__global__ void __launch_bounds__(DIAG_THREADS) k_diag_inverse_wide64(const double* __restrict__ A, int64_t lda,
                                                                      int nb, double* __restrict__ Linv, int64_t ldl) {
    diag_inverse_body<double>(A, lda, nb, Linv, ldl, blockIdx.x);
}

// One workgroup per diagonal block of several factorisations.
template <typename T>
__global__ void __launch_bounds__(DF_THREADS) k_diag_grouped(const DiagDesc* __restrict__ d, T tiny) {
    const DiagDesc& g = d[blockIdx.x];
    diag_body<T>((T*)g.A11, g.lda, g.nb, (T*)g.sgn, g.info, g.k0, tiny, (T*)g.Linv, FACT_NB);
}

// Column block blockIdx.x of the inverse of diagonal block blockIdx.y.
template <typename T>
__global__ void __launch_bounds__(DIAG_THREADS) k_diag_inverse_grouped(const DiagDesc* __restrict__ d) {
    const DiagDesc& g = d[blockIdx.y];
    if constexpr (sizeof(T) == 8)
        diag_inverse_body_f64((const double*)g.A11, g.lda, g.nb, (double*)g.Linv, FACT_NB, blockIdx.x);
    else
        diag_inverse_body<T>((const T*)g.A11, g.lda, g.nb, (T*)g.Linv, FACT_NB, blockIdx.x);
}

// Element-wise operations of several matrices (group = blockIdx.y).
__global__ void __launch_bounds__(256) k_elem_grouped(const ElemDesc* __restrict__ d) {
    const ElemDesc& g = d[blockIdx.y];
    const int64_t total = g.m * g.n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / g.m, i = t - j * g.m;
        const int64_t si = i + j * g.lds, di = i + j * g.ldd;
        switch (g.op) {
            case ELEM_COPY_F64: ((double*)g.dst)[di] = ((const double*)g.src)[si]; break;
            case ELEM_COPY_F32: ((float*)g.dst)[di] = ((const float*)g.src)[si]; break;
            case ELEM_F64_TO_F32: ((float*)g.dst)[di] = (float)((const double*)g.src)[si]; break;
            case ELEM_F64_TO_F32_LOWER:
                if (i >= j) ((float*)g.dst)[di] = (float)((const double*)g.src)[si];
                break;
            case ELEM_ADD_F32_TO_F64: ((double*)g.dst)[di] += (double)((const float*)g.src)[si]; break;
            case ELEM_SCALE_ROWS_F32: ((float*)g.dst)[di] *= ((const float*)g.aux)[i]; break;
            case ELEM_SCALE_ROWS_F64: ((double*)g.dst)[di] *= ((const double*)g.aux)[i]; break;
            case ELEM_SCALE_COLS_F32:
                ((float*)g.dst)[di] = ((const float*)g.src)[si] * ((const float*)g.aux)[j];
                break;
            case ELEM_SCALE_COLS_F64:
                ((double*)g.dst)[di] = ((const double*)g.src)[si] * ((const double*)g.aux)[j];
                break;
            case ELEM_ZERO_F64: ((double*)g.dst)[di] = 0.0; break;
            case ELEM_ZERO_INT: ((int*)g.dst)[di] = 0; break;
            default: break;
        }
    }
}

// L21 = X * S1 (column scaling by the block's signs) written into A21.
template <typename T>
__global__ void k_scale_cols(const T* __restrict__ X, int64_t ldx, T* __restrict__ Aout, int64_t lda, int64_t m,
                             int nb, const T* __restrict__ s) {
    const int64_t total = m * nb;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m;
        Aout[i + j * lda] = X[i + j * ldx] * s[j];
    }
}

template <typename T>
__global__ void k_copy(const T* __restrict__ src, int64_t lds, T* __restrict__ dst, int64_t ldd, int64_t m,
                       int64_t n) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m;
        dst[i + j * ldd] = src[i + j * lds];
    }
}

template <typename T>
__global__ void k_scale_rows(T* __restrict__ W, int64_t ldw, int64_t m, int64_t n, const T* __restrict__ s) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m;
        W[i + j * ldw] *= s[i];
    }
}

// dst = src (m x n), or only its lower triangle (rows i >= j of column j):
// columns over grid.y (strided), rows over grid.x and the threads.
template <typename TS, typename TD>
__global__ void k_convert(const TS* __restrict__ src, int64_t lds, TD* __restrict__ dst, int64_t ldd, int64_t m,
                          int64_t n, int lower_only) {
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
        const TS* s = src + j * lds;
        TD* d = dst + j * ldd;
        const int64_t i0 = lower_only ? j : 0;
        for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
             i += (int64_t)gridDim.x * blockDim.x)
            d[i] = (TD)s[i];
    }
}

// x += d; with a row map, row i of d updates row rowmap[i] of x
__global__ void k_axpy_f32_to_f64(const float* __restrict__ d, int64_t ldd, double* __restrict__ x, int64_t ldx,
                                  int64_t m, int64_t n, const int32_t* __restrict__ rowmap) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m;
        const int64_t xi = rowmap ? (int64_t)rowmap[i] : i;
        x[xi + j * ldx] += (double)d[i + j * ldd];
    }
}

// Segment reorder.  rowmap[i] = the row of the original matrix that is row
// i of the reordered one (whole segments move, so consecutive i read
// consecutive rows inside a segment).
// dst's lower triangle (i >= j) = src(rowmap[i], rowmap[j]) of the full
// symmetric src: columns over grid.y, rows over grid.x and the threads
template <typename TD>
__global__ void k_sym_gather(const double* __restrict__ src, int64_t lds, TD* __restrict__ dst, int64_t ldd, int64_t n,
                             const int32_t* __restrict__ rowmap) {
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
        const double* s = src + (int64_t)rowmap[j] * lds;
        TD* d = dst + j * ldd;
        for (int64_t i = j + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
            d[i] = (TD)s[rowmap[i]];
    }
}

// The same reorder as block copies: block b of the reordered lower triangle
// (rows [dr, dr + h), columns [dc, dc + w) of dst; the diagonal blocks only
// their lower triangle) is the block of src at rows sr, columns sc: rows are
// contiguous on both sides, no per-element row map.  Blocks over grid.z,
// their columns over grid.y (strided), rows over grid.x and the threads.
constexpr int SG_MAX_BLOCKS = 64;
struct SymGatherBlocks {
    int64_t dr[SG_MAX_BLOCKS], dc[SG_MAX_BLOCKS], sr[SG_MAX_BLOCKS], sc[SG_MAX_BLOCKS];
    int32_t h[SG_MAX_BLOCKS], w[SG_MAX_BLOCKS];
    int32_t diag[SG_MAX_BLOCKS];
};
template <typename TD>
__global__ void k_sym_gather_blocks(const double* __restrict__ src, int64_t lds, TD* __restrict__ dst, int64_t ldd,
                                    SymGatherBlocks tb) {
    const int b = blockIdx.z;
    const int64_t h = tb.h[b], w = tb.w[b];
    for (int64_t j = blockIdx.y; j < w; j += gridDim.y) {
        const double* sp = src + tb.sr[b] + (tb.sc[b] + j) * lds;
        TD* dp = dst + tb.dr[b] + (tb.dc[b] + j) * ldd;
        const int64_t i0 = tb.diag[b] ? j : 0;
        for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < h;
             i += (int64_t)gridDim.x * blockDim.x)
            dp[i] = (TD)sp[i];
    }
}

// m x n rows: gather dst(i, :) = src(rowmap[i], :), or scatter
// dst(rowmap[i], :) = src(i, :)
template <typename TS, typename TD>
__global__ void k_rows_map(const TS* __restrict__ src, int64_t lds, TD* __restrict__ dst, int64_t ldd, int64_t m,
                           int64_t n, const int32_t* __restrict__ rowmap, int scatter) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m, r = rowmap[i];
        if (scatter) dst[r + j * ldd] = (TD)src[i + j * lds];
        else dst[i + j * ldd] = (TD)src[r + j * lds];
    }
}

// partial sums of squares, one per block
template <typename T>
__global__ void __launch_bounds__(256) k_sumsq(const T* __restrict__ a, int64_t lda, int64_t m, int64_t n,
                                               double* __restrict__ part) {
    const int64_t total = m * n;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / m, i = t - j * m;
        const double v = (double)a[i + j * lda];
        acc += v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// EH(e, :) = sum_k vals(e,k) * Hinv(cols(e,k), :) -- the sparse x dense
// product of sparse_matrix.cpp:36-51 for a 3-nnz-per-row Head2EEGMat.
// Column e of out (zeroed before) gets vals(e, k) at row rows(e, k), k < 3
// (one thread per column, so repeated rows add without atomics).
__global__ void k_sparse_cols(const int64_t* __restrict__ rows, const double* __restrict__ vals, int64_t ne, int64_t n,
                              double* __restrict__ out, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    for (int k = 0; k < 3; ++k) {
        const int64_t r = rows[3 * e + k];
        if (r >= 0 && r < n) out[r + e * ldo] += vals[3 * e + k];
    }
}

__global__ void k_sparse_rows(const int64_t* __restrict__ cols, const double* __restrict__ vals, int64_t ne,
                              const double* __restrict__ Hm, int64_t ldh, int64_t n, double* __restrict__ out,
                              int64_t ldo) {
    const int64_t total = ne * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t % ne, j = t / ne;
        double v = 0.0;
        for (int k = 0; k < 3; ++k) v += vals[3 * e + k] * Hm[cols[3 * e + k] + j * ldh];
        out[e + j * ldo] = v;
    }
}

// ---------------------------------------------------------------------------
// Split-FP16 trailing update.  The SYRK C -= X S X^T of a factor step (X the
// S_PANELS panels of a block group, S_KP columns) runs on the FP16 MFMA
// (v_mfma_f32_16x16x32_f16, FP32 accumulation): row i of X is scaled by a
// power of two s_i >= max_k |X_ik| and split into FP16 parts,
// X_ik / s_i = h + l to within 2^-22 of the row maximum, and
//   C_ij -= s_i s_j (Ah_i . Bh_j + Ah_i . Bl_j + Al_i . Bh_j),
// with B = X S (the signs folded into the B parts).  Operands are stored
// slice-major: element (row, k) at ((k / SK) mp + row) SK + k % SK, so a
// 32-wide k-slice of 128 rows is 8 KB of contiguous memory.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
constexpr int SK = 32;        // k-slice of the FP16 MFMA
constexpr int S_TILE = 128;   // output tile of the split SYRK
// Panels per split update (its rank / FACT_NB).  n = 102400: 2 panels
// 1.48 s, 3 panels 1.51 s, 4 panels 1.34 s for the whole factorisation
// (larger groups halve the trailing-matrix traffic again but lengthen the
// group prologue).
#ifndef OMH_SPLIT_PANELS
#define OMH_SPLIT_PANELS 4
#endif
constexpr int S_PANELS = OMH_SPLIT_PANELS;
constexpr int S_KP = S_PANELS * FACT_NB;      // k extent of the operands

// The S_PANELS consecutive panels of a split update (FP32, their rows
// starting at the update's first row at offsets off[]).
struct SplitSrc {
    const float* X[S_PANELS];
    int64_t ldx[S_PANELS];
    int64_t off[S_PANELS];
    const float* sgn[S_PANELS];
};

// Operands of the split update: row i gets one power-of-two scale over all
// S_PANELS panels; k in [256 p, 256 p + 256) comes from panel p.  64 rows per
// block, thread (row t % 64, 64-wide k group t / 64).  Padding rows
// (m <= i < mp) are zero.
__global__ void __launch_bounds__(64 * 4 * S_PANELS) k_split_group(SplitSrc src, int64_t m, int64_t mp,
                                                                  _Float16* __restrict__ Ah,
                                                                  _Float16* __restrict__ Al,
                                                                  _Float16* __restrict__ Bh,
                                                                  _Float16* __restrict__ Bl,
                                                                  float* __restrict__ scale) {
    __shared__ float red[4 * S_PANELS][64];
    const int ri = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * 64 + ri;
    const int pn = kq >> 2;                 // panel
    const float* X = src.X[pn] + src.off[pn];
    const int64_t ldx = src.ldx[pn];
    const float* sg = src.sgn[pn];
    const int kb = (kq & 3) * 64;           // first column inside its panel
    float v[64];
    float mx = 0.f;
#pragma unroll
    for (int kk = 0; kk < 64; ++kk) {
        v[kk] = row < m ? X[row + (int64_t)(kb + kk) * ldx] : 0.f;
        mx = fmaxf(mx, fabsf(v[kk]));
    }
    red[kq][ri] = mx;
    __syncthreads();
    mx = 0.f;
#pragma unroll
    for (int g = 0; g < 4 * S_PANELS; ++g) mx = fmaxf(mx, red[g][ri]);
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);   // mx = f 2^e, f in [0.5, 1)
    const float inv = ldexpf(1.f, -e);
    if (kq == 0) scale[row] = ldexpf(1.f, e);
#pragma unroll
    for (int kk = 0; kk < 64; kk += 8) {
        half8_t h, l, bh, bl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float x = v[kk + q] * inv;
            const _Float16 hx = (_Float16)x;
            const _Float16 lx = (_Float16)(x - (float)hx);
            const bool neg = sg[kb + kk + q] < 0.f;
            h[q] = hx;
            l[q] = lx;
            bh[q] = neg ? -hx : hx;
            bl[q] = neg ? -lx : lx;
        }
        const int k0 = kq * 64 + kk;
        const int64_t off = ((int64_t)(k0 / SK) * mp + row) * SK + (k0 % SK);
        *reinterpret_cast<half8_t*>(Ah + off) = h;
        *reinterpret_cast<half8_t*>(Al + off) = l;
        *reinterpret_cast<half8_t*>(Bh + off) = bh;
        *reinterpret_cast<half8_t*>(Bl + off) = bl;
    }
}

// LDS position of 16-byte chunk c (of 4 per 32-wide row) of tile row r: the
// chunk index is XOR-ed with ((r / 2) ^ (r / 4)) % 4 so that the rows of a
// fragment read hit disjoint banks whether the 16-byte reads are served 16
// lanes over 64 banks or 8 lanes over 32 banks.
__device__ __forceinline__ int swz(int r, int c) { return r * 4 + (c ^ (((r >> 1) ^ (r >> 2)) & 3)); }

// Slice clock of the split update (k-synchronised schedule): tiles that run
// at the same time on an XCD share operand bands, but each walks its k
// slices from its own start, so with the k extent of 1024 the bands of the
// 64 tiles in flight (16 bands of 512 KB) overflow the XCD's 4 MB L2.  With
// the clock on, a tile starts at the slice (realtime / period) mod NKS and
// wraps around, so that every tile in flight works on about the same slice
// and a band's slice is fetched once for all of them.  period[0] holds the
// period in ticks of the 100 MHz realtime counter (0: every tile starts at
// slice 0); workgroup 0 writes the period it measured on its tile to
// period_out (the next launch's period[0]).  The products summed per entry
// are the same; their order follows the clock.
struct SliceClock {
    const unsigned* period = nullptr;
    unsigned* period_out = nullptr;
};

// C (m x m, leading dimension ldc) -= split product of operand rows
// r_off + [0, m), over k extent KP, on tile (tm, tn) of C (tm >= tn);
// entries with row < col or an index >= m are not written.  256 threads, a
// 128 x 128 tile, 2 x 2 waves of 64 x 64 (4 x 4 MFMA blocks, 3 products
// each), k-slices double-buffered in LDS.  Operand rows up to
// r_off + 128 * ceil(m / 128) exist (zero padding), so loads need no guards.
template <int KP>
__device__ __forceinline__ void syrk_split_tile(int64_t m, int64_t mp, int64_t r_off, const _Float16* __restrict__ Ah,
                                                const _Float16* __restrict__ Al, const _Float16* __restrict__ Bh,
                                                const _Float16* __restrict__ Bl, const float* __restrict__ scale,
                                                float* __restrict__ C, int64_t ldc, int64_t tm, int64_t tn,
                                                SliceClock clk = SliceClock{}) {
    __shared__ half8_t S[2][4][S_TILE * 4];
    __shared__ int rot_sh;
    const int64_t row0 = r_off + tm * S_TILE, col0 = r_off + tn * S_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const _Float16* src[4] = {Ah, Al, Bh, Bl};
    constexpr int NKS = KP / SK;
    static_assert((NKS & (NKS - 1)) == 0, "the slice clock wraps k with a mask");
    // first slice of this tile (the slice clock, read once for the workgroup)
    uint64_t t_start = 0;
    int rot = 0;
    if (clk.period) {
        if (tid == 0) {
            t_start = __builtin_amdgcn_s_memrealtime();
            const unsigned per = clk.period[0];
            rot_sh = per ? (int)((t_start / per) & (uint64_t)(NKS - 1)) : 0;
        }
        __syncthreads();
        rot = rot_sh;
    }
    // operand slices travel global -> registers (two slots, loaded two
    // slices ahead) -> LDS (two buffers)
    half8_t stg[2][4][2];
    auto gload = [&](int ks, int slot) {
        const int64_t kk = (int64_t)((ks + rot) & (NKS - 1));
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int q = tid + 256 * p, r = q >> 2, c = q & 3;
                const int64_t base = a < 2 ? row0 : col0;
                stg[slot][a][p] =
                    *reinterpret_cast<const half8_t*>(src[a] + ((kk * mp + base + r) * SK + 8 * c));
            }
    };
    auto lstore = [&](int slot, int buf) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int q = tid + 256 * p, r = q >> 2, c = q & 3;
                S[buf][a][swz(r, c)] = stg[slot][a][p];
            }
    };
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int64_t i0 = tm * S_TILE, j0 = tn * S_TILE;   // tile origin in C
    const bool vec = tm != tn && i0 + S_TILE <= m && (ldc % 4 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0);
    float4 c_pre[4][4], si_pre[4];
    const int fr = lane & 15, fc = lane >> 4;
    gload(0, 0);
    lstore(0, 0);
    gload(1, 1);
    gload(2, 0);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {   // unrolled: register slots stay static
        const int buf = ks & 1;
        half8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int r = wm + 16 * a + fr;
            ah[a] = S[buf][0][swz(r, fc)];
            al[a] = S[buf][1][swz(r, fc)];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = wn + 16 * b + fr;
            bh[b] = S[buf][2][swz(r, fc)];
            bl[b] = S[buf][3][swz(r, fc)];
        }
        // the three products, each over all 16 accumulators in turn
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        // slice ks + 1 (in register slot (ks + 1) % 2) into the other buffer,
        // then refill that slot with slice ks + 3
        if (ks + 1 < NKS) lstore((ks + 1) & 1, buf ^ 1);
        __syncthreads();
        if (ks + 3 < NKS) gload(ks + 3, (ks + 1) & 1);
        if (ks == NKS - 2 && vec) {
            // the operand registers are free now: fetch this lane's part of
            // the C tile and the row scales, consumed after the last slice
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                si_pre[a] = *reinterpret_cast<const float4*>(scale + r_off + i0 + wm + 16 * a + 4 * fc);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    c_pre[a][b] = *reinterpret_cast<const float4*>(
                        C + (i0 + wm + 16 * a + 4 * fc) + (j0 + wn + 16 * b + fr) * ldc);
            }
        }
    }
    // epilogue: lane holds rows 4 (lane / 16) + r of column lane % 16 of a block
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t gi = i0 + wm + 16 * a + 4 * fc, gj = j0 + wn + 16 * b + fr;
            const float sj = scale[r_off + gj];
            if (vec) {
                const float4 si = si_pre[a];
                float4* cp = reinterpret_cast<float4*>(C + gi + gj * ldc);
                float4 c = c_pre[a][b];
                c.x -= si.x * sj * acc[a][b][0];
                c.y -= si.y * sj * acc[a][b][1];
                c.z -= si.z * sj * acc[a][b][2];
                c.w -= si.w * sj * acc[a][b][3];
                *cp = c;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = gi + r;
                    if (i < m && gj < m && i >= gj) C[i + gj * ldc] -= scale[r_off + i] * sj * acc[a][b][r];
                }
            }
        }
    if (clk.period_out && blockIdx.x == 0 && tid == 0) {   // this tile's time per slice, for the next launch
        const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t_start;
        clk.period_out[0] = (unsigned)max((uint64_t)1, min(dt / NKS, (uint64_t)1000000));
    }
}

// The same tile with a software pipeline: operand slices go global -> LDS
// directly (global_load_lds, 16 bytes per lane, two LDS buffers) and each
// wave's fragments are double-buffered in registers, so that a wave reads
// slice ks + 1 from LDS while its MFMAs work on slice ks, and one barrier per
// slice separates the buffers' reuse.  Slice ks + 2 is fetched into the
// buffer that slice ks left.  Every entry sums the same products in the same
// order as syrk_split_tile: the results are bitwise equal.
template <int KP>
__device__ __forceinline__ void syrk_split_tile_pp(int64_t m, int64_t mp, int64_t r_off,
                                                   const _Float16* __restrict__ Ah, const _Float16* __restrict__ Al,
                                                   const _Float16* __restrict__ Bh, const _Float16* __restrict__ Bl,
                                                   const float* __restrict__ scale, float* __restrict__ C,
                                                   int64_t ldc, int64_t tm, int64_t tn, SliceClock clk = SliceClock{}) {
    // two separate arrays, so that the compiler can tell a DMA into one
    // from fragment reads of the other
    __shared__ half8_t S0[4][S_TILE * 4];
    __shared__ half8_t S1[4][S_TILE * 4];
    __shared__ int rot_sh;
    const int64_t row0 = r_off + tm * S_TILE, col0 = r_off + tn * S_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    constexpr int NKS = KP / SK;
    static_assert(NKS >= 2 && (NKS & (NKS - 1)) == 0, "the slice clock wraps k with a mask");
    uint64_t t_start = 0;
    int rot = 0;
    if (clk.period) {
        if (tid == 0) {
            t_start = __builtin_amdgcn_s_memrealtime();
            const unsigned per = clk.period[0];
            rot_sh = per ? (int)((t_start / per) & (uint64_t)(NKS - 1)) : 0;
        }
        __syncthreads();
        rot = __builtin_amdgcn_readfirstlane(rot_sh);
    }
    // DMA: wave w moves rows 32 w + 16 p + lane / 4 (p = 0, 1) of each of
    // the four arrays; lane L lands in LDS slot 64 (8 w + 4 p) + L, so it
    // fetches the 16-byte chunk that the swizzle puts there (the swizzle
    // depends on the row's low 4 bits only)
    const int rl = lane >> 2;
    const uint32_t goff = (uint32_t)(((32 * wave + rl) * SK + 8 * ((lane & 3) ^ (((rl >> 1) ^ (rl >> 2)) & 3))) * 2);
    auto dma = [&](int ks, half8_t (*Sb)[S_TILE * 4]) {
        const int64_t so = (int64_t)((ks + rot) & (NKS - 1)) * mp * SK;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const char* base = reinterpret_cast<const char*>((a == 0 ? Ah : a == 1 ? Al : a == 2 ? Bh : Bl) + so +
                                                             (a < 2 ? row0 : col0) * SK);
#pragma unroll
            for (int p = 0; p < 2; ++p)
                __builtin_amdgcn_global_load_lds((const void*)(base + goff + p * 16 * SK * 2),
                                                 (lds_void*)&Sb[a][(32 * wave + 16 * p) * 4], 16, 0, 0);
        }
    };
    const int fr = lane & 15, fc = lane >> 4;
    const int lo = swz(fr, fc);   // slot of (row r0 + fr, chunk fc) is r0 * 4 + lo for r0 % 16 == 0
    // fragments: [0, 4) ah, [4, 8) al, [8, 12) bh, [12, 16) bl
    half8_t F[2][16];
    auto frags = [&](const half8_t (*Sb)[S_TILE * 4], half8_t (&f)[16]) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            f[a] = Sb[0][(wm + 16 * a) * 4 + lo];
            f[4 + a] = Sb[1][(wm + 16 * a) * 4 + lo];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            f[8 + b] = Sb[2][(wn + 16 * b) * 4 + lo];
            f[12 + b] = Sb[3][(wn + 16 * b) * 4 + lo];
        }
    };
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int64_t i0 = tm * S_TILE, j0 = tn * S_TILE;   // tile origin in C
    const bool vec = tm != tn && i0 + S_TILE <= m && (ldc % 4 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0);
    float4 c_pre[4][4], si_pre[4];
    // prologue: slices 0 and 1 in flight, slice 0's fragments in registers
    dma(0, S0);
    dma(1, S1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    frags(S0, F[0]);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();   // buffer 0 is free for slice 2
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
        half8_t (&f)[16] = F[ks & 1];
        half8_t (&g)[16] = F[(ks + 1) & 1];
        // MFMAs touch no memory: keep each slice's products inside its own
        // region, after the barrier that ends the previous one
        __builtin_amdgcn_sched_barrier(0);
        if (ks + 2 < NKS) dma(ks + 2, (ks & 1) ? S1 : S0);
        if (ks + 1 < NKS) {
            frags((ks & 1) ? S0 : S1, g);
        } else if (vec) {
            // last slice: the C tile and the row scales take the registers of
            // the next slice's fragments
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                si_pre[a] = *reinterpret_cast<const float4*>(scale + r_off + i0 + wm + 16 * a + 4 * fc);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    c_pre[a][b] = *reinterpret_cast<const float4*>(
                        C + (i0 + wm + 16 * a + 4 * fc) + (j0 + wn + 16 * b + fr) * ldc);
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[4 + a], f[8 + b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[a], f[12 + b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[a], f[8 + b], acc[a][b], 0, 0, 0);
        // fragment reads one per three products, the slice fetch first
        if (ks + 1 < NKS) {
            if (ks + 2 < NKS) __builtin_amdgcn_sched_group_barrier(0x020, 8, 0);   // VMEM read (DMA)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // DS read
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);   // MFMA
            }
        }
        if (ks + 2 < NKS) {
            // slice ks + 1's fragments are read everywhere and slice ks + 2
            // has landed everywhere before either buffer changes role
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
    // epilogue: lane holds rows 4 (lane / 16) + r of column lane % 16 of a block
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t gi = i0 + wm + 16 * a + 4 * fc, gj = j0 + wn + 16 * b + fr;
            const float sj = scale[r_off + gj];
            if (vec) {
                const float4 si = si_pre[a];
                float4* cp = reinterpret_cast<float4*>(C + gi + gj * ldc);
                float4 c = c_pre[a][b];
                c.x -= si.x * sj * acc[a][b][0];
                c.y -= si.y * sj * acc[a][b][1];
                c.z -= si.z * sj * acc[a][b][2];
                c.w -= si.w * sj * acc[a][b][3];
                *cp = c;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = gi + r;
                    if (i < m && gj < m && i >= gj) C[i + gj * ldc] -= scale[r_off + i] * sj * acc[a][b][r];
                }
            }
        }
    if (clk.period_out && blockIdx.x == 0 && tid == 0) {   // this tile's time per slice, for the next launch
        const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t_start;
        clk.period_out[0] = (unsigned)max((uint64_t)1, min(dt / NKS, (uint64_t)1000000));
    }
}

// Wide tiles of the split update: 256 rows x 128 columns of C, 512 threads
// as 4 x 2 waves of 64 x 64 (the same wave tile, products and summation
// order as syrk_split_tile, so the factor is bitwise the same), one
// workgroup per CU.  Tile coordinates are in units of 128 rows of the
// (possibly remapped) trailing matrix: column unit tb, and the row pair
// (2 ta, 2 ta + 1), each unit remapped on its own (rows are contiguous
// within a unit).  A slice (32 k) of the tile's operands is 48 KB of LDS,
// rows stacked as [Ah 256 | Al 256 | Bh 128 | Bl 128], chunks swizzled as
// in swz(); three slices live in a ring of buffers, so a slice is fetched by
// direct global-to-LDS loads two slices before it is read, and each wave's
// fragments are double-buffered in registers.  The barrier that closes a
// slice waits only for the loads of the next slice (the s_barrier is issued
// on its own: a workgroup fence would also wait for the loads in flight).
constexpr int WROWS = 256, WCOLS = 128, WSTACK = 2 * WROWS + 2 * WCOLS;

// Tiles (ta, tb) with tb <= min(2 ta + 1, ntm - 1) in bands of 4 row pairs
// (8 units): the band's columns left of its diagonal block, column by
// column, then its diagonal block column by column.  Full bands before band
// b hold 16 b^2 + 4 b tiles.
__device__ __forceinline__ void wide_band_tile(int64_t t, int64_t ntm, int64_t& ta, int64_t& tb) {
    const int64_t nta = (ntm + 1) / 2, nfull = nta / 4;
    int64_t b = (int64_t)((sqrt(16.0 + 64.0 * (double)t) - 4.0) / 32.0);
    while (b > 0 && 16 * b * b + 4 * b > t) --b;
    while (16 * (b + 1) * (b + 1) + 4 * (b + 1) <= t) ++b;
    if (b > nfull) b = nfull;
    const int64_t h = min((int64_t)4, nta - 4 * b);
    int64_t o = t - (16 * b * b + 4 * b);
    if (o < 8 * b * h) {
        tb = o / h;
        ta = 4 * b + o % h;
        return;
    }
    o -= 8 * b * h;
    for (int j = 0; j < 8; ++j) {   // column 8 b + j: rows a >= ceil((j - 1) / 2) of the band
        const int64_t c = 8 * b + j;
        if (c >= ntm) break;
        const int a0 = j <= 1 ? 0 : j / 2;
        const int64_t cnt = h - a0 > 0 ? h - a0 : 0;
        if (o < cnt) {
            tb = c;
            ta = 4 * b + a0 + o;
            return;
        }
        o -= cnt;
    }
    ta = tb = -1;   // past the last tile
}

template <int KP>
__global__ void __launch_bounds__(512) k_syrk_split_wide(int64_t m, int64_t mp, int64_t r_off,
                                                         const _Float16* __restrict__ Ah,
                                                         const _Float16* __restrict__ Al,
                                                         const _Float16* __restrict__ Bh,
                                                         const _Float16* __restrict__ Bl,
                                                         const float* __restrict__ scale, float* __restrict__ C,
                                                         int64_t ldc, int64_t ntm, int64_t ntiles, int64_t per,
                                                         TileRemap rm, SliceClock clk) {
    __shared__ half8_t S0[WSTACK * 4];
    __shared__ half8_t S1[WSTACK * 4];
    __shared__ half8_t S2[WSTACK * 4];
    __shared__ int rot_sh;
    const int64_t t = xcd_tile(blockIdx.x, per);
    if (t >= ntiles) return;
    int64_t ta, tb;
    wide_band_tile(t, ntm, ta, tb);
    if (ta < 0) return;
    // units (remapped): the row pair and the column; a second row unit past
    // the last one is not read (its rows alias the first unit's) nor written
    const bool h1 = 2 * ta + 1 < ntm;
    const int64_t u0 = rm.n ? remap_tile(rm, 2 * ta) : 2 * ta;
    const int64_t u1 = h1 ? (rm.n ? remap_tile(rm, 2 * ta + 1) : 2 * ta + 1) : u0;
    const int64_t uc = rm.n ? remap_tile(rm, tb) : tb;
    // the wave index is uniform: the DMA rows and their arrays stay scalar
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    constexpr int NKS = KP / SK;
    static_assert(NKS >= 3 && (NKS & (NKS - 1)) == 0, "the slice clock wraps k with a mask");
    uint64_t t_start = 0;
    int rot = 0;
    if (clk.period) {
        if (tid == 0) {
            t_start = __builtin_amdgcn_s_memrealtime();
            const unsigned pd = clk.period[0];
            rot_sh = pd ? (int)((t_start / pd) & (uint64_t)(NKS - 1)) : 0;
        }
        __syncthreads();
        rot = __builtin_amdgcn_readfirstlane(rot_sh);
    }
    // DMA: wave w moves 16-row groups of each array, rows 32 w + 16 q
    // (q < 2) of Ah and Al and rows 16 w of Bh and Bl; lane L lands in slot
    // 4 R + L (R the group's first stacked row), so it fetches the 16-byte
    // chunk that the swizzle puts there (the swizzle depends on the row's
    // low 4 bits only)
    const int rl = lane >> 2;
    const uint32_t goff = (uint32_t)((rl * SK + 8 * ((lane & 3) ^ (((rl >> 1) ^ (rl >> 2)) & 3))) * 2);
    const int64_t arow = (wave < 4 ? u0 : u1) * 128 + 32 * (wave & 3), brow = uc * 128 + 16 * wave;
    auto dma = [&](int ks, half8_t* Sb) {
        const int64_t so = (int64_t)((ks + rot) & (NKS - 1)) * mp * SK;
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const _Float16* src = p < 2 ? Ah : p < 4 ? Al : p == 4 ? Bh : Bl;
            const int64_t row = p < 4 ? arow + 16 * (p & 1) : brow;
            const int R = p < 2 ? 32 * wave + 16 * p : p < 4 ? WROWS + 32 * wave + 16 * (p - 2)
                                                     : p == 4 ? 2 * WROWS + 16 * wave : 2 * WROWS + WCOLS + 16 * wave;
            const char* base = reinterpret_cast<const char*>(src + so + (r_off + row) * SK);
            __builtin_amdgcn_global_load_lds((const void*)(base + goff), (lds_void*)&Sb[R * 4], 16, 0, 0);
        }
    };
    const int fr = lane & 15, fc = lane >> 4;
    const int lo = swz(fr, fc);
    half8_t F[2][16];
    auto frags = [&](const half8_t* Sb, half8_t (&f)[16]) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            f[a] = Sb[(wm + 16 * a) * 4 + lo];
            f[4 + a] = Sb[(WROWS + wm + 16 * a) * 4 + lo];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            f[8 + b] = Sb[(2 * WROWS + wn + 16 * b) * 4 + lo];
            f[12 + b] = Sb[(2 * WROWS + WCOLS + wn + 16 * b) * 4 + lo];
        }
    };
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    // this wave's 64 rows lie in one unit; C origin of the wave's block
    const bool wvalid = wm < 128 || h1;
    const int64_t gi0 = (wm < 128 ? u0 : u1) * 128 + (wm & 127), gj0 = uc * 128 + wn;
    const bool vec = gi0 >= gj0 + 64 && gi0 + 64 <= m && (ldc % 4 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0);
    float4 c_pre[4][4], si_pre[4];
    auto ring = [&](int i) -> half8_t* { return i == 0 ? S0 : i == 1 ? S1 : S2; };
    dma(0, S0);
    dma(1, S1);
    dma(2, S2);
    __builtin_amdgcn_s_waitcnt(0x7C);   // vmcnt(12) lgkmcnt(0): slice 0 has landed
    asm volatile("s_barrier" ::: "memory");
    frags(S0, F[0]);
    __builtin_amdgcn_s_waitcnt(0x76);   // vmcnt(6): slice 1 too; slice 0's fragments are read
    asm volatile("s_barrier" ::: "memory");
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
        half8_t (&f)[16] = F[ks & 1];
        half8_t (&g)[16] = F[(ks + 1) & 1];
        __builtin_amdgcn_sched_barrier(0);
        if (ks + 3 < NKS) dma(ks + 3, ring(ks % 3));
        if (ks + 1 < NKS) {
            frags(ring((ks + 1) % 3), g);
        } else if (vec && wvalid) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                si_pre[a] = *reinterpret_cast<const float4*>(scale + r_off + gi0 + 16 * a + 4 * fc);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    c_pre[a][b] = *reinterpret_cast<const float4*>(C + (gi0 + 16 * a + 4 * fc) +
                                                                   (gj0 + 16 * b + fr) * ldc);
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[4 + a], f[8 + b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[a], f[12 + b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(f[a], f[8 + b], acc[a][b], 0, 0, 0);
        if (ks + 1 < NKS) {
            if (ks + 3 < NKS) __builtin_amdgcn_sched_group_barrier(0x020, 6, 0);   // VMEM read (DMA)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // DS read
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);   // MFMA
            }
        }
        if (ks + 2 < NKS) {
            // slice ks + 2 has landed everywhere (only this slice's fetch of
            // ks + 3 may still be in flight) and slice ks + 1's fragments are
            // read everywhere before its buffer takes slice ks + 4
            __builtin_amdgcn_sched_barrier(0);
            if (ks + 3 < NKS) __builtin_amdgcn_s_waitcnt(0x76);   // vmcnt(6) lgkmcnt(0)
            else __builtin_amdgcn_s_waitcnt(0x70);                // vmcnt(0) lgkmcnt(0)
            asm volatile("s_barrier" ::: "memory");
        }
    }
    // epilogue: lane holds rows 4 (lane / 16) + r of column lane % 16 of a block
    if (wvalid) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int64_t gi = gi0 + 16 * a + 4 * fc, gj = gj0 + 16 * b + fr;
                const float sj = scale[r_off + gj];
                if (vec) {
                    const float4 si = si_pre[a];
                    float4* cp = reinterpret_cast<float4*>(C + gi + gj * ldc);
                    float4 c = c_pre[a][b];
                    c.x -= si.x * sj * acc[a][b][0];
                    c.y -= si.y * sj * acc[a][b][1];
                    c.z -= si.z * sj * acc[a][b][2];
                    c.w -= si.w * sj * acc[a][b][3];
                    *cp = c;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int64_t i = gi + r;
                        if (i < m && gj < m && i >= gj) C[i + gj * ldc] -= scale[r_off + i] * sj * acc[a][b][r];
                    }
                }
            }
    }
    if (clk.period_out && blockIdx.x == 0 && tid == 0) {
        const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t_start;
        clk.period_out[0] = (unsigned)max((uint64_t)1, min(dt / NKS, (uint64_t)1000000));
    }
}

// The same on the tile of map that this workgroup takes: the lower triangle,
// or (kind 6) the lower part of a strip of whole tile columns.
template <int KP, bool PP>
__device__ __forceinline__ void syrk_split_body(int64_t m, int64_t mp, int64_t r_off, const _Float16* __restrict__ Ah,
                                                const _Float16* __restrict__ Al, const _Float16* __restrict__ Bh,
                                                const _Float16* __restrict__ Bl, const float* __restrict__ scale,
                                                float* __restrict__ C, int64_t ldc, TileMap map,
                                                const TileRemap& rm, SliceClock clk = SliceClock{}) {
    int64_t t = blockIdx.x;
    if (map.kind >= 4) {
        t = xcd_tile(t, map.per);
        if (t >= map.ntiles) return;
    }
    int64_t tm, tn;
    map_tile(map, t, tm, tn);
    if (rm.n) {   // compressed order over the structurally non-zero tile rows
        tm = remap_tile(rm, tm);
        tn = remap_tile(rm, tn);
    }
    if (PP)
        syrk_split_tile_pp<KP>(m, mp, r_off, Ah, Al, Bh, Bl, scale, C, ldc, tm, tn, clk);
    else
        syrk_split_tile<KP>(m, mp, r_off, Ah, Al, Bh, Bl, scale, C, ldc, tm, tn, clk);
}

// The same update on 256 x 256 tiles: 512 threads, 8 waves of 64 x 128
// (4 x 8 MFMA blocks, 3 products each), so that every fragment read from LDS
// feeds more products (24 reads per 96 MFMAs against 16 per 48) while two
// waves per SIMD remain.  Operand slices go global -> LDS directly
// (global_load_lds, 16 bytes per lane, lane L to base + 16 L; each lane
// fetches the chunk whose swizzled LDS slot it lands in), two stages in
// separate LDS arrays; a slice's fragments are all read before the next
// slice's loads are issued, and the loads run under the slice's MFMAs.
// Operand rows up to r_off + 256 * ceil(m / 256) exist (zero padding).
constexpr int S2_TILE = 256;

template <int KP>
__global__ void __launch_bounds__(512) k_syrk_split256(int64_t m, int64_t mp, int64_t r_off,
                                                       const _Float16* __restrict__ Ah,
                                                       const _Float16* __restrict__ Al,
                                                       const _Float16* __restrict__ Bh,
                                                       const _Float16* __restrict__ Bl,
                                                       const float* __restrict__ scale, float* __restrict__ C,
                                                       int64_t ldc, TileMap map, TileRemap rm, SliceClock clk) {
    __shared__ half8_t S0[4][S2_TILE * 4];
    __shared__ half8_t S1[4][S2_TILE * 4];
    __shared__ int rot_sh;
    int64_t t = blockIdx.x;
    if (map.kind >= 4) {
        t = xcd_tile(t, map.per);
        if (t >= map.ntiles) return;
    }
    int64_t tm, tn;
    map_tile(map, t, tm, tn);
    if (rm.n) {
        tm = remap_tile(rm, tm);
        tn = remap_tile(rm, tn);
    }
    const int64_t row0 = r_off + tm * S2_TILE, col0 = r_off + tn * S2_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 128;
    // this lane's DMA sources: rows 32 wave + 16 p + lane / 4 of each array,
    // the chunk that belongs in LDS slot lane % 4 of that row (the swizzle
    // depends on the row's low 4 bits only: one 32-bit offset serves both p)
    const int rl = lane >> 2;
    const uint32_t goff = (uint32_t)(((32 * wave + rl) * SK + 8 * ((lane & 3) ^ (((rl >> 1) ^ (rl >> 2)) & 3))) * 2);
    constexpr int NKS = KP / SK;
    static_assert((NKS & (NKS - 1)) == 0, "the slice clock wraps k with a mask");
    uint64_t t_start = 0;
    int rot = 0;
    if (clk.period) {   // first slice of this tile (the slice clock)
        if (tid == 0) {
            t_start = __builtin_amdgcn_s_memrealtime();
            const unsigned per = clk.period[0];
            rot_sh = per ? (int)((t_start / per) & (uint64_t)(NKS - 1)) : 0;
        }
        __syncthreads();
        rot = rot_sh;
    }
    auto dma = [&](int ks, half8_t (*S)[S2_TILE * 4]) {
        const int64_t so = (int64_t)((ks + rot) & (NKS - 1)) * mp * SK;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const char* base = reinterpret_cast<const char*>((a == 0 ? Ah : a == 1 ? Al : a == 2 ? Bh : Bl) + so +
                                                             (a < 2 ? row0 : col0) * SK);
#pragma unroll
            for (int p = 0; p < 2; ++p)
                __builtin_amdgcn_global_load_lds((const void*)(base + goff + p * 16 * SK * 2),
                                                 (lds_void*)&S[a][(32 * wave + 16 * p) * 4], 16, 0, 0);
        }
    };
    float4_t acc[4][8];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int fr = lane & 15, fc = lane >> 4;
    // fragment (row r0 + fr, chunk fc) of a 16-row block at r0: slot
    // r0 * 4 + swz(fr, fc), r0 a multiple of 16
    const int lo = swz(fr, fc);
    auto step = [&](int ks, half8_t (*Sc)[S2_TILE * 4], half8_t (*Sn)[S2_TILE * 4]) {
        __builtin_amdgcn_sched_barrier(0);   // the previous slice's products stay above the wait
        __builtin_amdgcn_s_waitcnt(0);       // this wave's loads of slice ks have landed
        __syncthreads();                 // everyone's have; Sn is no longer read
        // product 1 (l x h) needs al, bh; products 2 and 3 ah, bl, bh: al is
        // dead before ah and bl are read (20 fragments live at most)
        const half8_t* A0 = &Sc[0][wm * 4 + lo];
        const half8_t* A1 = &Sc[1][wm * 4 + lo];
        const half8_t* B2 = &Sc[2][wn * 4 + lo];
        const half8_t* B3 = &Sc[3][wn * 4 + lo];
        half8_t ah[4], al[4], bh[8], bl[8];
#pragma unroll
        for (int a = 0; a < 4; ++a) al[a] = A1[64 * a];
#pragma unroll
        for (int b = 0; b < 8; ++b) bh[b] = B2[64 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < 4; ++a) ah[a] = A0[64 * a];
#pragma unroll
        for (int b = 0; b < 8; ++b) bl[b] = B3[64 * b];
        if (ks + 1 < NKS) dma(ks + 1, Sn);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
    };
    dma(0, S0);
    for (int ks = 0; ks < NKS; ks += 2) {
        step(ks, S0, S1);
        step(ks + 1, S1, S0);
    }
    // epilogue: lane holds rows 4 (lane / 16) + r of column lane % 16 of a block
    const int64_t i0 = tm * S2_TILE, j0 = tn * S2_TILE;
    const bool vec = tm != tn && i0 + S2_TILE <= m && (ldc % 4 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t gi = i0 + wm + 16 * a + 4 * fc;
        float4 si = make_float4(0.f, 0.f, 0.f, 0.f);
        if (vec) si = *reinterpret_cast<const float4*>(scale + r_off + gi);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int64_t gj = j0 + wn + 16 * b + fr;
            const float sj = scale[r_off + gj];
            if (vec) {
                float4* cp = reinterpret_cast<float4*>(C + gi + gj * ldc);
                float4 c = *cp;
                c.x -= si.x * sj * acc[a][b][0];
                c.y -= si.y * sj * acc[a][b][1];
                c.z -= si.z * sj * acc[a][b][2];
                c.w -= si.w * sj * acc[a][b][3];
                *cp = c;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = gi + r;
                    if (i < m && gj < m && i >= gj) C[i + gj * ldc] -= scale[r_off + i] * sj * acc[a][b][r];
                }
            }
        }
    }
    if (clk.period_out && blockIdx.x == 0 && tid == 0) {
        const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t_start;
        clk.period_out[0] = (unsigned)max((uint64_t)1, min(dt / NKS, (uint64_t)1000000));
    }
}

// The trailing matrix below the next block group (PP: the pipelined tile).
template <int KP, bool PP>
__global__ void __launch_bounds__(256) k_syrk_split(int64_t m, int64_t mp, int64_t r_off,
                                                    const _Float16* __restrict__ Ah, const _Float16* __restrict__ Al,
                                                    const _Float16* __restrict__ Bh, const _Float16* __restrict__ Bl,
                                                    const float* __restrict__ scale, float* __restrict__ C,
                                                    int64_t ldc, TileMap map, TileRemap rm, SliceClock clk) {
    syrk_split_body<KP, PP>(m, mp, r_off, Ah, Al, Bh, Bl, scale, C, ldc, map, rm, clk);
}

// The next block group's columns (a strip of whole tile columns, map kind 6).
template <int KP, bool PP>
__global__ void __launch_bounds__(256) k_syrk_strip(int64_t m, int64_t mp, int64_t r_off,
                                                    const _Float16* __restrict__ Ah, const _Float16* __restrict__ Al,
                                                    const _Float16* __restrict__ Bh, const _Float16* __restrict__ Bl,
                                                    const float* __restrict__ scale, float* __restrict__ C,
                                                    int64_t ldc, TileMap map) {
    syrk_split_body<KP, PP>(m, mp, r_off, Ah, Al, Bh, Bl, scale, C, ldc, map, TileRemap{});
}

// Grouped rank-nb splits: workgroup b takes 64-row block b - blk0 of the
// matrix whose blocks it falls in (blk0 ascending); thread (row t % 64,
// 64-wide k quarter t / 64).  Columns k >= nb and rows >= m are zero.
__device__ __forceinline__ int desc_of(const SplitDesc* __restrict__ d, int ng, int64_t b, bool tiles) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if ((tiles ? d[mid].tile0 : d[mid].blk0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_split_grouped(const SplitDesc* __restrict__ ds, int ng) {
    __shared__ float red[4][64];
    const SplitDesc& d = ds[desc_of(ds, ng, blockIdx.x, false)];
    const int ri = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int64_t row = ((int64_t)blockIdx.x - d.blk0) * 64 + ri;
    const int kb = kq * 64;
    float v[64];
    float mx = 0.f;
#pragma unroll
    for (int kk = 0; kk < 64; ++kk) {
        v[kk] = (row < d.m && kb + kk < d.nb) ? d.X[row + (int64_t)(kb + kk) * d.ldx] : 0.f;
        mx = fmaxf(mx, fabsf(v[kk]));
    }
    red[kq][ri] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0][ri], red[1][ri]), fmaxf(red[2][ri], red[3][ri]));
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);   // mx = f 2^e, f in [0.5, 1)
    const float inv = ldexpf(1.f, -e);
    const size_t arr = (size_t)d.mp * FACT_NB;
    _Float16* Ah = (_Float16*)d.Ah;
    _Float16* Al = Ah + arr;
    _Float16* Bh = Al + arr;
    _Float16* Bl = Bh + arr;
    if (kq == 0) reinterpret_cast<float*>(Bl + arr)[row] = ldexpf(1.f, e);
#pragma unroll
    for (int kk = 0; kk < 64; kk += 8) {
        half8_t h, l, bh, bl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float x = v[kk + q] * inv;
            const _Float16 hx = (_Float16)x;
            const _Float16 lx = (_Float16)(x - (float)hx);
            const bool neg = kb + kk + q < d.nb && d.sgn[kb + kk + q] < 0.f;
            h[q] = hx;
            l[q] = lx;
            bh[q] = neg ? -hx : hx;
            bl[q] = neg ? -lx : lx;
        }
        const int k0 = kb + kk;
        const int64_t off = ((int64_t)(k0 / SK) * d.mp + row) * SK + (k0 % SK);
        *reinterpret_cast<half8_t*>(Ah + off) = h;
        *reinterpret_cast<half8_t*>(Al + off) = l;
        *reinterpret_cast<half8_t*>(Bh + off) = bh;
        *reinterpret_cast<half8_t*>(Bl + off) = bl;
    }
}

// Grouped split updates: workgroup b takes tile b - tile0 (lower triangle,
// row by row) of the matrix it falls in.
__global__ void __launch_bounds__(256) k_syrk_split_grouped(const SplitDesc* __restrict__ ds, int ng) {
    const SplitDesc& d = ds[desc_of(ds, ng, blockIdx.x, true)];
    int64_t tm, tn;
    upper_tile((int64_t)blockIdx.x - d.tile0, tn, tm);   // tn <= tm
    const size_t arr = (size_t)d.mp * FACT_NB;
    const _Float16* Ah = (const _Float16*)d.Ah;
    syrk_split_tile_pp<FACT_NB>(d.m, d.mp, 0, Ah, Ah + arr, Ah + 2 * arr, Ah + 3 * arr,
                                reinterpret_cast<const float*>(Ah + 4 * arr), d.C, d.ldc, tm, tn);
}

// Rank-1024 grouped splits (the batch's super-steps): workgroup b takes
// 64-row block b - blk0 of the matrix it falls in; thread (row t % 64,
// 64-wide k group t / 64), k in [256 p, 256 p + 256) from panel p, one
// power-of-two scale per row over all 1024 columns.  Rows >= m are zero.
__device__ __forceinline__ int desc4_of(const Split4Desc* __restrict__ d, int ng, int64_t b, bool tiles) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if ((tiles ? d[mid].tile0 : d[mid].blk0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(64 * 4 * S_PANELS) k_split4_grouped(const Split4Desc* __restrict__ ds, int ng) {
    static_assert(S_PANELS == S4_PANELS, "super-steps of the factorisation's block groups");
    __shared__ float red[4 * S_PANELS][64];
    const Split4Desc& d = ds[desc4_of(ds, ng, blockIdx.x, false)];
    const int ri = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int64_t row = ((int64_t)blockIdx.x - d.blk0) * 64 + ri;
    const int pn = kq >> 2;
    const float* X = d.X[pn] + d.off[pn];
    const int64_t ldx = d.ldx[pn];
    const int kb = (kq & 3) * 64;   // first column inside its panel
    float v[64];
    float mx = 0.f;
#pragma unroll
    for (int kk = 0; kk < 64; ++kk) {
        v[kk] = row < d.m ? X[row + (int64_t)(kb + kk) * ldx] : 0.f;
        mx = fmaxf(mx, fabsf(v[kk]));
    }
    red[kq][ri] = mx;
    __syncthreads();
    mx = 0.f;
#pragma unroll
    for (int g = 0; g < 4 * S_PANELS; ++g) mx = fmaxf(mx, red[g][ri]);
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);   // mx = f 2^e, f in [0.5, 1)
    const float inv = ldexpf(1.f, -e);
    const size_t arr = (size_t)d.mp * S_KP;
    _Float16* Ah = (_Float16*)d.Ah;
    _Float16* Al = Ah + arr;
    _Float16* Bh = Al + arr;
    _Float16* Bl = Bh + arr;
    if (kq == 0) reinterpret_cast<float*>(Bl + arr)[row] = ldexpf(1.f, e);
    const float* sg = d.sgn + pn * FACT_NB;
#pragma unroll
    for (int kk = 0; kk < 64; kk += 8) {
        half8_t h, l, bh, bl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float x = v[kk + q] * inv;
            const _Float16 hx = (_Float16)x;
            const _Float16 lx = (_Float16)(x - (float)hx);
            const bool neg = sg[kb + kk + q] < 0.f;
            h[q] = hx;
            l[q] = lx;
            bh[q] = neg ? -hx : hx;
            bl[q] = neg ? -lx : lx;
        }
        const int k0 = kq * 64 + kk;
        const int64_t off = ((int64_t)(k0 / SK) * d.mp + row) * SK + (k0 % SK);
        *reinterpret_cast<half8_t*>(Ah + off) = h;
        *reinterpret_cast<half8_t*>(Al + off) = l;
        *reinterpret_cast<half8_t*>(Bh + off) = bh;
        *reinterpret_cast<half8_t*>(Bl + off) = bl;
    }
}

// Rank-1024 grouped split updates: workgroup b takes tile b - tile0 (lower
// triangle, row by row) of the matrix it falls in.
__global__ void __launch_bounds__(256) k_syrk_split4_grouped(const Split4Desc* __restrict__ ds, int ng) {
    const Split4Desc& d = ds[desc4_of(ds, ng, blockIdx.x, true)];
    int64_t tm, tn;
    upper_tile((int64_t)blockIdx.x - d.tile0, tn, tm);   // tn <= tm
    const size_t arr = (size_t)d.mp * S_KP;
    const _Float16* Ah = (const _Float16*)d.Ah;
    syrk_split_tile_pp<S_KP>(d.m, d.mp, 0, Ah, Ah + arr, Ah + 2 * arr, Ah + 3 * arr,
                             reinterpret_cast<const float*>(Ah + 4 * arr), d.C, d.ldc, tm, tn);
}

// flag = 1 if any entry of the m x n block is non-zero.
template <typename T>
__global__ void __launch_bounds__(256) k_any_nonzero(const T* __restrict__ A, int64_t lda, int64_t m, int64_t n,
                                                     int* __restrict__ flag) {
    const int64_t total = m * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        if (A[i + j * lda] != T(0)) {
            *flag = 1;
            return;
        }
    }
}

__global__ void k_set_int(int* p, int v) { p[0] = v; }

inline unsigned grid_for(int64_t total) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(total, 256), 16384)); }

}  // namespace

// ---------------------------------------------------------------------------
// host API
// ---------------------------------------------------------------------------
template <typename T, int WB, bool EDGE>
static void launch_tiles(bool ta, bool tb, bool lower, int64_t nblk, TileMap map, int64_t m, int64_t n, int64_t k,
                         T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
                         hipStream_t st) {
    if (nblk == 0) nblk = map.kind >= 4 ? 8 * map.per : map.ntiles;   // grid from the map
    if (nblk <= 0) return;
    if (nblk > 0xffffffffLL) throw std::invalid_argument("gemm grid too large");
#define OMH_GEMM_CASE(TA_, TB_, LO_)                                                                          \
    if (ta == TA_ && tb == TB_ && lower == LO_) {                                                             \
        k_gemm<T, TA_, TB_, LO_, WB, EDGE><<<(unsigned)nblk, 256, 0, st>>>(m, n, k, alpha, A, lda, B, ldb, beta, \
                                                                           C, ldc, map);                      \
        OMH_LAUNCH_CHECK();                                                                                   \
        return;                                                                                               \
    }
    OMH_GEMM_CASE(false, false, false)
    OMH_GEMM_CASE(true, false, false)
    OMH_GEMM_CASE(false, true, false)
    OMH_GEMM_CASE(true, true, false)
    OMH_GEMM_CASE(false, true, true)
#undef OMH_GEMM_CASE
    throw std::invalid_argument("unsupported gemm variant");
}

// Tile map of an interior launch: the L2-aware order (kinds 4/5) when the
// launch has enough tiles to spread over the XCDs, else the plain order.
static TileMap l2_map(bool lower, int64_t fm, int64_t fn) {
    const int64_t tiles = lower ? fm * (fm + 1) / 2 : fm * fn;
    if (tiles < 1024) return lower ? TileMap{1, fm, 0, 0, tiles, 0} : TileMap{0, fm, 0, 0, tiles, 0};
    const int64_t per = cdiv(tiles, 8);
    return lower ? TileMap{4, fm, 0, 0, tiles, per} : TileMap{5, fm, 0, fn, tiles, per};
}

// Interior tiles run the unguarded kernel; the partial last tile row and
// column (if any) run as edge strips.
template <typename T, int WB>
static void gemm_launch(bool ta, bool tb, bool lower, bool vec, int64_t m, int64_t n, int64_t k, T alpha, const T* A,
                        int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st,
                        const char* prof) {
    constexpr int TILE = 32 * WB;
    const int64_t ntm = cdiv(m, TILE), ntn = cdiv(n, TILE);
    if (!vec) {
        const int64_t nblk = lower ? ntm * (ntm + 1) / 2 : ntm * ntn;
        launch_tiles<T, WB, true>(ta, tb, lower, nblk, TileMap{lower ? 1 : 0, ntm, 0, 0}, m, n, k, alpha, A, lda, B,
                                  ldb, beta, C, ldc, st);
        return;
    }
    const int64_t fm = m / TILE, fn = n / TILE;   // full tile rows / columns
    if (lower) {
        {
            // optional timing of the interior launch; algorithmic work: the
            // lower triangle of the interior square, (fm T)(fm T + 1) k flop
            std::string name;
            if (prof) name = std::string(prof) + "_t" + std::to_string(TILE);
            const double side = (double)(fm * TILE);
            ProfScope ps(prof ? name.c_str() : "", prof ? side * (side + 1) * (double)k : 0.0,
                         prof ? st : nullptr);
            if (!prof) ps.cancel();
            launch_tiles<T, WB, false>(ta, tb, true, 0, l2_map(true, fm, fm), m, n, k, alpha, A, lda, B, ldb, beta,
                                       C, ldc, st);
        }
        if (fm < ntm)   // last (partial) tile row, all its lower tiles
            launch_tiles<T, WB, true>(ta, tb, true, ntm, TileMap{2, ntm, fm, 0}, m, n, k, alpha, A, lda, B, ldb,
                                      beta, C, ldc, st);
        return;
    }
    launch_tiles<T, WB, false>(ta, tb, false, 0, l2_map(false, fm, fn), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                               st);
    if (fm < ntm)   // partial last tile row (all tile columns)
        launch_tiles<T, WB, true>(ta, tb, false, ntn, TileMap{2, ntm, fm, 0}, m, n, k, alpha, A, lda, B, ldb, beta,
                                  C, ldc, st);
    if (fn < ntn)   // partial last tile column (full tile rows only)
        launch_tiles<T, WB, true>(ta, tb, false, fm, TileMap{3, ntm, 0, fn}, m, n, k, alpha, A, lda, B, ldb, beta,
                                  C, ldc, st);
}

template <typename T>
static bool aligned_for_vec(const T* p, int64_t ld) {
    constexpr int V = 16 / sizeof(T);
    return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && (ld % V == 0);
}

// ---------------------------------------------------------------------------
// grouped launches
// ---------------------------------------------------------------------------
static bool aligned_bytes(const void* p, int64_t ld, size_t elem) {
    return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ((ld * (int64_t)elem) % 16 == 0);
}

void build_grouped_gemm(const std::vector<GemmJob>& jobs, bool f32, bool ta, bool tb, bool lower, int wb,
                        std::vector<GemmDesc>& descs, std::vector<GemmLaunch>& out) {
    if (wb != 2 && wb != 4) throw std::invalid_argument("tile factor must be 2 or 4");
    if (!f32 && wb != 2) throw std::invalid_argument("FP64 grouped GEMM uses 64 x 64 tiles");
    if (lower && (ta || !tb)) throw std::invalid_argument("lower products are X * Y^T only");
    const int64_t TILE = 32 * wb;
    const size_t elem = f32 ? 4 : 8;
    std::vector<GemmDesc> in, ed;
    int64_t ti = 0, te = 0;
    auto add = [](std::vector<GemmDesc>& v, int64_t& tot, const GemmJob& j, TileMap map, int64_t ntiles) {
        if (ntiles <= 0) return;
        v.push_back(GemmDesc{j.A, j.B, j.C, j.lda, j.ldb, j.ldc, j.m, j.n, j.k, j.alpha, j.beta, map, tot});
        tot += ntiles;
    };
    for (const GemmJob& j : jobs) {
        if (j.m <= 0 || j.n <= 0) continue;
        if (lower && j.m != j.n) throw std::invalid_argument("lower gemm needs a square output");
        const int64_t ntm = cdiv(j.m, TILE), ntn = cdiv(j.n, TILE);
        const bool vec = aligned_bytes(j.A, j.lda, elem) && aligned_bytes(j.B, j.ldb, elem);
        if (!vec) {
            add(ed, te, j, TileMap{lower ? 1 : 0, ntm, 0, 0}, lower ? ntm * (ntm + 1) / 2 : ntm * ntn);
            continue;
        }
        const int64_t fm = j.m / TILE, fn = j.n / TILE;
        if (lower) {
            add(in, ti, j, TileMap{1, fm, 0, 0}, fm * (fm + 1) / 2);
            if (fm < ntm) add(ed, te, j, TileMap{2, ntm, fm, 0}, ntm);
        } else {
            add(in, ti, j, TileMap{0, fm, 0, 0}, fm * fn);
            if (fm < ntm) add(ed, te, j, TileMap{2, ntm, fm, 0}, ntn);
            if (fn < ntn) add(ed, te, j, TileMap{3, ntm, 0, fn}, fm);
        }
    }
    const int base = (f32 ? 1 : 0) | (ta ? 2 : 0) | (tb ? 4 : 0) | (lower ? 8 : 0) | (wb == 4 ? 16 : 0);
    if (!in.empty()) {
        out.push_back(GemmLaunch{base, descs.size(), (int)in.size(), ti});
        descs.insert(descs.end(), in.begin(), in.end());
    }
    if (!ed.empty()) {
        out.push_back(GemmLaunch{base | 32, descs.size(), (int)ed.size(), te});
        descs.insert(descs.end(), ed.begin(), ed.end());
    }
}

size_t split_operand_bytes(int64_t m) {
    const size_t mp = (size_t)cdiv(m, S_TILE) * S_TILE;
    return 4 * mp * FACT_NB * sizeof(_Float16) + mp * sizeof(float);
}

void launch_grouped_split(const SplitDesc* d_descs, int ng, int64_t blocks, hipStream_t st) {
    if (ng <= 0 || blocks <= 0) return;
    k_split_grouped<<<(unsigned)blocks, 256, 0, st>>>(d_descs, ng);
    OMH_LAUNCH_CHECK();
}

size_t split4_operand_bytes(int64_t m) {
    const size_t mp = (size_t)cdiv(m, S_TILE) * S_TILE;
    return 4 * mp * S_KP * sizeof(_Float16) + mp * sizeof(float);
}

void launch_grouped_split4(const Split4Desc* d_descs, int ng, int64_t blocks, hipStream_t st) {
    if (ng <= 0 || blocks <= 0) return;
    k_split4_grouped<<<(unsigned)blocks, 64 * 4 * S_PANELS, 0, st>>>(d_descs, ng);
    OMH_LAUNCH_CHECK();
}

void launch_grouped_split4_update(const Split4Desc* d_descs, int ng, int64_t tiles, hipStream_t st) {
    if (ng <= 0 || tiles <= 0) return;
    k_syrk_split4_grouped<<<(unsigned)tiles, 256, 0, st>>>(d_descs, ng);
    OMH_LAUNCH_CHECK();
}

void launch_grouped_split_update(const SplitDesc* d_descs, int ng, int64_t tiles, hipStream_t st) {
    if (ng <= 0 || tiles <= 0) return;
    k_syrk_split_grouped<<<(unsigned)tiles, 256, 0, st>>>(d_descs, ng);
    OMH_LAUNCH_CHECK();
}

void launch_grouped_gemm(const GemmLaunch& l, const GemmDesc* d_descs, int64_t grid_cap, hipStream_t st) {
    if (l.total <= 0) return;
    const int64_t g = grid_cap > 0 ? std::min(l.total, grid_cap) : l.total;
    if (g > 0x7fffffffLL) throw std::invalid_argument("grouped gemm grid too large");
    const unsigned grid = (unsigned)g;
    const GemmDesc* d = d_descs + l.off;
#define OMH_GG(F32, TA, TB, LO, W4, ED)                                                                    \
    case (F32) | ((TA) << 1) | ((TB) << 2) | ((LO) << 3) | ((W4) << 4) | ((ED) << 5):                    \
        k_gemm_grouped<typename std::conditional<(F32) != 0, float, double>::type, (TA) != 0, (TB) != 0,   \
                       (LO) != 0, (W4) ? 4 : 2, (ED) != 0><<<grid, 256, 0, st>>>(d, l.ng, l.total);        \
        break;
#define OMH_GG_OPS(F32, W4, ED) \
    OMH_GG(F32, 0, 0, 0, W4, ED) OMH_GG(F32, 1, 0, 0, W4, ED) OMH_GG(F32, 0, 1, 0, W4, ED) OMH_GG(F32, 0, 1, 1, W4, ED)
    switch (l.variant) {
        OMH_GG_OPS(1, 0, 0)
        OMH_GG_OPS(1, 0, 1)
        OMH_GG_OPS(1, 1, 0)
        OMH_GG_OPS(1, 1, 1)
        OMH_GG_OPS(0, 0, 0)
        OMH_GG_OPS(0, 0, 1)
        default: throw std::invalid_argument("unsupported grouped gemm variant");
    }
#undef OMH_GG_OPS
#undef OMH_GG
    OMH_LAUNCH_CHECK();
}

void launch_grouped_diag(bool f32, const DiagDesc* d_descs, int ng, hipStream_t st) {
    if (ng <= 0) return;
    if (f32) {
        k_diag_grouped<float><<<ng, DF_THREADS, 0, st>>>(d_descs, 1e-30f);
        OMH_LAUNCH_CHECK();
        k_diag_inverse_grouped<float><<<dim3(inv_blocks<float>(), (unsigned)ng), DIAG_THREADS, 0, st>>>(d_descs);
    } else {
        k_diag_grouped<double><<<ng, DF_THREADS, 0, st>>>(d_descs, 1e-290);
        OMH_LAUNCH_CHECK();
        k_diag_inverse_grouped<double><<<dim3(inv_blocks<double>(), (unsigned)ng), DIAG_THREADS, 0, st>>>(d_descs);
    }
    OMH_LAUNCH_CHECK();
}

void launch_grouped_elem(const ElemDesc* d_descs, int ng, int64_t max_elems, hipStream_t st) {
    if (ng <= 0 || max_elems <= 0) return;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(max_elems, 256), 2048));
    k_elem_grouped<<<dim3(gx, (unsigned)ng), 256, 0, st>>>(d_descs);
    OMH_LAUNCH_CHECK();
}

// FP64 A B^T on the 128 x 128 DMA tiles over the interior [0, 128 fm) x
// [0, 128 fn) (lower: its lower triangle of tiles), the rest by gemm_launch
// on the strips around it.  Returns false (nothing launched) when the
// product does not fit the kernel: operands 16-byte aligned with even
// leading dimensions, enough whole tiles.
// min_tiles: fewer interior tiles than this leave the product to gemm's
// smaller tiles (a launch of its own should fill the CUs; products that run
// concurrently on several streams can take fewer each).
// partial_rows (rectangular products with an even m): the rows past the last
// whole tile row run as a partial tile row of the same kernel instead of a
// 64 x 64 edge strip (whose few rows would walk all of k on a handful of
// workgroups).
static bool gemm_f64_dma(bool lower, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                         const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t st,
                         const char* prof, int64_t min_tiles = 256, bool partial_rows = false) {
    const char* e = std::getenv("OMH_GEMM_F64_DMA");
    if (e && e[0] == '0') return false;
    if (k <= 0 || lda % 2 || ldb % 2 || reinterpret_cast<uintptr_t>(A) % 16 ||
        reinterpret_cast<uintptr_t>(B) % 16)
        return false;
    const bool part = partial_rows && !lower && m % 2 == 0 && m % DT != 0;
    const int64_t fm = part ? cdiv(m, DT) : m / DT, fn = lower ? fm : n / DT;
    const int64_t tiles = lower ? fm * (fm + 1) / 2 : fm * fn;
    if (tiles < std::max<int64_t>(min_tiles, 1)) return false;
    const int64_t mrows = part ? m : fm * DT, arow_max = part ? m - 2 : fm * DT;
    const TileMap map = l2_map(lower, fm, fn);
    const unsigned grid = (unsigned)(map.kind >= 4 ? 8 * map.per : map.ntiles);
    {
        // optional timing; algorithmic work: the interior's lower triangle
        // (lower) or rectangle, 2 k flop per entry
        std::string name;
        if (prof) name = std::string(prof) + "_t" + std::to_string(DT);
        const double side = (double)std::min(fm * DT, m);
        const double work = lower ? side * (side + 1) * (double)k : 2.0 * side * (double)(fn * DT) * (double)k;
        ProfScope ps(prof ? name.c_str() : "", prof ? work : 0.0, prof ? st : nullptr);
        if (!prof) ps.cancel();
        if (lower)
            k_gemm_f64_dma<true, false><<<grid, 512, 0, st>>>(k, alpha, A, lda, B, ldb, beta, C, ldc, map, mrows,
                                                              arow_max);
        else if (part)
            k_gemm_f64_dma<false, true><<<grid, 512, 0, st>>>(k, alpha, A, lda, B, ldb, beta, C, ldc, map, mrows,
                                                              arow_max);
        else
            k_gemm_f64_dma<false, false><<<grid, 512, 0, st>>>(k, alpha, A, lda, B, ldb, beta, C, ldc, map, mrows,
                                                               arow_max);
        OMH_LAUNCH_CHECK();
    }
    const int64_t r1 = std::min(fm * DT, m), c1 = fn * DT;
    if (lower) {
        if (r1 < m) {   // the rows below the interior square: a rectangle and the last lower square
            gemm<double>(false, true, false, m - r1, r1, k, alpha, A + r1, lda, B, ldb, beta, C + r1, ldc, st);
            gemm<double>(false, true, true, m - r1, m - r1, k, alpha, A + r1, lda, B + r1, ldb, beta,
                         C + r1 + r1 * ldc, ldc, st);
        }
        return true;
    }
    if (r1 < m)   // rows below the interior, all columns
        gemm<double>(false, true, false, m - r1, n, k, alpha, A + r1, lda, B, ldb, beta, C + r1, ldc, st);
    if (c1 < n)   // columns right of the interior, its rows
        gemm<double>(false, true, false, r1, n - c1, k, alpha, A, lda, B + c1, ldb, beta, C + c1 * ldc, ldc, st);
    return true;
}

template <typename T>
void gemm(bool ta, bool tb, bool lower, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
          const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st, const char* prof) {
    if (m <= 0 || n <= 0) return;
    if (lower && m != n) throw std::invalid_argument("lower gemm needs a square output");
    if (sizeof(T) == 8 && !ta && tb &&
        gemm_f64_dma(lower, m, n, k, (double)alpha, (const double*)A, lda, (const double*)B, ldb, (double)beta,
                     (double*)C, ldc, st, prof, 256, true))
        return;
    const bool vec = aligned_for_vec(A, lda) && aligned_for_vec(B, ldb);
    // FP64: 64 x 64 tiles (128 x 128 accumulators would limit FP64 to one wave
    // per SIMD).  FP32: 128 x 128 tiles when they give at least ~2 waves of
    // workgroups over the 256 CUs.
    const int64_t big_tiles = lower ? cdiv(m, 128) * (cdiv(m, 128) + 1) / 2 : cdiv(m, 128) * cdiv(n, 128);
    if (sizeof(T) == 4 && big_tiles >= 512)
        gemm_launch<T, 4>(ta, tb, lower, vec, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st, prof);
    else
        gemm_launch<T, 2>(ta, tb, lower, vec, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st, prof);
}

// Split-K launch: interior tiles unguarded (16-byte aligned operands and
// chunk starts), the partial last tile row and column guarded.
template <typename T, bool TA, bool TB>
static void splitk_tiles(bool vec, int64_t m, int64_t n, int64_t k, int64_t kc, int nsplit, const T* A, int64_t lda,
                         const T* B, int64_t ldb, T* W, hipStream_t st) {
    constexpr int WB = 2, TILE = 64;
    const int64_t ntm = cdiv(m, TILE), ntn = cdiv(n, TILE);
    if (!vec) {
        k_gemm_splitk<T, TA, TB, WB, true><<<dim3((unsigned)(ntm * ntn), nsplit), 256, 0, st>>>(
            m, n, k, kc, A, lda, B, ldb, W, TileMap{0, ntm, 0, 0});
        OMH_LAUNCH_CHECK();
        return;
    }
    const int64_t fm = m / TILE, fn = n / TILE;
    if (fm * fn > 0)
        k_gemm_splitk<T, TA, TB, WB, false><<<dim3((unsigned)(fm * fn), nsplit), 256, 0, st>>>(
            m, n, k, kc, A, lda, B, ldb, W, TileMap{0, fm, 0, 0});
    if (fm < ntm)   // partial last tile row (all tile columns)
        k_gemm_splitk<T, TA, TB, WB, true><<<dim3((unsigned)ntn, nsplit), 256, 0, st>>>(
            m, n, k, kc, A, lda, B, ldb, W, TileMap{2, ntm, fm, 0});
    if (fn < ntn && fm > 0)   // partial last tile column (full tile rows only)
        k_gemm_splitk<T, TA, TB, WB, true><<<dim3((unsigned)fm, nsplit), 256, 0, st>>>(
            m, n, k, kc, A, lda, B, ldb, W, TileMap{3, ntm, 0, fn});
    OMH_LAUNCH_CHECK();
}

template <typename T>
void gemm_entry(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const int64_t tiles = cdiv(m, 64) * cdiv(n, 64);
    if (tiles >= 512 || k < 8 * std::max(m, n) || k < 4096) {
        gemm<T>(ta, tb, false, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st);
        return;
    }
    // about 2048 workgroups, chunks of whole 64-index slices
    const int nsplit = (int)std::min<int64_t>(16, std::max<int64_t>(2, 2048 / tiles));
    const int64_t kc = cdiv(cdiv(k, nsplit), 64) * 64;
    const int ns = (int)cdiv(k, kc);
    T* W = (T*)scratch(SCR_GEMM, sizeof(T) * (size_t)ns * (size_t)m * (size_t)n, st);
    const bool vec = aligned_for_vec(A, lda) && aligned_for_vec(B, ldb);
    if (ta && tb) splitk_tiles<T, true, true>(vec, m, n, k, kc, ns, A, lda, B, ldb, W, st);
    else if (ta) splitk_tiles<T, true, false>(vec, m, n, k, kc, ns, A, lda, B, ldb, W, st);
    else if (tb) splitk_tiles<T, false, true>(vec, m, n, k, kc, ns, A, lda, B, ldb, W, st);
    else splitk_tiles<T, false, false>(vec, m, n, k, kc, ns, A, lda, B, ldb, W, st);
    k_splitk_sum<T><<<(unsigned)std::min<int64_t>(cdiv(m * n, 256), 4096), 256, 0, st>>>(W, ns, m, n, alpha, beta, C,
                                                                                         ldc);
    OMH_LAUNCH_CHECK();
}

template void gemm_entry<double>(bool, bool, int64_t, int64_t, int64_t, double, const double*, int64_t,
                                 const double*, int64_t, double, double*, int64_t, hipStream_t);
template void gemm_entry<float>(bool, bool, int64_t, int64_t, int64_t, float, const float*, int64_t, const float*,
                                int64_t, float, float*, int64_t, hipStream_t);

template void gemm<double>(bool, bool, bool, int64_t, int64_t, int64_t, double, const double*, int64_t, const double*,
                           int64_t, double, double*, int64_t, hipStream_t, const char*);
template void gemm<float>(bool, bool, bool, int64_t, int64_t, int64_t, float, const float*, int64_t, const float*,
                          int64_t, float, float*, int64_t, hipStream_t, const char*);

// Workspace: panel X | diagonal inverses | signs (padded) | second panel
// (the look-ahead computes the next panel while the current one is in use).
static size_t sign_span(int64_t n) { return (size_t)((n + 63) / 64 * 64); }

// Split-FP16 operands of one block group (FP32 factorisation): four arrays
// of mp x S_KP halves and mp row scales, mp = n rounded up to S_TILE; then
// the panels beyond the two of the dense workspace.
static size_t split_set_bytes(int64_t n) {
    const size_t mp = (size_t)cdiv(n, S_TILE) * S_TILE;
    return 4 * mp * S_KP * sizeof(_Float16) + mp * sizeof(float) +
           (size_t)(S_PANELS > 2 ? S_PANELS - 2 : 0) * ((size_t)n * FACT_NB * sizeof(float) + 256) +
           // the second set of operands (the look-ahead splits the next
           // group's panels while the main update reads the current set)
           256 + 4 * mp * S_KP * sizeof(_Float16) + mp * sizeof(float);
}

static size_t dense_workspace_bytes(int64_t n, size_t elem) {
    const int64_t nblk = cdiv(n, FACT_NB);
    return elem * ((size_t)n * FACT_NB            // panel X
                   + (size_t)nblk * FACT_NB * FACT_NB   // diagonal inverses
                   + sign_span(n)                 // signs
                   + (size_t)n * FACT_NB);        // second panel
}

size_t factor_workspace_bytes(int64_t n, size_t elem) {
    size_t b = dense_workspace_bytes(n, elem) + 64;
    if (elem == 4) b = (b + 255) / 256 * 256 + split_set_bytes(n) + 256;   // split operands
    return b;
}

// Side stream and events of the look-ahead (one set per device, default
// priority).
constexpr int SCALE_EVENTS = 32;   // chunks of groups whose solve scales are formed ahead (half per direction)
struct LookAhead {
    hipStream_t side = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    hipEvent_t scale_ev[SCALE_EVENTS] = {};
};

static LookAhead& lookahead() {
    static std::mutex mu;
    static std::map<int, LookAhead> per_dev;
    int dev = 0;
    OMH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    LookAhead& la = per_dev[dev];
    if (!la.side) {
        OMH_HIP(hipStreamCreateWithFlags(&la.side, hipStreamNonBlocking));
        OMH_HIP(hipEventCreateWithFlags(&la.ready, hipEventDisableTiming));
        OMH_HIP(hipEventCreateWithFlags(&la.done, hipEventDisableTiming));
        for (int i = 0; i < SCALE_EVENTS; ++i) OMH_HIP(hipEventCreateWithFlags(&la.scale_ev[i], hipEventDisableTiming));
    }
    return la;
}

template <typename T>
static void diag_factor(T* A, int64_t lda, int64_t k, int nb, T* Dinv, T* sgn, int* info, T tiny, hipStream_t st) {
    T* A11 = A + k + k * lda;
    T* Li = Dinv + (size_t)(k / FACT_NB) * FACT_NB * FACT_NB;
    ProfScope ps(sizeof(T) == 8 ? "diag_factor_f64" : "diag_factor_f32", (double)nb * nb * nb / 2.0, st);
    k_diag_factor<T><<<1, DF_THREADS, 0, st>>>(A11, lda, nb, sgn, info, k, tiny, Li, FACT_NB);
    OMH_LAUNCH_CHECK();
    const char* wide64 = sizeof(T) == 8 ? std::getenv("OMH_DIAG_INV64_WIDE") : nullptr;
    if (wide64 && wide64[0] == '1')
        k_diag_inverse_wide64<<<cdiv(nb, IB), DIAG_THREADS, 0, st>>>((const double*)A11, lda, nb, (double*)Li,
                                                                     FACT_NB);
    else
        k_diag_inverse<T><<<cdiv(nb, sizeof(T) == 8 ? INV64_CW : IB), DIAG_THREADS, 0, st>>>(A11, lda, nb, Li,
                                                                                          FACT_NB);
    OMH_LAUNCH_CHECK();
}

// Right-looking blocked factorisation with a one-block look-ahead: after
// the panel of block k, the columns of block k+1 are updated first, then
// block k+1 is factored on the side stream while the rest of the trailing
// matrix is updated on the caller's stream.
// FP32 factorisation with split-FP16 trailing updates, in super-steps of
// S_PANELS blocks (a group) so that the large update has rank S_KP and reads
// and writes the trailing matrix once per S_PANELS blocks:
//   prologue of a group: each block in turn is factored, forms its panel and
//     updates the group's later columns (FP32 GEMMs);
//   the next group's columns -= all panels of the group (FP32 GEMMs);
//   the next group's prologue on the side stream, while the rest of the
//   trailing matrix -= [X_1 .. X_P] S [X_1 .. X_P]^T as one split-FP16 SYRK.
// Row intervals [lo, hi) at or below row r that the update with the panel
// of columns [c0, c1) can reach (segments whose rows may be non-zero there
// in L), merged.  Returns false when that is every row from r on (iv is then
// empty); true with the intervals (possibly none) otherwise.
static bool active_rows(const BlockFill& bf, int64_t n, int64_t c0, int64_t c1, int64_t r,
                        std::vector<std::pair<int64_t, int64_t>>& iv) {
    iv.clear();
    const int ns = bf.nseg();
    for (int s = 0; s < ns; ++s) {
        bool act = false;
        for (int t = 0; t < ns && !act; ++t)
            if (bf.start[t] < c1 && bf.start[t + 1] > c0 && bf.nz(s, t)) act = true;
        const int64_t lo = std::max(bf.start[s], r), hi = std::min(bf.start[s + 1], n);
        if (!act || hi <= lo) continue;
        if (!iv.empty() && iv.back().second == lo) iv.back().second = hi;
        else iv.emplace_back(lo, hi);
    }
    if (iv.size() == 1 && iv[0].first == r && iv[0].second == n) {
        iv.clear();
        return false;
    }
    return true;
}

// Column intervals [lo, hi) below column c that rows [r0, r1) of L can
// reach (the column segments of their possibly non-zero blocks), merged;
// the same convention as active_rows.
static bool active_cols(const BlockFill& bf, int64_t r0, int64_t r1, int64_t c,
                        std::vector<std::pair<int64_t, int64_t>>& iv) {
    iv.clear();
    const int ns = bf.nseg();
    for (int t = 0; t < ns; ++t) {
        bool act = false;
        for (int s = 0; s < ns && !act; ++s)
            if (bf.start[s] < r1 && bf.start[s + 1] > r0 && s >= t && bf.nz(s, t)) act = true;
        const int64_t lo = bf.start[t], hi = std::min(bf.start[t + 1], c);
        if (!act || hi <= lo) continue;
        if (!iv.empty() && iv.back().second == lo) iv.back().second = hi;
        else iv.emplace_back(lo, hi);
    }
    if (iv.size() == 1 && iv[0].first == 0 && iv[0].second == c) {
        iv.clear();
        return false;
    }
    return true;
}

// Tile rows (of `tile` rows, relative to row r) that the rank-S_KP update of the
// block group with columns [c0, c1) can touch: the segments whose rows may be
// non-zero below those columns.  Returns false when that is every row.
static bool active_tile_rows(const BlockFill& bf, int64_t n, int64_t c0, int64_t c1, int64_t r, int64_t tile,
                             TileRemap& rm, int64_t& rows) {
    const int ns = bf.nseg();
    std::vector<std::pair<int64_t, int64_t>> iv;   // tile intervals
    rows = 0;
    for (int s = 0; s < ns; ++s) {
        bool act = false;
        for (int t = 0; t < ns && !act; ++t)
            if (bf.start[t] < c1 && bf.start[t + 1] > c0 && bf.nz(s, t)) act = true;
        const int64_t lo = std::max(bf.start[s], r), hi = std::min(bf.start[s + 1], n);
        if (!act || hi <= lo) continue;
        rows += hi - lo;
        iv.emplace_back((lo - r) / tile, cdiv(hi - r, tile));
    }
    std::sort(iv.begin(), iv.end());
    std::vector<std::pair<int64_t, int64_t>> mg;
    for (const auto& v : iv) {
        if (!mg.empty() && v.first <= mg.back().second) mg.back().second = std::max(mg.back().second, v.second);
        else mg.push_back(v);
    }
    const int64_t ntm = cdiv(n - r, tile);
    if (mg.empty() || mg.size() > (size_t)MAX_REMAP || (mg.size() == 1 && mg[0].first == 0 && mg[0].second >= ntm))
        return false;
    rm.n = (int)mg.size();
    int64_t c = 0;
    for (int k = 0; k < rm.n; ++k) {
        rm.c[k] = c;
        rm.s[k] = mg[k].first;
        c += mg[k].second - mg[k].first;
    }
    rm.c[rm.n] = c;
    return true;
}

// Tile of the main split update: 128 (default) or 256 (OMH_SYRK_TILE=256;
// both give the same factor, each entry's products summed in the same
// order).  At c2 both run at ~850 TFLOP/s: the update is bound by operand
// bands fetched again from beyond L2 (the 256-row tiles halve the operand
// bytes per output but, one workgroup per CU, share fewer bands per XCD: 11%
// less fetch).
// The split update's slice clock (k-synchronised tile schedule); off with
// OMH_SYRK_CLOCK=0 (every tile then sums its k slices in order 0, 1, ...).
static bool slice_clock_on() {
    const char* e = std::getenv("OMH_SYRK_CLOCK");
    return !(e && e[0] == '0');
}

// 256 x 128 tiles for the main split update when it has enough of them
// (OMH_SYRK_WIDE=0: 128 x 128 tiles throughout)
static bool syrk_wide() {
    const char* e = std::getenv("OMH_SYRK_WIDE");
    return !(e && e[0] == '0');
}

// Tiles of k_syrk_split_wide over ntm units of 128 rows: row pair ta holds
// min(2 ta + 2, ntm) of them.
static int64_t wide_tiles(int64_t ntm) {
    int64_t t = 0;
    for (int64_t ta = 0; 2 * ta < ntm; ++ta) t += std::min<int64_t>(2 * ta + 2, ntm);
    return t;
}

// the software-pipelined 128 x 128 tile (OMH_SYRK_PP=0: the register-staged one)
static bool syrk_pipelined() {
    const char* e = std::getenv("OMH_SYRK_PP");
    return !(e && e[0] == '0');
}

static int syrk_tile() {
    const char* e = std::getenv("OMH_SYRK_TILE");
    return (e && std::atoi(e) == 256) ? 256 : 128;
}

// The end (exclusive) of the rows that L's blocks reach from columns [c0,
// c1): the columns' panels are zero below it (n without a block structure).
static int64_t reach_end(const BlockFill* bf, int64_t n, int64_t c0, int64_t c1) {
    if (!bf) return n;
    int64_t e = c1;
    const int ns = bf->nseg();
    for (int t = 0; t < ns; ++t) {
        if (bf->start[t] >= c1 || bf->start[t + 1] <= c0) continue;
        for (int s2 = t; s2 < ns; ++s2)
            if (bf->nz(s2, t)) e = std::max(e, bf->start[s2 + 1]);
    }
    return std::min(e, n);
}

static FactorInfo sym_factor_split(float* A, int64_t n, int64_t lda, void* work, hipStream_t st, const BlockFill* bf) {
    float* Dinv = (float*)work + (size_t)n * FACT_NB;
    const int64_t nblk = cdiv(n, FACT_NB);
    float* sgn = Dinv + (size_t)nblk * FACT_NB * FACT_NB;
    const int64_t mp_max = cdiv(n, S_TILE) * S_TILE;
    char* sp = (char*)work + (dense_workspace_bytes(n, 4) + 64 + 255) / 256 * 256;
    const size_t arr = (size_t)mp_max * S_KP * sizeof(_Float16);
    // two sets of split operands: set g % 2 holds group g's (row stride mp of
    // that group)
    struct SplitSet {
        _Float16 *Ah, *Al, *Bh, *Bl;
        float* sc;
        int64_t mp;
    } sets[2];
    auto place = [&](char* p, SplitSet& q) {
        q.Ah = (_Float16*)p;
        q.Al = (_Float16*)(p + arr);
        q.Bh = (_Float16*)(p + 2 * arr);
        q.Bl = (_Float16*)(p + 3 * arr);
        q.sc = (float*)(p + 4 * arr);
        q.mp = 0;
    };
    place(sp, sets[0]);
    float* Xbuf[S_PANELS];
    Xbuf[0] = (float*)work;
    Xbuf[1] = sgn + sign_span(n);
    {
        char* xp = (char*)(sets[0].sc + mp_max);
        for (int b = 2; b < S_PANELS; ++b) {
            xp = (char*)(((uintptr_t)xp + 255) / 256 * 256);
            Xbuf[b] = (float*)xp;
            xp += (size_t)n * FACT_NB * sizeof(float);
        }
        place((char*)(((uintptr_t)xp + 255) / 256 * 256), sets[1]);
    }
    int* info = (int*)scratch(SCR_SMALL, 128, st);
    k_set_int<<<1, 1, 0, st>>>(info, 0);
    OMH_LAUNCH_CHECK();
    // slice clock of the main update (two period slots, alternating between
    // launches: one is read while the other is written); the first launch
    // assumes 1.8 us per k-slice (180 ticks of 10 ns)
    unsigned* clk_slots = nullptr;
    if (slice_clock_on()) {
        clk_slots = (unsigned*)(info + 16);
        k_set_int<<<1, 1, 0, st>>>((int*)clk_slots, 180);
        k_set_int<<<1, 1, 0, st>>>((int*)clk_slots + 1, 180);
        OMH_LAUNCH_CHECK();
    }
    int clk_launch = 0;
    const bool pp = syrk_pipelined();
    const bool wide_ok = pp && syrk_wide();
    const float tiny = 1e-30f;
    LookAhead& la = lookahead();
    auto start_of = [&](int64_t b) { return b * FACT_NB; };
    auto width_of = [&](int64_t b) { return (int)std::min<int64_t>(FACT_NB, n - b * FACT_NB); };
    auto ldx_of = [&](int64_t e) { return (n - e + 3) / 4 * 4; };   // panel whose rows start at e
    // The end (exclusive) of the rows that L's blocks reach from group g's
    // columns: below it the group's panels, the in-group updates, the split
    // operands and the strip run; beyond it every panel of the group is zero
    // (n without a block structure).
    auto reach_of = [&](int64_t g) -> int64_t {
        return reach_end(bf, n, start_of(g * S_PANELS), std::min<int64_t>(start_of((g + 1) * S_PANELS), n));
    };
    // panel of block b into Xbuf[b % S_PANELS]: X = A21 Linv^T, L21 = X S1,
    // rows below the block up to re (its group's reach)
    auto panel = [&](int64_t b, int64_t re, hipStream_t s) {
        const int64_t k = start_of(b);
        const int nb = width_of(b);
        const int64_t e = k + nb, m = re - e;
        if (m <= 0) return;
        float* X = Xbuf[b % S_PANELS];
        const float* Li = Dinv + (size_t)b * FACT_NB * FACT_NB;
        float* A21 = A + e + k * lda;
        gemm<float>(false, true, false, m, nb, nb, 1.f, A21, lda, Li, FACT_NB, 0.f, X, ldx_of(e), s);
        k_scale_cols<float><<<grid_for(m * nb), 256, 0, s>>>(X, ldx_of(e), A21, lda, m, nb, sgn + k);
        OMH_LAUNCH_CHECK();
    };
    // C(rows [r, re), cols [r, r + w)) -= X (rows from r) L21^T with the
    // panel of (full) block b: lower w x w block plus the rows below (re >= r
    // + w: the group's reach covers its own columns)
    auto update_cols = [&](int64_t b, int64_t r, int64_t w, int64_t re, hipStream_t s) {
        const int64_t k = start_of(b), e = k + FACT_NB;
        const int64_t ldx = ldx_of(e);
        const float* X = Xbuf[b % S_PANELS] + (r - e);
        const float* L21 = A + r + k * lda;
        float* C = A + r + r * lda;
        gemm<float>(false, true, true, w, w, FACT_NB, -1.f, X, ldx, L21, lda, 1.f, C, lda, s);
        if (re - r > w) gemm<float>(false, true, false, re - r - w, w, FACT_NB, -1.f, X + w, ldx, L21, lda, 1.f,
                                    C + w, lda, s);
    };
    // everything of group g that touches only its own columns
    auto group_prologue = [&](int64_t g, hipStream_t s) {
        const int64_t b0 = g * S_PANELS, b1 = std::min<int64_t>(b0 + S_PANELS, nblk);
        const int64_t gend = std::min<int64_t>(start_of(b1), n);
        const int64_t re = reach_of(g);
        for (int64_t b = b0; b < b1; ++b) {
            diag_factor<float>(A, lda, start_of(b), width_of(b), Dinv, sgn, info, tiny, s);
            panel(b, re, s);
            const int64_t r = start_of(b) + width_of(b);
            if (b + 1 < b1 && r < gend) update_cols(b, r, gend - r, re, s);
        }
    };
    const int64_t ngroups = cdiv(nblk, S_PANELS);
    // rows of group g's update: the next group's columns start at r0 (a
    // strip of w2 columns), the rest of the trailing matrix at r2; the
    // group's panels reach the rows below re (the operands are zero beyond);
    // the strip covers at least the next group's own columns
    struct GroupRows {
        int64_t r0, w2, re, r2, mb, mx, ms, mbr, mp;
    };
    auto rows_of = [&](int64_t g) {
        GroupRows q;
        q.r0 = start_of((g + 1) * S_PANELS);
        q.w2 = std::min<int64_t>(S_KP, n - q.r0);
        q.re = reach_of(g);
        q.r2 = q.r0 + q.w2;
        q.mb = n - q.r2;
        q.mx = std::max<int64_t>(q.re - q.r0, 0);
        q.ms = std::max(q.mx, q.w2);
        q.mbr = std::max<int64_t>(q.re - q.r2, 0);
        // operand rows (zero padded): whole 128-row tiles of the strip and
        // whole 256-row tiles of the main update's reached rows
        q.mp = std::max(cdiv(q.ms, S_TILE) * S_TILE, cdiv(q.w2 + cdiv(q.mbr, S2_TILE) * S2_TILE, S_TILE) * S_TILE);
        return q;
    };
    // split operands of group g (its S_PANELS panels) into set g % 2
    auto split_of = [&](int64_t g, hipStream_t s) {
        const GroupRows q = rows_of(g);
        SplitSet& o = sets[g & 1];
        o.mp = q.mp;
        SplitSrc src;
        for (int i = 0; i < S_PANELS; ++i) {
            const int64_t b = g * S_PANELS + i, e = start_of(b) + FACT_NB;
            src.X[i] = Xbuf[b % S_PANELS];
            src.ldx[i] = ldx_of(e);
            src.off[i] = q.r0 - e;
            src.sgn[i] = sgn + start_of(b);
        }
        k_split_group<<<(unsigned)(q.mp / 64), 64 * 4 * S_PANELS, 0, s>>>(src, q.mx, q.mp, o.Ah, o.Al, o.Bh, o.Bl,
                                                                          o.sc);
        OMH_LAUNCH_CHECK();
    };
    // group g has an update (its columns are not the last ones)
    auto has_update = [&](int64_t g) { return g + 1 < ngroups && start_of((g + 1) * S_PANELS) < n; };
    group_prologue(0, st);
    if (has_update(0)) split_of(0, st);
    for (int64_t g = 0; g + 1 < ngroups; ++g) {
        const int64_t b0 = g * S_PANELS;
        if (!has_update(g)) break;
        const GroupRows q = rows_of(g);
        const int64_t r0 = q.r0, w2 = q.w2, r2 = q.r2, mb = q.mb, ms = q.ms;
        const SplitSet& o = sets[g & 1];
        const int64_t mp = o.mp;
        _Float16 *Ah = o.Ah, *Al = o.Al, *Bh = o.Bh, *Bl = o.Bl;
        float* sc = o.sc;
        // the next group's columns: a strip of whole tile columns (w2 = S_KP
        // whenever rows remain below it, else the strip is the remainder)
        {
            const int64_t ntm = cdiv(ms, S_TILE), nc = cdiv(w2, S_TILE);
            const int64_t tiles = nc * (nc + 1) / 2 + (ntm - nc) * nc;
            const TileMap map{6, ntm, 0, nc, tiles, cdiv(tiles, 8)};
            ProfScope ps("factor_strip_f32s", (double)(2 * ms - w2 + 1) * (double)w2 * (double)S_KP, st);
            if (pp)
                k_syrk_strip<S_KP, true><<<(unsigned)(8 * map.per), 256, 0, st>>>(ms, mp, 0, Ah, Al, Bh, Bl, sc,
                                                                                A + r0 + r0 * lda, lda, map);
            else
                k_syrk_strip<S_KP, false><<<(unsigned)(8 * map.per), 256, 0, st>>>(ms, mp, 0, Ah, Al, Bh, Bl, sc,
                                                                                 A + r0 + r0 * lda, lda, map);
            OMH_LAUNCH_CHECK();
        }
        // the next group's prologue and its split operands (into the other
        // set) on the side stream: its columns are outside the SYRK's region,
        // and the panel buffers and that set are free once this group's
        // split and the previous group's update are done (both precede this
        // point on the main stream)
        OMH_HIP(hipEventRecord(la.ready, st));
        OMH_HIP(hipStreamWaitEvent(la.side, la.ready, 0));
        group_prologue(g + 1, la.side);
        if (has_update(g + 1)) split_of(g + 1, la.side);
        OMH_HIP(hipEventRecord(la.done, la.side));
        // the rest of the trailing matrix
        if (mb > 0) {
            // only the tile rows the group's panels can reach (block
            // structure); 256 x 256 tiles while they give >= 512 workgroups
            const int64_t c0 = start_of(b0), c1 = std::min(n, start_of(b0 + S_PANELS));
            TileRemap rm;
            int64_t rows = mb;
            int64_t ntm = cdiv(mb, S2_TILE);
            if (bf && active_tile_rows(*bf, n, c0, c1, r2, S2_TILE, rm, rows)) ntm = rm.c[rm.n];
            bool big = syrk_tile() == 256 && ntm * (ntm + 1) / 2 >= 512;
            if (!big) {
                rm = TileRemap{};
                rows = mb;
                ntm = cdiv(mb, S_TILE);
                if (bf && active_tile_rows(*bf, n, c0, c1, r2, S_TILE, rm, rows)) ntm = rm.c[rm.n];
            }
            const int64_t tiles = ntm * (ntm + 1) / 2;
            const TileMap map = tiles >= 1024 ? TileMap{4, ntm, 0, 0, tiles, cdiv(tiles, 8)}
                                              : TileMap{1, ntm, 0, 0, tiles, 0};
            const unsigned grid = (unsigned)(map.kind == 4 ? 8 * map.per : tiles);
            // 256 x 128 tiles (units of 128 rows, the same remap) while they
            // give at least 4 per CU
            const int64_t wtiles = (!big && wide_ok) ? wide_tiles(ntm) : 0;
            const bool wide = wtiles >= 1024;
            ProfScope ps("factor_update_f32s", (double)rows * (double)(rows + 1) * (double)S_KP, st);
            if (tiles > 0) {
                SliceClock clk;
                if (clk_slots && (map.kind == 4 || wide)) {
                    clk.period = clk_slots + (clk_launch & 1);
                    clk.period_out = clk_slots + ((clk_launch + 1) & 1);
                    ++clk_launch;
                }
                if (wide) {
                    const int64_t per = cdiv(wtiles, 8);
                    k_syrk_split_wide<S_KP><<<(unsigned)(8 * per), 512, 0, st>>>(
                        mb, mp, w2, Ah, Al, Bh, Bl, sc, A + r2 + r2 * lda, lda, ntm, wtiles, per, rm, clk);
                } else if (big)
                    k_syrk_split256<S_KP><<<grid, 512, 0, st>>>(mb, mp, w2, Ah, Al, Bh, Bl, sc, A + r2 + r2 * lda,
                                                                lda, map, rm, clk);
                else if (pp)
                    k_syrk_split<S_KP, true><<<grid, 256, 0, st>>>(mb, mp, w2, Ah, Al, Bh, Bl, sc, A + r2 + r2 * lda,
                                                                   lda, map, rm, clk);
                else
                    k_syrk_split<S_KP, false><<<grid, 256, 0, st>>>(mb, mp, w2, Ah, Al, Bh, Bl, sc, A + r2 + r2 * lda,
                                                                    lda, map, rm, clk);
            }
            OMH_LAUNCH_CHECK();
        }
        OMH_HIP(hipStreamWaitEvent(st, la.done, 0));
    }
    FactorInfo fi;
    int h_info = 0;
    OMH_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    fi.info = h_info;
    fi.dinv = Dinv;
    fi.signs = sgn;
    return fi;
}

template <typename T>
FactorInfo sym_factor(T* A, int64_t n, int64_t lda, void* work, hipStream_t st, bool split16, const BlockFill* bf) {
    T* Dinv = (T*)work + (size_t)n * FACT_NB;
    const int64_t nblk = cdiv(n, FACT_NB);
    T* sgn = Dinv + (size_t)nblk * FACT_NB * FACT_NB;
    T* Xbuf[2] = {(T*)work, sgn + sign_span(n)};
    if (split16 && sizeof(T) == 4) return sym_factor_split((float*)A, n, lda, work, st, bf);
    int* info = (int*)scratch(SCR_SMALL, 64, st);
    k_set_int<<<1, 1, 0, st>>>(info, 0);
    OMH_LAUNCH_CHECK();
    const T tiny = sizeof(T) == 8 ? T(1e-290) : T(1e-30);
    LookAhead& la = lookahead();
    const char* upd = sizeof(T) == 8 ? "factor_update_f64" : "factor_update_f32";
    // panel of block k: X = A21 Linv^T (leading dimension m rounded up to 4
    // for aligned vector loads), L21 = X S1 written over A21
    auto panel = [&](int64_t k, int nb, int b, hipStream_t s) {
        const int64_t e = k + nb, m = reach_end(bf, n, k, e) - e;   // rows the block's columns reach
        if (m <= 0) return;
        const T* Li = Dinv + (size_t)(k / FACT_NB) * FACT_NB * FACT_NB;
        const int64_t ldx = (n - e + 3) / 4 * 4;   // the layout the update reads (all rows below the block)
        T* A21 = A + e + k * lda;
        T* X = Xbuf[b];
        gemm<T>(false, true, false, m, nb, nb, T(1), A21, lda, Li, FACT_NB, T(0), X, ldx, s);
        k_scale_cols<T><<<grid_for(m * nb), 256, 0, s>>>(X, ldx, A21, lda, m, nb, sgn + k);
        OMH_LAUNCH_CHECK();
    };
    diag_factor<T>(A, lda, 0, (int)std::min<int64_t>(FACT_NB, n), Dinv, sgn, info, tiny, st);
    panel(0, (int)std::min<int64_t>(FACT_NB, n), 0, st);
    for (int64_t k = 0, it = 0; k < n; k += FACT_NB, ++it) {
        const int nb = (int)std::min<int64_t>(FACT_NB, n - k);
        const int64_t e = k + nb, m = n - e;
        if (m <= 0) break;
        T* X = Xbuf[it & 1];
        const int64_t ldx = (m + 3) / 4 * 4;
        T* A21 = A + e + k * lda;
        // trailing update A22 -= X L21^T (lower)
        const int nb2 = (int)std::min<int64_t>(FACT_NB, m);   // next block
        T* A22 = A + e + e * lda;
        // (a) columns of the next block, over the rows this block's panel
        // reaches (its rows beyond are zero and were not formed)
        const int64_t mr = reach_end(bf, n, k, e) - e;
        const int wa = (int)std::min<int64_t>(nb2, mr);
        if (wa > 0) gemm<T>(false, true, true, wa, wa, nb, T(-1), X, ldx, A21, lda, T(1), A22, lda, st);
        if (mr > nb2)
            gemm<T>(false, true, false, mr - nb2, nb2, nb, T(-1), X + nb2, ldx, A21, lda, T(1), A22 + nb2, lda, st);
        // factor the next block and form its panel on the side stream (it
        // touches only the next block's columns, which (b) does not).  Beside
        // the FP64 update the diagonal kernels wait for CU slots (~8 ms
        // instead of 0.42 ms alone: the inverse's 142 KB of LDS fit only on a
        // CU without update workgroups), yet the 4-layer 10242-vertex FP64
        // factorisation still takes 4.02 s this way against 4.20 s with the
        // diagonal block factored on the main stream before (b).
        OMH_HIP(hipEventRecord(la.ready, st));
        OMH_HIP(hipStreamWaitEvent(la.side, la.ready, 0));
        diag_factor<T>(A, lda, e, nb2, Dinv, sgn, info, tiny, la.side);
        panel(e, nb2, (int)((it + 1) & 1), la.side);
        OMH_HIP(hipEventRecord(la.done, la.side));
        // (b) the rest of the trailing matrix: with a block structure, only
        // the row intervals the panel can reach, pair by pair
        if (m > nb2) {
            std::vector<std::pair<int64_t, int64_t>> iv;
            if (!bf || !active_rows(*bf, n, k, e, e + nb2, iv)) {
                gemm<T>(false, true, true, m - nb2, m - nb2, nb, T(-1), X + nb2, ldx, A21 + nb2, lda, T(1),
                        A22 + nb2 + nb2 * lda, lda, st, upd);
            } else {
                for (size_t a = 0; a < iv.size(); ++a)
                    for (size_t b = 0; b <= a; ++b) {
                        const int64_t r0 = iv[a].first, r1 = iv[a].second, c0 = iv[b].first, c1 = iv[b].second;
                        // C[I_a, I_b] -= X[I_a] L21[I_b]^T
                        gemm<T>(false, true, a == b, r1 - r0, c1 - c0, nb, T(-1), X + (r0 - e), ldx, A + c0 + k * lda,
                                lda, T(1), A + r0 + c0 * lda, lda, st, upd);
                    }
            }
        }
        OMH_HIP(hipStreamWaitEvent(st, la.done, 0));
    }
    FactorInfo fi;
    int h_info = 0;
    OMH_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    fi.info = h_info;
    fi.dinv = Dinv;
    fi.signs = sgn;
    return fi;
}

int64_t factor_sign_mismatch(const void* work, int64_t n, size_t elem, int64_t n_pos, hipStream_t st) {
    std::vector<int8_t> expect((size_t)n, (int8_t)-1);
    for (int64_t i = 0; i < std::min(n, n_pos); ++i) expect[i] = 1;
    return factor_sign_mismatch(work, n, elem, expect.data(), st);
}

int64_t factor_sign_mismatch(const void* work, int64_t n, size_t elem, const int8_t* expect, hipStream_t st) {
    const int64_t nblk = cdiv(n, FACT_NB);
    const size_t off = elem * ((size_t)n * FACT_NB + (size_t)nblk * FACT_NB * FACT_NB);
    std::vector<double> s((size_t)n);
    if (elem == 8) {
        OMH_HIP(hipMemcpyAsync(s.data(), (const char*)work + off, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        OMH_HIP(hipStreamSynchronize(st));
    } else {
        std::vector<float> f((size_t)n);
        OMH_HIP(hipMemcpyAsync(f.data(), (const char*)work + off, sizeof(float) * n, hipMemcpyDeviceToHost, st));
        OMH_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < n; ++i) s[i] = f[i];
    }
    for (int64_t i = 0; i < n; ++i)
        if ((expect[i] > 0) != (s[i] > 0)) return i;
    return -1;
}

template FactorInfo sym_factor<double>(double*, int64_t, int64_t, void*, hipStream_t, bool, const BlockFill*);
template FactorInfo sym_factor<float>(float*, int64_t, int64_t, void*, hipStream_t, bool, const BlockFill*);

template <typename T>
int block_fill_violated(const T* A, int64_t n, int64_t lda, const BlockFill& bf, hipStream_t st) {
    int* flag = (int*)scratch(SCR_SMALL, 64, st);
    k_set_int<<<1, 1, 0, st>>>(flag, 0);
    OMH_LAUNCH_CHECK();
    const int ns = bf.nseg();
    for (int s = 0; s < ns; ++s)
        for (int t = 0; t < s; ++t) {
            if (bf.hnz(s, t)) continue;
            const int64_t r0 = bf.start[s], r1 = std::min(bf.start[s + 1], n);
            const int64_t c0 = bf.start[t], c1 = std::min(bf.start[t + 1], n);
            if (r1 <= r0 || c1 <= c0) continue;
            const int64_t m = r1 - r0, k = c1 - c0;
            k_any_nonzero<T><<<std::min<unsigned>(grid_for(m * k), 4096), 256, 0, st>>>(A + r0 + c0 * lda, lda, m, k,
                                                                                       flag);
            OMH_LAUNCH_CHECK();
        }
    int h = 0;
    OMH_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    return h;
}

template int block_fill_violated<float>(const float*, int64_t, int64_t, const BlockFill&, hipStream_t);
template int block_fill_violated<double>(const double*, int64_t, int64_t, const BlockFill&, hipStream_t);

// Diagonal blocks per group of the triangular solves.
#ifndef OMH_SOLVE_GROUP
#define OMH_SOLVE_GROUP 4
#endif
constexpr int SOLVE_GROUP = OMH_SOLVE_GROUP;

namespace {

// acc[c] = sum_k A(i, k) X[k][c] for the workgroup's thread i < m, k < kk,
// A(i, k) = A[i + k lda] (or A[k + i lda] when TRANS), X in LDS (read as a
// broadcast).  A is read straight from memory, SKU values of k in flight per
// thread (one workgroup per CU runs this, so the loads in flight hide the
// cache latency): coalesced across the threads, or (TRANS) 16-byte loads
// along each thread's own column when the column is aligned for them.
constexpr int SKU = 32;
template <typename T, bool TRANS, int SV>
__device__ __forceinline__ void group_product(const T* __restrict__ A, int64_t lda, int m, int kk, const T (*X)[SV],
                                              T acc[SV]) {
    constexpr int VW = 16 / sizeof(T);   // elements per 16-byte load
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    const int tid = threadIdx.x;
#pragma unroll
    for (int c = 0; c < SV; ++c) acc[c] = T(0);
    if (tid >= m) return;
    const T* a0 = TRANS ? A + (int64_t)tid * lda : A + tid;
    const int64_t ks = TRANS ? 1 : lda;
    const bool vload = TRANS && (lda % VW) == 0 && (reinterpret_cast<uintptr_t>(A) % 16) == 0;
    int k = 0;
    for (; k + SKU <= kk; k += SKU) {
        T a[SKU];
        if (vload) {
#pragma unroll
            for (int u = 0; u < SKU; u += VW) {
                const vec_t v = *reinterpret_cast<const vec_t*>(a0 + k + u);
#pragma unroll
                for (int e = 0; e < VW; ++e) a[u + e] = v[e];
            }
        } else {
#pragma unroll
            for (int u = 0; u < SKU; ++u) a[u] = a0[(int64_t)(k + u) * ks];
        }
#pragma unroll
        for (int u = 0; u < SKU; ++u)
#pragma unroll
            for (int c = 0; c < SV; ++c) acc[c] = fma(a[u], X[k + u][c], acc[c]);
    }
    for (; k < kk; ++k) {
        const T a = a0[(int64_t)k * ks];
#pragma unroll
        for (int c = 0; c < SV; ++c) acc[c] = fma(a, X[k][c], acc[c]);
    }
}

// Triangular solve inside one group of diagonal blocks (rows [k0, gend))
// for SV right-hand sides per workgroup (a narrow strip, so that a few
// hundred right-hand sides still spread over many CUs), the group's rows of
// W in LDS:
// forward, block by block, W_b <- Linv_b W_b, then W_b' -= L_b'b W_b for
// the group's later blocks; backward (BACK) W_b <- Linv_b^T W_b, then
// W_b' -= L_bb'^T W_b for its earlier blocks.
// Split-FP16 copy of a solved group's rows of W (the B operand of the split
// solve products): column c scaled by tw[c], a power of two at or above the
// column's largest magnitude over the group's rows, then W / tw[c] = h + l
// with h and l in FP16, stored slice-major like the split factor's operands:
// (row i, column c) at ((i / 32) rp + c) 32 + i % 32.
struct WSplit {
    _Float16* h = nullptr;
    _Float16* l = nullptr;
    float* tw = nullptr;
    int64_t rp = 0;   // columns per slice (the right-hand sides rounded up to 256)
};

template <typename T, bool BACK, int SV>
__global__ void __launch_bounds__(256) k_solve_group(const T* __restrict__ L, int64_t lda, const T* __restrict__ Dinv,
                                                     int64_t k0, int64_t gend, T* __restrict__ W, int64_t ldw,
                                                     int64_t r, WSplit ws) {
    __shared__ T Wg[SOLVE_GROUP * FACT_NB][SV];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * SV;
    const int nc = (int)min<int64_t>(SV, r - c0);
    const int g = (int)(gend - k0);
    for (int e = tid; e < g * SV; e += 256) {
        const int i = e % g, c = e / g;
        Wg[i][c] = c < nc ? W[k0 + i + (c0 + c) * ldw] : T(0);
    }
    __syncthreads();
    const int nbk = (g + FACT_NB - 1) / FACT_NB;
    T acc[SV];
    for (int bi = 0; bi < nbk; ++bi) {
        const int b = BACK ? nbk - 1 - bi : bi;
        const int64_t kb = k0 + (int64_t)b * FACT_NB;
        const int nb = (int)min<int64_t>(FACT_NB, gend - kb);
        const T* Li = Dinv + (size_t)(kb / FACT_NB) * FACT_NB * FACT_NB;
        group_product<T, BACK, SV>(Li, FACT_NB, nb, nb, &Wg[b * FACT_NB], acc);
        __syncthreads();   // every thread has read W_b
        if (tid < nb) {
#pragma unroll
            for (int c = 0; c < SV; ++c) Wg[b * FACT_NB + tid][c] = acc[c];
        }
        __syncthreads();
        for (int b2 = BACK ? 0 : b + 1; BACK ? b2 < b : b2 < nbk; ++b2) {
            const int64_t kb2 = k0 + (int64_t)b2 * FACT_NB;
            const int nb2 = (int)min<int64_t>(FACT_NB, gend - kb2);
            const T* A = BACK ? L + kb + kb2 * lda : L + kb2 + kb * lda;
            group_product<T, BACK, SV>(A, lda, nb2, nb, &Wg[b * FACT_NB], acc);
            if (tid < nb2) {
#pragma unroll
                for (int c = 0; c < SV; ++c) Wg[b2 * FACT_NB + tid][c] -= acc[c];
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < g * SV; e += 256) {
        const int i = e % g, c = e / g;
        if (c < nc) W[k0 + i + (c0 + c) * ldw] = Wg[i][c];
    }
    if (sizeof(T) == 4 && ws.h) {   // FP32 solves only: the split copy for the products that follow
        __shared__ float red[SV][256];
#pragma unroll
        for (int c = 0; c < SV; ++c) {
            float mx = 0.f;
            for (int i = tid; i < g; i += 256) mx = fmaxf(mx, fabsf((float)Wg[i][c]));
            red[c][tid] = mx;
        }
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h)
#pragma unroll
                for (int c = 0; c < SV; ++c) red[c][tid] = fmaxf(red[c][tid], red[c][tid + h]);
            __syncthreads();
        }
        for (int e = tid; e < g * SV; e += 256) {
            const int i = e % g, c = e / g;
            int ex = 0;
            if (red[c][0] > 0.f) (void)frexpf(red[c][0], &ex);   // max = f 2^ex, f in [0.5, 1)
            const float x = (float)Wg[i][c] * ldexpf(1.f, -ex);
            const _Float16 hx = (_Float16)x;
            const int64_t o = ((int64_t)(i / 32) * ws.rp + c0 + c) * 32 + i % 32;
            ws.h[o] = hx;
            ws.l[o] = (_Float16)(x - (float)hx);
            if (i == 0) ws.tw[c0 + c] = ldexpf(1.f, ex);
        }
    }
}

// Row intervals [start, start + len) of a split solve product, cut into
// tiles of `tile` rows: tile x lies in the interval k with tile0[k] <= x <
// tile0[k + 1].
constexpr int MAX_SPANS = 8;
struct Spans {
    int n = 0;
    int64_t start[MAX_SPANS] = {};
    int64_t len[MAX_SPANS] = {};
    int64_t tile0[MAX_SPANS + 1] = {};
};

__device__ __forceinline__ bool span_tile(const Spans& sp, int64_t x, int tile, int64_t& row0, int64_t& rend) {
    if (x >= sp.tile0[sp.n]) return false;
    int k = 0;
    while (k + 1 < sp.n && x >= sp.tile0[k + 1]) ++k;
    row0 = sp.start[k] + (x - sp.tile0[k]) * tile;
    rend = sp.start[k] + sp.len[k];
    return true;
}

// Power of two at or above a non-negative value's magnitude (1 for 0), so
// that value / scale lies in [0, 1).
__device__ __forceinline__ float pow2_above(float mx) {
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);
    return ldexpf(1.f, e);
}

// Forward scales: s[i] = pow2_above(max_j |L(i, k0 + j)|, j < SOLVE_GROUP NB)
// for the rows i of the spans (64-row tiles; 256 threads = 64 rows x 4
// quarters of the group's columns).
__global__ void __launch_bounds__(256) k_span_row_scale(const float* __restrict__ L, int64_t lda, int64_t k0, Spans sp,
                                                        float* __restrict__ s) {
    __shared__ float red[4][64];
    constexpr int GW = SOLVE_GROUP * FACT_NB;
    int64_t row0, rend;
    if (!span_tile(sp, blockIdx.x, 64, row0, rend)) return;
    const int ri = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t i = row0 + ri;
    float mx = 0.f;
    if (i < rend) {
        const float* p = L + i + (k0 + (int64_t)q * (GW / 4)) * lda;
#pragma unroll 8
        for (int j = 0; j < GW / 4; ++j) mx = fmaxf(mx, fabsf(p[(int64_t)j * lda]));
    }
    red[q][ri] = mx;
    __syncthreads();
    if (q == 0 && i < rend) s[i] = pow2_above(fmaxf(fmaxf(red[0][ri], red[1][ri]), fmaxf(red[2][ri], red[3][ri])));
}

// Backward scales: s[j] = pow2_above(max_i |L(k0 + i, j)|, i < SOLVE_GROUP NB)
// for the columns j of the spans (4-column tiles, one wave per column; L's
// leading dimension and k0 are multiples of 4 and L is 16-byte aligned).
__global__ void __launch_bounds__(256) k_span_col_scale(const float* __restrict__ L, int64_t lda, int64_t k0, Spans sp,
                                                        float* __restrict__ s) {
    constexpr int GW = SOLVE_GROUP * FACT_NB;
    int64_t col0, cend;
    if (!span_tile(sp, blockIdx.x, 4, col0, cend)) return;
    const int lane = threadIdx.x & 63;
    const int64_t j = col0 + (threadIdx.x >> 6);
    if (j >= cend) return;
    const float4_t* p = reinterpret_cast<const float4_t*>(L + k0 + j * lda);
    float mx = 0.f;
#pragma unroll
    for (int u = 0; u < GW / 256; ++u) {
        const float4_t v = p[lane + 64 * u];
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) s[j] = pow2_above(mx);
}

// Split-FP16 product of the triangular solves:
//   W(row, c) -= s[row] tw[c] sum_k A(row, k) / s[row] . (h + l)(k, c)
// over k < SOLVE_GROUP NB, for the rows of the spans (128-row tiles) and the
// columns c < r (256-column tiles), with
//   forward (TRANS = false): A(row, k) = L(row, k0 + k), the group's columns
//     of the rows below it, s the row scales;
//   backward (TRANS = true): A(row, k) = L(k0 + k, row), the group's rows of
//     the earlier columns, s the column scales;
// and (h + l) = W(k0 + k, c) / tw[c] the group's split rows of W.  A / s is
// split into FP16 h + l as it is staged into LDS; the three products
// Al Bh + Ah Bl + Ah Bh are summed in FP32 by MFMA.  512 threads, 2 x 4 waves
// of 64 x 64, so that a tile of L is read once for 256 right-hand sides;
// k-slices of 32 double-buffered in LDS (the XOR swizzle of the split
// factor's update), operands staged through two register slots (slot s % 2
// holds slice s, loaded two slices ahead).  Workgroups go to the XCDs in runs
// of consecutive tiles (the tiles of one row tile share its A rows).
constexpr int SW_COLS = 256;
template <bool TRANS>
__global__ void __launch_bounds__(512) k_solve_split_wide(const float* __restrict__ L, int64_t lda, int64_t k0,
                                                          Spans sp, int64_t nct, int64_t per,
                                                          const float* __restrict__ s, WSplit ws,
                                                          float* __restrict__ W, int64_t ldw, int64_t r) {
    __shared__ half8_t SA[2][2][S_TILE * 4];
    __shared__ half8_t SB[2][2][SW_COLS * 4];
    constexpr int NKS = SOLVE_GROUP * FACT_NB / SK;
    const int64_t t = xcd_tile(blockIdx.x, per);
    int64_t row0, rend;
    if (!span_tile(sp, t / nct, S_TILE, row0, rend)) return;
    const int64_t c0 = (t % nct) * SW_COLS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    // A staging: forward, row tid % 128 and k in [8 (tid / 128), + 8);
    // backward, row tid / 4 and k in [8 (tid % 4), + 8)
    const int ar = TRANS ? tid >> 2 : tid & 127, akq = TRANS ? tid & 3 : tid >> 7;
    const int64_t arow = row0 + ar;
    const bool a_in = arow < rend;
    const float a_inv = a_in ? 1.f / s[arow] : 0.f;   // exact: s is a power of two
    const float* arow_p = TRANS ? L + arow * lda : L + arow;
    float av[2][8];
    half8_t bv[2][2][2];
    auto gload = [&](int ks, int slot) {
        const int64_t kb = k0 + (int64_t)ks * SK + 8 * akq;
        if (TRANS) {
            const float4_t* p = reinterpret_cast<const float4_t*>(arow_p + kb);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const float4_t v = a_in ? p[u] : float4_t{0.f, 0.f, 0.f, 0.f};
                av[slot][4 * u] = v[0];
                av[slot][4 * u + 1] = v[1];
                av[slot][4 * u + 2] = v[2];
                av[slot][4 * u + 3] = v[3];
            }
        } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) av[slot][u] = a_in ? arow_p[(kb + u) * lda] : 0.f;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int q = tid + 512 * p, cr = q >> 2, ch = q & 3;
                const _Float16* src = a == 0 ? ws.h : ws.l;
                bv[slot][a][p] =
                    *reinterpret_cast<const half8_t*>(src + (((int64_t)ks * ws.rp + c0 + cr) * SK + 8 * ch));
            }
    };
    auto lstore = [&](int slot, int buf) {
        half8_t h, l;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float x = av[slot][e] * a_inv;
            const _Float16 hx = (_Float16)x;
            h[e] = hx;
            l[e] = (_Float16)(x - (float)hx);
        }
        SA[buf][0][swz(ar, akq)] = h;
        SA[buf][1][swz(ar, akq)] = l;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int q = tid + 512 * p, cr = q >> 2, ch = q & 3;
                SB[buf][a][swz(cr, ch)] = bv[slot][a][p];
            }
    };
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    gload(0, 0);
    lstore(0, 0);
    gload(1, 1);
    gload(2, 0);
    __syncthreads();
    const int fr = lane & 15, fc = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {   // unrolled: register slots stay static
        const int buf = ks & 1;
        half8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            ah[a] = SA[buf][0][swz(wm + 16 * a + fr, fc)];
            al[a] = SA[buf][1][swz(wm + 16 * a + fr, fc)];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            bh[b] = SB[buf][0][swz(wn + 16 * b + fr, fc)];
            bl[b] = SB[buf][1][swz(wn + 16 * b + fr, fc)];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        // slice ks + 1 (register slot (ks + 1) % 2) into the other buffer,
        // then that slot gets slice ks + 3
        if (ks + 1 < NKS) lstore((ks + 1) & 1, buf ^ 1);
        __syncthreads();
        if (ks + 3 < NKS) gload(ks + 3, (ks + 1) & 1);
    }
    // epilogue: lane holds rows 4 (lane / 16) + e of column lane % 16 of a block
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t c = c0 + wn + 16 * b + fr;
        if (c >= r) continue;
        const float tc = ws.tw[c];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = row0 + wm + 16 * a + 4 * fc + e;
                if (i < rend) W[i + c * ldw] -= s[i] * tc * acc[a][b][e];
            }
    }
}
}  // namespace

// Buffers of the split-FP16 solve products (float solves): the split rows of
// the current group of W and one scale per row or column of L.
struct SplitSolveBuf {
    WSplit ws;
    float* s = nullptr;
    // optional: the products' row (forward) and column (backward) scales of
    // every group, n each (they depend on L only); `cached` once a solve has
    // filled them, so that the next solve with the same factor and block
    // structure reads them instead
    float* cache = nullptr;
    int64_t cache_n = 0, cache_groups = 0;
    bool cached = false;
};

size_t solve_scale_cache_bytes(int64_t n) {
    return 2 * (size_t)cdiv(n, (int64_t)SOLVE_GROUP * FACT_NB) * (size_t)n * sizeof(float) + 256;
}

size_t split_solve_bytes(int64_t n, int64_t r) {
    const size_t rp = (size_t)cdiv(r, SW_COLS) * SW_COLS;   // whole column tiles of the product kernel
    return 2 * (size_t)SOLVE_GROUP * FACT_NB * rp * sizeof(_Float16) + rp * sizeof(float) + (size_t)n * sizeof(float) +
           3 * 256;
}

static SplitSolveBuf split_solve_buffers(void* p, int64_t n, int64_t r) {
    SplitSolveBuf b;
    const size_t rp = (size_t)cdiv(r, SW_COLS) * SW_COLS;
    char* q = (char*)p;
    b.ws.rp = (int64_t)rp;
    b.ws.h = (_Float16*)q;
    q += (size_t)SOLVE_GROUP * FACT_NB * rp * sizeof(_Float16);
    b.ws.l = (_Float16*)q;
    q += (size_t)SOLVE_GROUP * FACT_NB * rp * sizeof(_Float16);
    b.ws.tw = (float*)q;
    q += (rp * sizeof(float) + 255) / 256 * 256;
    b.s = (float*)q;
    return b;
}

static Spans make_spans(const std::vector<std::pair<int64_t, int64_t>>& iv, int tile) {
    Spans sp;
    sp.n = (int)iv.size();
    for (int k = 0; k < sp.n; ++k) {
        sp.start[k] = iv[k].first;
        sp.len[k] = iv[k].second - iv[k].first;
        sp.tile0[k + 1] = sp.tile0[k] + cdiv(sp.len[k], tile);
    }
    return sp;
}

// W -= (split product of the group [k0, k0 + SOLVE_GROUP NB)) on the rows
// (forward: rows below the group, TRANS = false) or columns (backward:
// columns before it, TRANS = true) listed in iv.
// The array group k0's scales go to: its cache slot, or the one scratch
// array without a cache.
static float* scale_slot(bool trans, int64_t k0, const SplitSolveBuf& sb) {
    if (!sb.cache) return sb.s;
    return sb.cache +
           (size_t)((trans ? sb.cache_groups : 0) + k0 / ((int64_t)SOLVE_GROUP * FACT_NB)) * (size_t)sb.cache_n;
}

// Row (forward) or column (backward) scales of the group at k0 over the
// spans iv, into scl.
static void form_scales(bool trans, const float* L, int64_t lda, int64_t k0,
                        const std::vector<std::pair<int64_t, int64_t>>& iv, float* scl, hipStream_t st) {
    for (size_t p = 0; p < iv.size(); p += MAX_SPANS) {
        const std::vector<std::pair<int64_t, int64_t>> part(iv.begin() + p,
                                                             iv.begin() + std::min(iv.size(), p + MAX_SPANS));
        int64_t rows = 0;
        for (const auto& v : part) rows += v.second - v.first;
        if (rows <= 0) continue;
        const Spans sc = make_spans(part, trans ? 4 : 64);
        if (trans) k_span_col_scale<<<(unsigned)sc.tile0[sc.n], 256, 0, st>>>(L, lda, k0, sc, scl);
        else k_span_row_scale<<<(unsigned)sc.tile0[sc.n], 256, 0, st>>>(L, lda, k0, sc, scl);
        OMH_LAUNCH_CHECK();
    }
}

// ready: the scales were formed ahead on another stream; wait for them
// instead of forming them here.
static void split_solve_update(bool trans, const float* L, int64_t lda, int64_t k0,
                               const std::vector<std::pair<int64_t, int64_t>>& iv, const SplitSolveBuf& sb, float* W,
                               int64_t ldw, int64_t r, hipStream_t st, hipEvent_t ready = nullptr) {
    float* scl = scale_slot(trans, k0, sb);
    if (ready) OMH_HIP(hipStreamWaitEvent(st, ready, 0));
    else if (!sb.cache || !sb.cached) form_scales(trans, L, lda, k0, iv, scl, st);
    for (size_t p = 0; p < iv.size(); p += MAX_SPANS) {
        const std::vector<std::pair<int64_t, int64_t>> part(iv.begin() + p,
                                                             iv.begin() + std::min(iv.size(), p + MAX_SPANS));
        int64_t rows = 0;
        for (const auto& v : part) rows += v.second - v.first;
        if (rows <= 0) continue;
        const Spans sp = make_spans(part, S_TILE);
        ProfScope ps("solve_update_f32s", 2.0 * (double)rows * (double)r * (double)(SOLVE_GROUP * FACT_NB), st);
        const int64_t nct = cdiv(r, SW_COLS), total = sp.tile0[sp.n] * nct, per = cdiv(total, 8);
        if (trans)
            k_solve_split_wide<true><<<(unsigned)(8 * per), 512, 0, st>>>(L, lda, k0, sp, nct, per, scl, sb.ws, W,
                                                                          ldw, r);
        else
            k_solve_split_wide<false><<<(unsigned)(8 * per), 512, 0, st>>>(L, lda, k0, sp, nct, per, scl, sb.ws, W,
                                                                           ldw, r);
        OMH_LAUNCH_CHECK();
    }
}

template <typename T>
static void sym_solve_impl(const T* L, int64_t n, int64_t lda, const void* work, T* W, int64_t ldw, int64_t r,
                           hipStream_t st, const BlockFill* bf, const SplitSolveBuf* sb) {
    if (r <= 0) return;
    const T* Dinv = (const T*)work + (size_t)n * FACT_NB;
    const int64_t nblk = cdiv(n, FACT_NB);
    const T* sgn = Dinv + (size_t)nblk * FACT_NB * FACT_NB;
    // Blocks are taken in groups of SOLVE_GROUP: inside a group block by
    // block in one kernel (k_solve_group: the diagonal inverse, then the
    // group's later rows), then the rest of W with the whole group in one
    // product of depth SOLVE_GROUP x NB.
    constexpr int64_t GW = (int64_t)SOLVE_GROUP * FACT_NB;
    std::vector<std::pair<int64_t, int64_t>> iv;
    // The split products need 16-byte aligned columns of L (backward) and
    // serve the full groups; the others take the FP32 products.
    const bool split_ok = sb && sizeof(T) == 4 && lda % 4 == 0 && reinterpret_cast<uintptr_t>(L) % 16 == 0;
    // strips of 2 right-hand sides up to 1024 of them (128 workgroups for
    // 256; 8-wide strips took twice as long per group at 256), else 8
    const bool narrow = r <= 1024;
    auto group_solve = [&](bool back, int64_t k0, int64_t gend, bool split) {
        WSplit ws;
        if (split) ws = sb->ws;
        if (narrow) {
            const unsigned g = (unsigned)cdiv(r, 2);
            if (back) k_solve_group<T, true, 2><<<g, 256, 0, st>>>(L, lda, Dinv, k0, gend, W, ldw, r, ws);
            else k_solve_group<T, false, 2><<<g, 256, 0, st>>>(L, lda, Dinv, k0, gend, W, ldw, r, ws);
        } else {
            const unsigned g = (unsigned)cdiv(r, 8);
            if (back) k_solve_group<T, true, 8><<<g, 256, 0, st>>>(L, lda, Dinv, k0, gend, W, ldw, r, ws);
            else k_solve_group<T, false, 8><<<g, 256, 0, st>>>(L, lda, Dinv, k0, gend, W, ldw, r, ws);
        }
        OMH_LAUNCH_CHECK();
    };
    // rows below the group at k0 (forward) / columns before it (backward):
    // with a block structure only those L's blocks reach
    auto rows_below = [&](int64_t k0, int64_t gend) {
        const bool some = bf && active_rows(*bf, n, k0, gend, gend, iv);
        if (!some) iv.assign(1, {gend, n});
    };
    auto cols_before = [&](int64_t k0, int64_t gend) {
        const bool some = bf && active_cols(*bf, k0, gend, k0, iv);
        if (!some) iv.assign(1, {0, k0});
    };
    // The products' scales depend on L only: when a scale cache is to be
    // filled, every group's are formed on the look-ahead stream ahead of the
    // products (beside the latency-bound group solves), with one event per
    // chunk of groups that the products wait for.
    const bool ahead = split_ok && sb->cache && !sb->cached;
    LookAhead* la = ahead ? &lookahead() : nullptr;
    const int64_t ngroups = cdiv(n, GW);
    const int64_t chunk = std::max<int64_t>(1, cdiv(ngroups, SCALE_EVENTS / 2));
    if (ahead) {
        OMH_HIP(hipEventRecord(la->ready, st));
        OMH_HIP(hipStreamWaitEvent(la->side, la->ready, 0));
        for (int64_t k0 = 0, gi = 0; k0 < n; k0 += GW, ++gi) {
            const int64_t gend = std::min<int64_t>(k0 + GW, n);
            if (gend - k0 == GW && gend < n) {
                rows_below(k0, gend);
                form_scales(false, (const float*)L, lda, k0, iv, scale_slot(false, k0, *sb), la->side);
            }
            if ((gi + 1) % chunk == 0 || gend >= n) OMH_HIP(hipEventRecord(la->scale_ev[gi / chunk], la->side));
        }
        for (int64_t k0 = ((n - 1) / GW) * GW, bi = 0; k0 >= 0; k0 -= GW, ++bi) {
            const int64_t gend = std::min<int64_t>(k0 + GW, n);
            if (gend - k0 == GW && k0 > 0) {
                cols_before(k0, gend);
                form_scales(true, (const float*)L, lda, k0, iv, scale_slot(true, k0, *sb), la->side);
            }
            if ((bi + 1) % chunk == 0 || k0 == 0)
                OMH_HIP(hipEventRecord(la->scale_ev[SCALE_EVENTS / 2 + bi / chunk], la->side));
        }
    }
    // forward: W <- L^-1 W
    for (int64_t k0 = 0, gi = 0; k0 < n; k0 += GW, ++gi) {
        const int64_t gend = std::min<int64_t>(k0 + GW, n);
        const bool split = split_ok && gend - k0 == GW && gend < n;
        group_solve(false, k0, gend, split);
        if (gend < n) {
            rows_below(k0, gend);
            if (split)
                split_solve_update(false, (const float*)L, lda, k0, iv, *sb, (float*)W, ldw, r, st,
                                   ahead ? la->scale_ev[gi / chunk] : nullptr);
            else
                for (const auto& v : iv)
                    gemm<T>(false, false, false, v.second - v.first, r, gend - k0, T(-1), L + v.first + k0 * lda,
                            lda, W + k0, ldw, T(1), W + v.first, ldw, st);
        }
    }
    // W <- S W
    k_scale_rows<T><<<grid_for(n * r), 256, 0, st>>>(W, ldw, n, r, sgn);
    OMH_LAUNCH_CHECK();
    // backward: W <- L^-T W (rows of a group already carry the later
    // groups' contributions)
    for (int64_t k0 = ((n - 1) / GW) * GW, bi = 0; k0 >= 0; k0 -= GW, ++bi) {
        const int64_t gend = std::min<int64_t>(k0 + GW, n);
        const bool split = split_ok && gend - k0 == GW && k0 > 0;
        group_solve(true, k0, gend, split);
        // rows above the group: only the columns its rows reach in L
        if (k0 > 0) {
            cols_before(k0, gend);
            if (split)
                split_solve_update(true, (const float*)L, lda, k0, iv, *sb, (float*)W, ldw, r, st,
                                   ahead ? la->scale_ev[SCALE_EVENTS / 2 + bi / chunk] : nullptr);
            else
                for (const auto& v : iv)
                    gemm<T>(true, false, false, v.second - v.first, r, gend - k0, T(-1), L + k0 + v.first * lda, lda,
                            W + k0, ldw, T(1), W + v.first, ldw, st);
        }
    }
}

template <typename T>
void sym_solve(const T* L, int64_t n, int64_t lda, const void* work, T* W, int64_t ldw, int64_t r, hipStream_t st,
               const BlockFill* bf) {
    sym_solve_impl<T>(L, n, lda, work, W, ldw, r, st, bf, nullptr);
}

template void sym_solve<double>(const double*, int64_t, int64_t, const void*, double*, int64_t, int64_t, hipStream_t,
                                const BlockFill*);
template void sym_solve<float>(const float*, int64_t, int64_t, const void*, float*, int64_t, int64_t, hipStream_t,
                               const BlockFill*);

void convert_f64_to_f32(const double* src, int64_t lds, float* dst, int64_t ldd, int64_t m, int64_t n, bool lower,
                        hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)std::min<int64_t>(cdiv(m, 256), 32), (unsigned)std::min<int64_t>(n, 16384));
    k_convert<double, float><<<grid, 256, 0, st>>>(src, lds, dst, ldd, m, n, lower ? 1 : 0);
    OMH_LAUNCH_CHECK();
}

std::vector<int32_t> segment_rowmap(const std::vector<int64_t>& start, const std::vector<int32_t>& order) {
    std::vector<int32_t> map;
    map.reserve((size_t)start.back());
    for (const int32_t s : order)
        for (int64_t i = start[s]; i < start[s + 1]; ++i) map.push_back((int32_t)i);
    return map;
}

bool sym_reorder_blocks(const double* src, int64_t lds, void* dst, int64_t ldd, int elem,
                        const std::vector<int64_t>& start, const std::vector<int32_t>& order, hipStream_t st) {
    const int ns = (int)order.size();
    if (ns * (ns + 1) / 2 > SG_MAX_BLOCKS) return false;
    std::vector<int64_t> fs(ns + 1, 0);
    for (int q = 0; q < ns; ++q) fs[q + 1] = fs[q] + start[order[q] + 1] - start[order[q]];
    SymGatherBlocks tb{};
    int nb = 0;
    int64_t hmax = 0, wmax = 0;
    for (int p = 0; p < ns; ++p)
        for (int q = p; q < ns; ++q) {
            tb.dr[nb] = fs[q];
            tb.dc[nb] = fs[p];
            tb.sr[nb] = start[order[q]];
            tb.sc[nb] = start[order[p]];
            tb.h[nb] = (int32_t)(fs[q + 1] - fs[q]);
            tb.w[nb] = (int32_t)(fs[p + 1] - fs[p]);
            tb.diag[nb] = q == p;
            hmax = std::max<int64_t>(hmax, tb.h[nb]);
            wmax = std::max<int64_t>(wmax, tb.w[nb]);
            ++nb;
        }
    const dim3 grid((unsigned)std::min<int64_t>(cdiv(hmax, 256), 32), (unsigned)std::min<int64_t>(wmax, 4096),
                    (unsigned)nb);
    if (elem == 8) k_sym_gather_blocks<double><<<grid, 256, 0, st>>>(src, lds, (double*)dst, ldd, tb);
    else k_sym_gather_blocks<float><<<grid, 256, 0, st>>>(src, lds, (float*)dst, ldd, tb);
    OMH_LAUNCH_CHECK();
    return true;
}

void sym_reorder(const double* src, int64_t lds, void* dst, int64_t ldd, int elem, int64_t n, const int32_t* rowmap,
                 hipStream_t st) {
    if (n <= 0) return;
    const dim3 grid((unsigned)std::min<int64_t>(cdiv(n, 256), 32), (unsigned)std::min<int64_t>(n, 16384));
    if (elem == 8) k_sym_gather<double><<<grid, 256, 0, st>>>(src, lds, (double*)dst, ldd, n, rowmap);
    else k_sym_gather<float><<<grid, 256, 0, st>>>(src, lds, (float*)dst, ldd, n, rowmap);
    OMH_LAUNCH_CHECK();
}

void rows_reorder_f64(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t ncol,
                      const int32_t* rowmap, bool scatter, hipStream_t st) {
    if (m <= 0 || ncol <= 0) return;
    k_rows_map<double, double><<<grid_for(m * ncol), 256, 0, st>>>(src, lds, dst, ldd, m, ncol, rowmap, scatter ? 1 : 0);
    OMH_LAUNCH_CHECK();
}

void copy_f64(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n, hipStream_t st) {
    k_copy<double><<<grid_for(m * n), 256, 0, st>>>(src, lds, dst, ldd, m, n);
    OMH_LAUNCH_CHECK();
}

// Frobenius norm through NORM_PARTS partial sums in `part` (device), added
// on the host.
template <typename T>
static double frob_norm_with(const T* a, int64_t lda, int64_t m, int64_t n, double* part, hipStream_t st) {
    const unsigned g = std::min<unsigned>(grid_for(m * n), (unsigned)NORM_PARTS);
    k_sumsq<T><<<g, 256, 0, st>>>(a, lda, m, n, part);
    OMH_LAUNCH_CHECK();
    double h[NORM_PARTS];
    OMH_HIP(hipMemcpyAsync(h, part, sizeof(double) * g, hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (unsigned i = 0; i < g; ++i) s += h[i];
    return std::sqrt(s);
}

double frob_norm_f64(const double* a, int64_t lda, int64_t m, int64_t n, hipStream_t st) {
    double* part = (double*)scratch(SCR_SMALL, sizeof(double) * NORM_PARTS, st);
    return frob_norm_with<double>(a, lda, m, n, part, st);
}

double frob_norm_f32(const float* a, int64_t lda, int64_t m, int64_t n, hipStream_t st) {
    double* part = (double*)scratch(SCR_SMALL, sizeof(double) * NORM_PARTS, st);
    return frob_norm_with<float>(a, lda, m, n, part, st);
}

RefineState refine_state(int it, double rel, double prev, double tol) {
    if (!(rel == rel)) return REFINE_STALLED;                                     // NaN
    if (rel <= tol) return REFINE_CONVERGED;                                      // the update itself
    if (it >= 1 && rel < prev && rel * (rel / prev) <= tol) return REFINE_CONVERGED;   // update x contraction
    if (it >= 2 && rel > 0.5 * prev)                                              // no longer contracting
        return rel <= std::min(1e3 * tol, REFINE_FLOOR) ? REFINE_CONVERGED : REFINE_STALLED;
    return REFINE_CONTINUE;
}

// Side streams (default priority) for independent products that are too
// small to fill the GPU one at a time.
constexpr int N_SIDE = 3;
struct SidePool {
    hipStream_t s[N_SIDE] = {};
    hipEvent_t fork = nullptr, join[N_SIDE] = {};
};

static SidePool& side_pool() {
    static std::mutex mu;
    static std::map<int, SidePool> per_dev;
    int dev = 0;
    OMH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    SidePool& p = per_dev[dev];
    if (!p.fork) {
        OMH_HIP(hipEventCreateWithFlags(&p.fork, hipEventDisableTiming));
        for (int i = 0; i < N_SIDE; ++i) {
            OMH_HIP(hipStreamCreateWithFlags(&p.s[i], hipStreamNonBlocking));
            OMH_HIP(hipEventCreateWithFlags(&p.join[i], hipEventDisableTiming));
        }
    }
    return p;
}

// R <- R - H X; with a block structure row segment by row segment, each
// run of consecutive non-zero column blocks of H in one product.  The row
// segments write disjoint rows of R and run on up to N_SIDE + 1 streams, so
// that the last, partly filled round of tiles of one product overlaps the
// others.
// XT (r x n, leading dimension r) = X^T for X n x r (leading dimension ldx):
// 32 x 32 tiles through LDS, reads and writes along contiguous memory.
__global__ void __launch_bounds__(256) k_transpose_f64(const double* __restrict__ X, int64_t ldx, int64_t n, int64_t r,
                                                       double* __restrict__ XT) {
    __shared__ double t[32][33];
    const int64_t i0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int q = ty; q < 32; q += 8) {
        const int64_t i = i0 + tx, j = j0 + q;
        t[q][tx] = (i < n && j < r) ? X[i + j * ldx] : 0.0;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const int64_t j = j0 + tx, i = i0 + q;
        if (i < n && j < r) XT[j + i * r] = t[tx][q];
    }
}

static void residual_update(const double* H, int64_t n, int64_t ldh, const double* X, int64_t ldx, int64_t r,
                            double* R, int64_t ldr, hipStream_t st, const BlockFill* bf) {
    if (!bf) {
        gemm<double>(false, false, false, n, r, n, -1.0, H, ldh, X, ldx, 1.0, R, ldr, st);
        return;
    }
    // X^T, so that the segment products are A B^T with contiguous rows of
    // both operands (the 128 x 128 FP64 tiles); the same sums in the same
    // order as the products with X
    double* XT = (double*)scratch(SCR_RESID_XT, sizeof(double) * (size_t)n * (size_t)r, st);
    k_transpose_f64<<<dim3((unsigned)cdiv(n, 32), (unsigned)cdiv(r, 32)), 256, 0, st>>>(X, ldx, n, r, XT);
    OMH_LAUNCH_CHECK();
    SidePool& sp = side_pool();
    OMH_HIP(hipEventRecord(sp.fork, st));
    for (int i = 0; i < N_SIDE; ++i) OMH_HIP(hipStreamWaitEvent(sp.s[i], sp.fork, 0));
    const int ns = bf->nseg();
    // row segments to streams, the largest products first, each onto the
    // stream with the least work so far
    std::vector<std::pair<double, int>> work;
    for (int s = 0; s < ns; ++s) {
        double w = 0.0;
        for (int t = 0; t < ns; ++t)
            if (bf->hnz(s, t)) w += (double)(bf->start[t + 1] - bf->start[t]);
        work.emplace_back(w * (double)(bf->start[s + 1] - bf->start[s]), s);
    }
    std::sort(work.begin(), work.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
        return a.first > b.first || (a.first == b.first && a.second < b.second);
    });
    double load[N_SIDE + 1] = {};
    for (const auto& ws : work) {
        const int s = ws.second;
        int q = 0;
        for (int i = 1; i <= N_SIDE; ++i)
            if (load[i] < load[q]) q = i;
        load[q] += ws.first;
        const int64_t r0 = bf->start[s], r1 = bf->start[s + 1];
        hipStream_t ss = q == 0 ? st : sp.s[q - 1];
        for (int t = 0; t < ns;) {
            if (!bf->hnz(s, t)) { ++t; continue; }
            int u = t + 1;
            while (u < ns && bf->hnz(s, u)) ++u;
            const int64_t c0 = bf->start[t], c1 = bf->start[u];
            // the row segments run concurrently: 128 x 128 tiles down to 64
            // of them per product (a 10242-row segment has 160), the rows
            // past the last whole tile row as a partial tile row
            if (!gemm_f64_dma(false, r1 - r0, r, c1 - c0, -1.0, H + r0 + c0 * ldh, ldh, XT + c0 * r, r, 1.0, R + r0,
                              ldr, ss, nullptr, 64, true))
                gemm<double>(false, true, false, r1 - r0, r, c1 - c0, -1.0, H + r0 + c0 * ldh, ldh, XT + c0 * r, r,
                             1.0, R + r0, ldr, ss);
            t = u;
        }
    }
    for (int i = 0; i < N_SIDE; ++i) {
        OMH_HIP(hipEventRecord(sp.join[i], sp.s[i]));
        OMH_HIP(hipStreamWaitEvent(st, sp.join[i], 0));
    }
}

int mixed_refine(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                 const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                 double* final_rel_update, bool* converged, hipStream_t st, const BlockFill* bf,
                 const FactorOrder* fo) {
    RefineBuffers buf;
    buf.R32 = (float*)scratch(SCR_LINALG, sizeof(float) * (size_t)n * (size_t)r, st);
    buf.R = (double*)scratch(SCR_ASSEMBLY, sizeof(double) * (size_t)n * (size_t)r, st);
    buf.part = (double*)scratch(SCR_SMALL, sizeof(double) * NORM_PARTS, st);
    buf.split = scratch(SCR_LINALG2, split_solve_bytes(n, r), st);
    const char* sce = std::getenv("OMH_SOLVE_SCALE_CACHE");   // 0: every solve forms its scales
    if (!(sce && sce[0] == '0'))
        buf.scale_cache = (float*)scratch(SCR_SOLVE_SCALES, solve_scale_cache_bytes(n), st);
    return mixed_refine_in(H, n, ldh, L32, ld32, work32, B, ldb, X, ldx, r, max_iter, tol, final_rel_update, converged,
                           buf, st, bf, fo);
}

int mixed_refine_in(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                    const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                    double* final_rel_update, bool* converged, const RefineBuffers& buf, hipStream_t st,
                    const BlockFill* bf, const FactorOrder* fo) {
    // the factor's unknown order: the residual R is in H's, the correction
    // solve runs in the factor's (R32 gathered, added back scattered)
    const BlockFill* bfl = fo ? &fo->fill : bf;
    const int32_t* rowmap = fo ? fo->rowmap : nullptr;
    float* R32 = buf.R32;
    double* R = buf.R;
    SplitSolveBuf sb;
    if (buf.split) {
        sb = split_solve_buffers(buf.split, n, r);
        if (buf.scale_cache) {
            sb.cache = buf.scale_cache;
            sb.cache_n = n;
            sb.cache_groups = cdiv(n, (int64_t)SOLVE_GROUP * FACT_NB);
        }
    }
    fill_f64(X, n, r, ldx, 0.0, st);
    copy_f64(B, ldb, R, n, n, r, st);
    int it = 0;
    double rel = 1.0, prev = 2.0;
    RefineState state = REFINE_CONTINUE;
    for (; it < max_iter; ++it) {
        if (rowmap) {
            k_rows_map<double, float><<<grid_for(n * r), 256, 0, st>>>(R, n, R32, n, n, r, rowmap, 0);
            OMH_LAUNCH_CHECK();
        } else {
            convert_f64_to_f32(R, n, R32, n, n, r, false, st);
        }
        sym_solve_impl<float>(L32, n, ld32, work32, R32, n, r, st, bfl, buf.split ? &sb : nullptr);
        sb.cached = sb.cache != nullptr;   // the same factor and structure on the next pass
        k_axpy_f32_to_f64<<<grid_for(n * r), 256, 0, st>>>(R32, n, X, ldx, n, r, rowmap);
        OMH_LAUNCH_CHECK();
        if (it == 0) {
            // X was zero, so the first update is X itself (relative size 1):
            // no norms, no wait for the device before the residual (an
            // all-zero right-hand side converges on the next pass instead)
            rel = 1.0;
        } else {
            const double dn = frob_norm_with<float>(R32, n, n, r, buf.part, st);
            const double xn = frob_norm_with<double>(X, ldx, n, r, buf.part, st);
            rel = xn > 0 ? dn / xn : (dn > 0 ? 1.0 : 0.0);
        }
        state = refine_state(it, rel, prev, tol);
        if (state != REFINE_CONTINUE) { ++it; break; }
        prev = rel;
        // R = B - H X
        copy_f64(B, ldb, R, n, n, r, st);
        residual_update(H, n, ldh, X, ldx, r, R, n, st, bf);
    }
    if (final_rel_update) *final_rel_update = rel;
    if (converged) *converged = state == REFINE_CONVERGED;
    return it;
}

void sparse_transpose_dense(const int64_t* d_rows, const double* d_vals, int64_t ne, int64_t n, double* out, int64_t ldo,
                            hipStream_t st) {
    fill_f64(out, n, ne, ldo, 0.0, st);
    if (ne <= 0) return;
    k_sparse_cols<<<(unsigned)cdiv(ne, 256), 256, 0, st>>>(d_rows, d_vals, ne, n, out, ldo);
    OMH_LAUNCH_CHECK();
}

void sparse_rows(const int64_t* d_cols, const double* d_vals, int64_t ne, const double* Hm, int64_t ldh, int64_t n,
                 double* out, int64_t ldo, hipStream_t st) {
    k_sparse_rows<<<grid_for(ne * n), 256, 0, st>>>(d_cols, d_vals, ne, Hm, ldh, n, out, ldo);
    OMH_LAUNCH_CHECK();
}

}  // namespace omh
