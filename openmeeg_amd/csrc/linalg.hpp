// Dense MFMA linear algebra used by the HeadMat solve and the gains.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace omh {

constexpr int FACT_NB = 256;   // outer block of the factorisation

// Tile enumeration of a GEMM launch: 0 full grid (ntm row tiles), 1 lower
// triangle of tiles, 2 one row strip (tile row tm0), 3 one column strip
// (tile column tn0); 4 and 5: the lower triangle / full grid in an L2-aware
// order (see map_tile), over a grid of 8 * per workgroups for ntiles tiles.
struct TileMap {
    int kind;
    int64_t ntm, tm0, tn0;
    int64_t ntiles = 0, per = 0;
};

// ---------------------------------------------------------------------------
// Grouped launches: one kernel launch runs the same operation on several
// independent matrices (e.g. the heads of a batch), each described by a
// descriptor in device memory.
// ---------------------------------------------------------------------------
struct GemmDesc {               // C = alpha op(A) op(B) + beta C on the tiles of map
    const void* A;
    const void* B;
    void* C;
    int64_t lda, ldb, ldc, m, n, k;
    double alpha, beta;
    TileMap map;
    int64_t tile0;              // index of the first tile of this group in the launch
};

struct DiagDesc {               // diagonal block of one factorisation (see sym_factor)
    void* A11;
    int64_t lda;
    int nb;
    void* Linv;                 // FACT_NB x FACT_NB, leading dimension FACT_NB
    void* sgn;                  // signs of the whole factorisation
    int* info;
    int64_t k0;                 // global index of the block's first column
};

enum ElemOp {
    ELEM_COPY_F64, ELEM_COPY_F32, ELEM_F64_TO_F32, ELEM_F64_TO_F32_LOWER, ELEM_ADD_F32_TO_F64,
    ELEM_SCALE_ROWS_F32, ELEM_SCALE_ROWS_F64, ELEM_SCALE_COLS_F32, ELEM_SCALE_COLS_F64, ELEM_ZERO_F64, ELEM_ZERO_INT
};

struct ElemDesc {               // dst (m x n, ldd) <- op(src (lds), aux)
    int op;
    const void* src;
    int64_t lds;
    void* dst;
    int64_t ldd;
    int64_t m, n;
    const void* aux;            // row or column scale factors
};

// One product of a grouped GEMM, before tiling.
struct GemmJob {
    int64_t m, n, k;
    double alpha;
    const void* A;
    int64_t lda;
    const void* B;
    int64_t ldb;
    double beta;
    void* C;
    int64_t ldc;
};

// A grouped launch: descriptors [off, off + ng) of a descriptor table.
struct GemmLaunch {
    int variant;                // encodes T, op(A), op(B), lower, tile size, edge handling
    size_t off;
    int ng;
    int64_t total;              // tiles
};

// Tiles the jobs (all with the same T = float if f32 else double, ta, tb,
// lower) into descriptors appended to `descs`: an interior launch with
// unguarded vector loads for the full tiles of aligned products and an edge
// launch with guarded loads for the rest (the rules of gemm()).
void build_grouped_gemm(const std::vector<GemmJob>& jobs, bool f32, bool ta, bool tb, bool lower, int wb,
                        std::vector<GemmDesc>& descs, std::vector<GemmLaunch>& out);
// Launches l with descriptors at d_descs (device); grid_cap > 0 caps the grid
// (workgroups then loop over tiles).
void launch_grouped_gemm(const GemmLaunch& l, const GemmDesc* d_descs, int64_t grid_cap, hipStream_t st);

// One matrix's rank-nb trailing update C -= X S X^T (lower triangle) as a
// split-FP16 product in a grouped launch: X (m x nb <= FACT_NB, FP32) is
// split into its FP16 parts (one power-of-two scale per row, slice-major,
// mp = m rounded up to 128 rows of zero padding), then every 128 x 128 tile of
// C's lower triangle gets the three products.
struct SplitDesc {
    const float* X;
    int64_t ldx;
    const float* sgn;               // the nb signs S
    int64_t m, mp;
    int nb;
    void* Ah;                       // 4 arrays of mp x FACT_NB FP16 values, then mp row scales
    float* C;
    int64_t ldc;
    int64_t blk0;                   // first 64-row block of this matrix in the split launch
    int64_t tile0;                  // first tile of this matrix in the update launch
};
// Bytes of SplitDesc::Ah for m rows.
size_t split_operand_bytes(int64_t m);

// Rank-1024 split update of one matrix of a batch (a group of S4_PANELS = 4
// panels of FACT_NB columns):
// row r of the update is row off[p] + r of panel p (FP32, leading dimension
// ldx[p]); sgn holds the 1024 signs of the group's columns.  Ah: 4 arrays of
// mp x 1024 FP16 values, then mp row scales (split4_operand_bytes).
constexpr int S4_PANELS = 4;
struct Split4Desc {
    const float* X[4];
    int64_t ldx[4];
    int64_t off[4];
    const float* sgn;
    int64_t m, mp;
    void* Ah;
    float* C;
    int64_t ldc;
    int64_t blk0;                   // first 64-row block of this matrix in the split launch
    int64_t tile0;                  // first tile of this matrix in the update launch
};
size_t split4_operand_bytes(int64_t m);
void launch_grouped_split4(const Split4Desc* d_descs, int ng, int64_t blocks, hipStream_t st);
void launch_grouped_split4_update(const Split4Desc* d_descs, int ng, int64_t tiles, hipStream_t st);
void launch_grouped_split(const SplitDesc* d_descs, int ng, int64_t blocks, hipStream_t st);
void launch_grouped_split_update(const SplitDesc* d_descs, int ng, int64_t tiles, hipStream_t st);
void launch_grouped_diag(bool f32, const DiagDesc* d_descs, int ng, hipStream_t st);
void launch_grouped_elem(const ElemDesc* d_descs, int ng, int64_t max_elems, hipStream_t st);

// C = alpha op(A) op(B) + beta C (column-major).  lower: square C, only
// entries with row >= col are computed (only ta=false, tb=true).
// prof: optional name under which the interior launch of a lower (SYRK)
// product is timed (ProfScope name + "_t<tile>").
template <typename T>
void gemm(bool ta, bool tb, bool lower, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
          const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st, const char* prof = nullptr);

// The same product (not lower) for callers outside the library: outputs of
// fewer than 512 64 x 64 tiles with an inner dimension of at least 4096 and
// 8 x the larger output side run split-K (chunks of k concurrently into the
// SCR_GEMM scratch slot, then summed in order).
template <typename T>
void gemm_entry(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t st);

struct FactorInfo {
    int info = 0;              // 0, or 1 + index of the first (near-)zero pivot
    const void* dinv = nullptr;
    const void* signs = nullptr;
};

// Bytes of workspace sym_factor / sym_solve need for an n x n matrix of
// element size elem.
size_t factor_workspace_bytes(int64_t n, size_t elem);

// Block structure of a symmetric matrix: the unknowns are cut into
// consecutive segments (start[s] .. start[s + 1], start[nseg] = n);
// hfill[s * nseg + t] (s > t) is 0 when the block (rows of s, columns of t)
// of the matrix is zero, and fill, its symbolic factorisation, is 0 when
// that block stays zero in L.  Factor and solve updates skip the work that
// only multiplies such blocks (it would add exact zeros).
struct BlockFill {
    std::vector<int64_t> start;
    std::vector<uint8_t> hfill, fill;
    int nseg() const { return (int)start.size() - 1; }
    bool nz(int s, int t) const { return s <= t || fill[(size_t)s * nseg() + t] != 0; }     // L, s >= t
    bool hnz(int s, int t) const {                                                          // the matrix
        if (s == t) return true;
        return s > t ? hfill[(size_t)s * nseg() + t] != 0 : hfill[(size_t)t * nseg() + s] != 0;
    }
    // fill from hfill by elimination in segment order
    void symbolic() {
        const int ns = nseg();
        fill = hfill;
        for (int k = 0; k < ns; ++k)
            for (int a = k + 1; a < ns; ++a) {
                if (!fill[(size_t)a * ns + k]) continue;
                for (int b = a + 1; b < ns; ++b)
                    if (fill[(size_t)b * ns + k]) fill[(size_t)b * ns + a] = 1;
            }
    }
};

// In-place pivot-free signed Cholesky H = L S L^T of the lower triangle of A
// (n x n, leading dimension lda).  The strict upper triangle is not touched.
// split16 (FP32 only): the trailing updates run as split-FP16 MFMA products
// (each panel row scaled and split into two FP16 parts, three products,
// FP32 accumulation; about 22 significant bits per operand).  bf: optional
// block structure whose zero blocks the trailing updates skip.
template <typename T>
FactorInfo sym_factor(T* A, int64_t n, int64_t lda, void* work, hipStream_t st, bool split16 = false,
                      const BlockFill* bf = nullptr);

// First unknown whose pivot sign in a finished factorisation (workspace
// `work` of sym_factor, element size elem) differs from +1 for the first
// n_pos unknowns and -1 for the rest, or -1.  In that pattern the matrix is
// quasi-definite (its leading n_pos block positive definite, the Schur
// complement negative definite) and the pivot-free factor is stable; a
// mismatch means that this order needs pivoting.
int64_t factor_sign_mismatch(const void* work, int64_t n, size_t elem, int64_t n_pos, hipStream_t st);
// The same against an expected sign per unknown (host array of n, +1 / -1).
int64_t factor_sign_mismatch(const void* work, int64_t n, size_t elem, const int8_t* expect, hipStream_t st);

// The unknown order a factor was computed in, when it differs from the
// matrix's: whole segments of the matrix reordered.  rowmap (device, n
// entries) gives, for each row of the factor's order, the matrix row it
// holds; fill is the block structure in the factor's order.
struct FactorOrder {
    const int32_t* rowmap = nullptr;
    BlockFill fill;
};
// rowmap of the segments start[0..nseg] (start[nseg] = n) taken in the
// order order[0..nseg-1] (a permutation of the segments).
std::vector<int32_t> segment_rowmap(const std::vector<int64_t>& start, const std::vector<int32_t>& order);
// dst's lower triangle (element size elem, 8 or 4) = the reordered matrix
// src(rowmap[i], rowmap[j]) of the full symmetric FP64 src.
void sym_reorder(const double* src, int64_t lds, void* dst, int64_t ldd, int elem, int64_t n, const int32_t* rowmap,
                 hipStream_t st);
// The same as block copies between segment pairs (no row map), for up to 10
// segments; false (nothing done) beyond.
bool sym_reorder_blocks(const double* src, int64_t lds, void* dst, int64_t ldd, int elem,
                        const std::vector<int64_t>& start, const std::vector<int32_t>& order, hipStream_t st);
// m x ncol rows: dst(i, :) = src(rowmap[i], :), or with scatter
// dst(rowmap[i], :) = src(i, :).
void rows_reorder_f64(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t ncol,
                      const int32_t* rowmap, bool scatter, hipStream_t st);

// 1 if any entry of the blocks a block structure declares zero (below the
// diagonal) is non-zero.
template <typename T>
int block_fill_violated(const T* A, int64_t n, int64_t lda, const BlockFill& bf, hipStream_t st);

// W <- H^-1 W for r right-hand sides with the factor of sym_factor (bf:
// optional block structure of H; the updates skip L's zero blocks).
template <typename T>
void sym_solve(const T* L, int64_t n, int64_t lda, const void* work, T* W, int64_t ldw, int64_t r, hipStream_t st,
               const BlockFill* bf = nullptr);

void convert_f64_to_f32(const double* src, int64_t lds, float* dst, int64_t ldd, int64_t m, int64_t n, bool lower,
                        hipStream_t st);
void copy_f64(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n, hipStream_t st);
double frob_norm_f64(const double* a, int64_t lda, int64_t m, int64_t n, hipStream_t st);
double frob_norm_f32(const float* a, int64_t lda, int64_t m, int64_t n, hipStream_t st);

// Stopping rule of the refinement after correction `it` (0-based) with
// relative size rel = ||dX||_F / ||X||_F (prev: that of the correction
// before).  REFINE_CONVERGED: rel <= tol; or the corrections contract
// (rel < prev) and the estimated error left, rel times the contraction
// ratio rel / prev, is <= tol; or they stopped contracting (rel > prev / 2
// from the third correction on) at the FP64 rounding floor, rel <=
// min(1e3 tol, REFINE_FLOOR).  REFINE_STALLED: they stopped contracting
// above that floor (a factor too coarse for the matrix, or divergence).
// REFINE_CONTINUE otherwise.
enum RefineState { REFINE_CONTINUE = 0, REFINE_CONVERGED = 1, REFINE_STALLED = 2 };
constexpr double REFINE_FLOOR = 1e-10;
RefineState refine_state(int it, double rel, double prev, double tol);

// Iterative refinement: X = H^-1 B with the FP32 factor (L32, work32) and
// FP64 residuals R = B - H X.  Stops by refine_state or after max_iter
// corrections; returns the number of corrections applied and sets
// *converged (REFINE_CONVERGED reached).  The triangular solves run the
// products of their full 1024-column groups as split-FP16 MFMA: the FP32
// factor's tiles over a power-of-two scale per row (forward) or column
// (backward) and the group's rows of W over one per column, each split into
// FP16 h + l, three products summed in FP32.
int mixed_refine(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                 const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                 double* final_rel_update, bool* converged, hipStream_t st, const BlockFill* bf = nullptr,
                 const FactorOrder* fo = nullptr);

// Device buffers of the refinement: the FP64 residual R (n x r), its FP32
// copy R32 (n x r) and NORM_PARTS doubles of norm partial sums.
constexpr int NORM_PARTS = 1024;
struct RefineBuffers {
    float* R32 = nullptr;
    double* R = nullptr;
    double* part = nullptr;
    // optional: buffers of the split-FP16 products of the FP32 triangular
    // solves; without them the solves use FP32 products
    void* split = nullptr;
    // optional: the split products' scales of every group (their row and
    // column scales depend on the factor only; solve_scale_cache_bytes(n)),
    // filled by the first solve of a call and read by the later ones
    float* scale_cache = nullptr;
};
size_t solve_scale_cache_bytes(int64_t n);
// Bytes of RefineBuffers::split for n unknowns and r right-hand sides.
size_t split_solve_bytes(int64_t n, int64_t r);
// mixed_refine with caller-given buffers (allocates nothing on the device;
// with a block structure the residual's side streams are created on first
// use).
int mixed_refine_in(const double* H, int64_t n, int64_t ldh, const float* L32, int64_t ld32, const void* work32,
                    const double* B, int64_t ldb, double* X, int64_t ldx, int64_t r, int max_iter, double tol,
                    double* final_rel_update, bool* converged, const RefineBuffers& buf, hipStream_t st,
                    const BlockFill* bf = nullptr, const FactorOrder* fo = nullptr);

// out(e, j) = sum_k vals(e,k) Hm(cols(e,k), j)  (3 non-zeros per row).
// out (n x ne, leading dimension ldo) = the dense transpose of a sparse
// ne x n matrix with 3 entries per row (rows(e, k), vals(e, k)).
void sparse_transpose_dense(const int64_t* d_rows, const double* d_vals, int64_t ne, int64_t n, double* out, int64_t ldo,
                            hipStream_t st);
void sparse_rows(const int64_t* d_cols, const double* d_vals, int64_t ne, const double* Hm, int64_t ldh, int64_t n,
                 double* out, int64_t ldo, hipStream_t st);

}  // namespace omh
