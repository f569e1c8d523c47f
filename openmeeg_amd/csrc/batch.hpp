// Batched EEG gains over many heads in one (graph-capturable) pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "geometry.hpp"
#include "linalg.hpp"

namespace omh {

struct BatchHeadInput {
    Geometry* g = nullptr;          // not owned; must outlive the plan
    std::vector<double> dipoles;    // n_dip x 6 (x y z qx qy qz)
    std::vector<double> electrodes; // n_el x 3
};

struct BatchTuneResult {
    int tile;                       // 64 or 128
    int64_t grid_cap;               // 0: one workgroup per tile
    bool split16;                   // split-FP16 update (rank-1024 super-steps, 128 x 128 tiles); seconds per
                                    // 256 columns
    double seconds;                 // sampled trailing updates, per repetition
};

class BatchPlan {
public:
    // precision 0: FP64 factor and solve; 1: FP32 factor with FP64 refinement
    BatchPlan(const std::vector<BatchHeadInput>& in, int precision);
    ~BatchPlan();
    BatchPlan(const BatchPlan&) = delete;
    BatchPlan& operator=(const BatchPlan&) = delete;

    // One pass, ordered after the work already on `caller` (which then waits
    // for it).  The first pass, and every pass with use_graph false, runs
    // eagerly and refines adaptively; that fixes the refinement count.  With
    // use_graph, later passes replay one captured graph of the whole pass
    // (captured on first use).  Returns whether the graph ran.
    bool run(hipStream_t caller, bool use_graph);
    // Times the trailing-update candidates and keeps the fastest; drops the
    // captured graph.  The matrices are overwritten (run again afterwards).
    std::vector<BatchTuneResult> tune(hipStream_t caller);

    int n_heads() const { return (int)heads_.size(); }
    const double* gain(int h, int64_t& ne, int64_t& nd) const {
        const Head& H = heads_.at(h);
        ne = H.ne;
        nd = H.nd;
        return H.G;
    }
    void stats(int h, int& factor_info, double& rel_update, hipStream_t caller);
    int refine_steps() const { return refine_steps_; }
    int trailing_tile() const { return trail_split_ ? 128 : trail_wb_ * 32; }
    bool trailing_split16() const { return trail_split_; }
    int64_t trailing_grid_cap() const { return trail_cap_; }
    int64_t graph_nodes() const { return graph_nodes_; }
    int64_t launches() const;   // grouped launches of the linear algebra per pass

private:
    struct Head {
        Geometry* g = nullptr;
        int64_t n = 0, nv = 0, nt = 0, ld = 0, nd = 0, ne = 0, nblk = 0;
        double* H = nullptr;        // FP64 HeadMat, triangles-first, ld x n
        void* A = nullptr;          // matrix factored (FP32 copy, or H itself)
        void* X = nullptr;          // panels (mixed: the four of a block group, ld x FACT_NB apart)
        void* Dinv = nullptr;       // inverses of the diagonal blocks
        void* sgn = nullptr;        // signs
        double* dip = nullptr;
        double* cD = nullptr;
        double* cP = nullptr;
        std::vector<char> used;
        double* D = nullptr;        // sources, ld x nd
        double* B = nullptr;        // E^T, ld x ne
        double* Y = nullptr;        // H^-1 E^T, ld x ne
        void* W = nullptr;          // right-hand sides in factor precision
        double* R = nullptr;        // FP64 residual (mixed)
        void* tmp = nullptr;        // FACT_NB x ne
        void* sp = nullptr;         // split-FP16 operands of the rank-1024 trailing update (mixed)
        double* G = nullptr;        // gain, ne x nd
        int* info = nullptr;
        double rel_update = 0.0;
    };
    struct Op {
        enum Type { GEMM, DIAG, ELEM, SPLIT, SPLIT_UPDATE, SPLIT4, SPLIT4_UPDATE } type = ELEM;
        GemmLaunch gemm{};
        bool trailing = false;
        bool f32 = false;
        size_t off = 0;
        int ng = 0;
        int64_t max_elems = 0;
        int64_t blocks = 0;         // SPLIT(4): 64-row blocks; SPLIT(4)_UPDATE: tiles
    };
    typedef std::vector<Op> Segment;

    void build_schedule();
    void factor_ops(Segment& seg);
    void solve_ops(Segment& seg);
    void add_gemm(Segment& seg, const std::vector<GemmJob>& jobs, bool f32, bool ta, bool tb, bool lower,
                  bool trailing);
    void add_elem(Segment& seg, const std::vector<ElemDesc>& ds);
    void add_diag(Segment& seg, const std::vector<DiagDesc>& ds, bool f32);
    void add_split(Segment& seg, std::vector<SplitDesc>& ds);
    void add_split4(Segment& seg, std::vector<Split4Desc>& ds);
    std::vector<Split4Desc> split4_group(int64_t q) const;
    void upload_tables();
    void launch(const Segment& seg, hipStream_t st);
    void enqueue_assembly(hipStream_t st);
    void enqueue(hipStream_t st, int refine_steps);
    void run_eager(hipStream_t st);
    void capture();

    int precision_;
    int device_ = 0;
    std::vector<Head> heads_;
    int64_t max_nblk_ = 0;
    int* info_ = nullptr;
    std::vector<void*> owned_;
    ScratchArena arena_;
    // assembly: heads spread over ASM_STREAMS streams forked from the pass's
    // stream (each with its own scratch arena), joined before the algebra
    static constexpr int ASM_STREAMS = 4;
    ScratchArena asm_arena_[ASM_STREAMS];
    hipStream_t asm_stream_[ASM_STREAMS] = {};
    hipEvent_t asm_fork_ = nullptr, asm_join_[ASM_STREAMS] = {};
    std::vector<int> asm_of_;     // stream of each head
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_in_ = nullptr, ev_out_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    int64_t graph_nodes_ = 0;
    bool calibrated_ = false;
    int refine_steps_ = 0;
    int max_refine_ = 10;
    double refine_tol_ = 1e-8;    // estimated error left (refine_state), as the single-head gain pipelines
    bool refine_ok_ = true;       // the last eager pass reached refine_tol_
    int trail_wb_ = 4;
    int64_t trail_cap_ = 0;
    bool trail_split_ = true;     // mixed: split-FP16 trailing update (else FP32 grouped GEMM)

    std::vector<GemmDesc> gemm_descs_;
    std::vector<DiagDesc> diag_descs_;
    std::vector<ElemDesc> elem_descs_;
    std::vector<SplitDesc> split_descs_;
    std::vector<Split4Desc> split4_descs_;
    GemmDesc* d_gemm_ = nullptr;
    DiagDesc* d_diag_ = nullptr;
    ElemDesc* d_elem_ = nullptr;
    SplitDesc* d_split_ = nullptr;
    Split4Desc* d_split4_ = nullptr;
    size_t cap_gemm_ = 0, cap_diag_ = 0, cap_elem_ = 0, cap_split_ = 0, cap_split4_ = 0;
    Segment prep_, factor_, solve_, ir_init_, ir_pre_, ir_post_, ir_resid_, gain_;
};

}  // namespace omh
