// Batched EEG gains: many heads (each a geometry with its own dipoles and
// electrodes, of uneven sizes) computed in one pass whose launches are
// captured into a single HIP graph.
//
// Per head the pass is the adjoint gain path of the single-head pipeline
// (HeadMat and DipSourceMat assembly, assembleHeadMat.cpp:74-135 and
// assembleSourceMat.cpp:66-104; HeadMat factor; Y = H^-1 E^T; G = Y^T D,
// gain.h:78-89).  The dense linear algebra of all heads advances in lock
// step: step s of the blocked factorisation is one grouped diagonal-block
// launch (one workgroup per head), one grouped panel GEMM and one grouped
// trailing update over the tiles of every head that still has a block s;
// the triangular solves and refinement residuals are grouped the same way.
// Small heads therefore share launches with large ones instead of running
// their own latency-bound sequences.  All descriptors (pointers, sizes, tile
// ranges) are built on the host once, so the pass is a fixed sequence of
// kernel launches: capture-safe, with plan-owned buffers and scratch.
//
// Tuning: the trailing update dominates; its tile size (64 or 128) and
// grid cap (one workgroup per tile, or a persistent grid of fixed size whose
// workgroups walk the tile list) are chosen by timing the candidates on the
// plan's own shapes (tune()).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "assembly.hpp"
#include "batch.hpp"
#include "common.hpp"
#include "device_geometry.hpp"
#include "linalg.hpp"
#include "sources.hpp"

namespace omh {

namespace {

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

template <typename T>
T* dalloc(size_t count, std::vector<void*>& owned) {
    void* p = nullptr;
    OMH_HIP(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
    owned.push_back(p);
    return (T*)p;
}

}  // namespace

BatchPlan::BatchPlan(const std::vector<BatchHeadInput>& in, int precision) : precision_(precision) {
    if (in.empty()) throw std::invalid_argument("a batch needs at least one head");
    if (precision != 0 && precision != 1) throw std::invalid_argument("precision must be 0 (FP64) or 1 (mixed)");
    OMH_HIP(hipGetDevice(&device_));
    OMH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    OMH_HIP(hipEventCreateWithFlags(&ev_in_, hipEventDisableTiming));
    OMH_HIP(hipEventCreateWithFlags(&ev_out_, hipEventDisableTiming));
    const bool f32 = precision_ == 1;
    const size_t te = f32 ? sizeof(float) : sizeof(double);
    info_ = dalloc<int>(in.size(), owned_);
    for (size_t h = 0; h < in.size(); ++h) {
        const BatchHeadInput& hi = in[h];
        if (!hi.g) throw std::invalid_argument("null geometry");
        Geometry& g = *hi.g;
        Head H;
        H.g = hi.g;
        H.n = g.size();
        H.nv = g.n_vert_unknowns;
        H.nt = g.n_tri_unknowns;
        H.nd = (int64_t)hi.dipoles.size() / 6;
        H.ne = (int64_t)hi.electrodes.size() / 3;
        if (H.n <= 0) throw std::invalid_argument("empty geometry");
        if (H.nd <= 0 || H.ne <= 0) throw std::invalid_argument("each head needs dipoles and electrodes");
        H.ld = round_up(H.n, 32);
        H.nblk = cdiv(H.n, FACT_NB);
        const Layout L{H.ld, H.nv, H.nt, 1};
        device_geometry(g);
        // dipoles, their domains and the source coefficients
        H.dip = dalloc<double>(H.nd * 6, owned_);
        OMH_HIP(hipMemcpy(H.dip, hi.dipoles.data(), sizeof(double) * 6 * H.nd, hipMemcpyHostToDevice));
        const std::vector<int> dom = locate_domains(g, H.dip, 6, H.nd, stream_);
        std::vector<double> cD, cP;
        dsm_coefficients(g, H.nd, dom, cD, cP, H.used);
        H.cD = dalloc<double>(cD.size(), owned_);
        H.cP = dalloc<double>(cP.size(), owned_);
        OMH_HIP(hipMemcpy(H.cD, cD.data(), sizeof(double) * cD.size(), hipMemcpyHostToDevice));
        OMH_HIP(hipMemcpy(H.cP, cP.data(), sizeof(double) * cP.size(), hipMemcpyHostToDevice));
        // electrodes -> E^T (n x ne, triangles-first rows)
        double* d_el = dalloc<double>(H.ne * 3, owned_);
        OMH_HIP(hipMemcpy(d_el, hi.electrodes.data(), sizeof(double) * 3 * H.ne, hipMemcpyHostToDevice));
        std::vector<int64_t> cols(3 * H.ne);
        std::vector<double> vals(3 * H.ne);
        head2eeg(g, d_el, H.ne, cols.data(), vals.data(), stream_);
        std::vector<double> Bh((size_t)H.ld * H.ne, 0.0);
        for (int64_t e = 0; e < H.ne; ++e)
            for (int k = 0; k < 3; ++k) Bh[dev_index(L, cols[3 * e + k]) + e * H.ld] += vals[3 * e + k];
        H.B = dalloc<double>(Bh.size(), owned_);
        OMH_HIP(hipMemcpy(H.B, Bh.data(), sizeof(double) * Bh.size(), hipMemcpyHostToDevice));
        // matrices and workspaces
        H.H = dalloc<double>((size_t)H.ld * H.n, owned_);
        H.A = f32 ? (void*)dalloc<float>((size_t)H.ld * H.n, owned_) : (void*)H.H;
        H.X = dalloc<char>(te * (size_t)H.ld * FACT_NB * (f32 ? S4_PANELS : 1), owned_);
        H.Dinv = dalloc<char>(te * (size_t)H.nblk * FACT_NB * FACT_NB, owned_);
        H.sgn = dalloc<char>(te * (size_t)H.ld, owned_);
        H.D = dalloc<double>((size_t)H.ld * H.nd, owned_);
        H.Y = dalloc<double>((size_t)H.ld * H.ne, owned_);
        H.W = f32 ? (void*)dalloc<float>((size_t)H.ld * H.ne, owned_) : (void*)H.Y;
        H.R = f32 ? dalloc<double>((size_t)H.ld * H.ne, owned_) : nullptr;
        H.tmp = dalloc<char>(te * (size_t)FACT_NB * H.ne, owned_);
        if (f32 && H.n > FACT_NB) H.sp = dalloc<char>(std::max(split_operand_bytes(H.n), split4_operand_bytes(H.n)), owned_);
        H.G = dalloc<double>((size_t)H.ne * H.nd, owned_);
        H.info = info_ + h;
        max_nblk_ = std::max(max_nblk_, H.nblk);
        heads_.push_back(std::move(H));
    }
    OMH_HIP(hipStreamSynchronize(stream_));
    // assembly streams: the largest heads first, each to the stream with the
    // least assembly so far (pairs of unknowns as the measure)
    OMH_HIP(hipEventCreateWithFlags(&asm_fork_, hipEventDisableTiming));
    for (int s = 0; s < ASM_STREAMS; ++s) {
        OMH_HIP(hipStreamCreateWithFlags(&asm_stream_[s], hipStreamNonBlocking));
        OMH_HIP(hipEventCreateWithFlags(&asm_join_[s], hipEventDisableTiming));
    }
    std::vector<int> order(heads_.size());
    for (size_t h = 0; h < order.size(); ++h) order[h] = (int)h;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return heads_[a].n > heads_[b].n; });
    asm_of_.assign(heads_.size(), 0);
    double load[ASM_STREAMS] = {};
    for (int h : order) {
        const int s = (int)(std::min_element(load, load + ASM_STREAMS) - load);
        asm_of_[h] = s;
        load[s] += (double)heads_[h].n * (double)heads_[h].n;
    }
    build_schedule();
}

BatchPlan::~BatchPlan() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    arena_.release();
    for (int s = 0; s < ASM_STREAMS; ++s) {
        if (asm_stream_[s]) (void)hipStreamSynchronize(asm_stream_[s]);
        asm_arena_[s].release();
        if (asm_join_[s]) (void)hipEventDestroy(asm_join_[s]);
        if (asm_stream_[s]) (void)hipStreamDestroy(asm_stream_[s]);
    }
    if (asm_fork_) (void)hipEventDestroy(asm_fork_);
    for (void* p : owned_) (void)hipFree(p);
    if (d_gemm_) (void)hipFree(d_gemm_);
    if (d_diag_) (void)hipFree(d_diag_);
    if (d_elem_) (void)hipFree(d_elem_);
    if (d_split_) (void)hipFree(d_split_);
    if (d_split4_) (void)hipFree(d_split4_);
    if (ev_in_) (void)hipEventDestroy(ev_in_);
    if (ev_out_) (void)hipEventDestroy(ev_out_);
    if (stream_) (void)hipStreamDestroy(stream_);
    (void)hipSetDevice(cur);
}

// ---------------------------------------------------------------------------
// schedule
// ---------------------------------------------------------------------------
void BatchPlan::add_gemm(Segment& seg, const std::vector<GemmJob>& jobs, bool f32, bool ta, bool tb, bool lower,
                         bool trailing) {
    const int wb = (trailing && f32) ? trail_wb_ : 2;
    std::vector<GemmLaunch> ls;
    build_grouped_gemm(jobs, f32, ta, tb, lower, wb, gemm_descs_, ls);
    for (const GemmLaunch& l : ls) {
        Op op;
        op.type = Op::GEMM;
        op.gemm = l;
        op.trailing = trailing;
        seg.push_back(op);
    }
}

void BatchPlan::add_elem(Segment& seg, const std::vector<ElemDesc>& ds) {
    if (ds.empty()) return;
    Op op;
    op.type = Op::ELEM;
    op.off = elem_descs_.size();
    op.ng = (int)ds.size();
    for (const ElemDesc& d : ds) op.max_elems = std::max(op.max_elems, d.m * d.n);
    elem_descs_.insert(elem_descs_.end(), ds.begin(), ds.end());
    seg.push_back(op);
}

// A split of every listed matrix, then their split updates (descriptors
// get their block and tile offsets here).
void BatchPlan::add_split(Segment& seg, std::vector<SplitDesc>& ds) {
    if (ds.empty()) return;
    int64_t blocks = 0, tiles = 0;
    for (SplitDesc& d : ds) {
        d.blk0 = blocks;
        d.tile0 = tiles;
        blocks += d.mp / 64;
        const int64_t t = cdiv(d.m, 128);
        tiles += t * (t + 1) / 2;
    }
    Op op;
    op.off = split_descs_.size();
    op.ng = (int)ds.size();
    split_descs_.insert(split_descs_.end(), ds.begin(), ds.end());
    op.type = Op::SPLIT;
    op.blocks = blocks;
    seg.push_back(op);
    op.type = Op::SPLIT_UPDATE;
    op.blocks = tiles;
    seg.push_back(op);
}

// The rank-1024 split of every listed matrix, then their updates.
void BatchPlan::add_split4(Segment& seg, std::vector<Split4Desc>& ds) {
    if (ds.empty()) return;
    int64_t blocks = 0, tiles = 0;
    for (Split4Desc& d : ds) {
        d.blk0 = blocks;
        d.tile0 = tiles;
        blocks += d.mp / 64;
        const int64_t t = cdiv(d.m, 128);
        tiles += t * (t + 1) / 2;
    }
    Op op;
    op.off = split4_descs_.size();
    op.ng = (int)ds.size();
    split4_descs_.insert(split4_descs_.end(), ds.begin(), ds.end());
    op.type = Op::SPLIT4;
    op.blocks = blocks;
    seg.push_back(op);
    op.type = Op::SPLIT4_UPDATE;
    op.blocks = tiles;
    seg.push_back(op);
}

// Descriptors of block group q's rank-1024 update: the heads whose group
// is whole and has rows below it (the four panels' rows from the group's
// end).
std::vector<Split4Desc> BatchPlan::split4_group(int64_t q) const {
    std::vector<Split4Desc> sd;
    for (const Head& H : heads_) {
        const int64_t k0 = q * S4_PANELS * FACT_NB, gend = k0 + S4_PANELS * FACT_NB;
        if (gend >= H.n) continue;
        Split4Desc d{};
        for (int p = 0; p < S4_PANELS; ++p) {
            const int64_t e = k0 + (p + 1) * FACT_NB;   // panel p holds the rows from e
            d.X[p] = (const float*)((const char*)H.X + sizeof(float) * (size_t)H.ld * FACT_NB * p);
            d.ldx[p] = round_up(H.n - e, 4);
            d.off[p] = gend - e;
        }
        d.sgn = (const float*)H.sgn + k0;
        d.m = H.n - gend;
        d.mp = round_up(d.m, 128);
        d.Ah = H.sp;
        d.C = (float*)H.A + (size_t)gend + (size_t)gend * H.ld;
        d.ldc = H.ld;
        sd.push_back(d);
    }
    return sd;
}

void BatchPlan::add_diag(Segment& seg, const std::vector<DiagDesc>& ds, bool f32) {
    if (ds.empty()) return;
    Op op;
    op.type = Op::DIAG;
    op.f32 = f32;
    op.off = diag_descs_.size();
    op.ng = (int)ds.size();
    diag_descs_.insert(diag_descs_.end(), ds.begin(), ds.end());
    seg.push_back(op);
}

void BatchPlan::factor_ops(Segment& seg) {
    const bool f32 = precision_ == 1;
    const size_t te = f32 ? sizeof(float) : sizeof(double);
    auto at = [&](const Head& H, void* base, int64_t i, int64_t j) {
        return (void*)((char*)base + te * (size_t)(i + j * H.ld));
    };
    std::vector<ElemDesc> zero;
    for (const Head& H : heads_) zero.push_back(ElemDesc{ELEM_ZERO_INT, nullptr, 1, H.info, 1, 1, 1, nullptr});
    add_elem(seg, zero);
    if (f32 && trail_split_) {
        // Super-steps of S4_PANELS blocks (the single-head factor's grouping):
        // inside a group each block's diagonal factor, panel (its own slot of
        // X) and the update of the group's later columns; then one rank-1024
        // split update of everything below the group, a quarter of the
        // trailing matrix's reads and writes of rank-256 steps.
        for (int64_t q = 0; q * S4_PANELS < max_nblk_; ++q) {
            for (int b = 0; b < S4_PANELS; ++b) {
                const int64_t s = q * S4_PANELS + b, k = s * FACT_NB;
                std::vector<DiagDesc> diag;
                std::vector<GemmJob> panel, ulow, urect;
                std::vector<ElemDesc> scale;
                for (const Head& H : heads_) {
                    if (k >= H.n) continue;
                    const int nb = (int)std::min<int64_t>(FACT_NB, H.n - k);
                    void* Li = (char*)H.Dinv + te * (size_t)s * FACT_NB * FACT_NB;
                    diag.push_back(DiagDesc{at(H, H.A, k, k), H.ld, nb, Li, H.sgn, H.info, k});
                    const int64_t e = k + nb, m = H.n - e;
                    if (m <= 0) continue;
                    const int64_t ldx = round_up(m, 4);
                    void* Xb = (char*)H.X + te * (size_t)H.ld * FACT_NB * b;
                    void* A21 = at(H, H.A, e, k);
                    panel.push_back(GemmJob{m, nb, nb, 1.0, A21, H.ld, Li, FACT_NB, 0.0, Xb, ldx});
                    scale.push_back(ElemDesc{ELEM_SCALE_COLS_F32, Xb, ldx, A21, H.ld, m, nb, (char*)H.sgn + te * k});
                    // the group's later columns [e, e + w): their lower
                    // square and the rows below it
                    const int64_t w = std::min<int64_t>((q + 1) * S4_PANELS * FACT_NB, H.n) - e;
                    if (w > 0) {
                        ulow.push_back(GemmJob{w, w, nb, -1.0, Xb, ldx, A21, H.ld, 1.0, at(H, H.A, e, e), H.ld});
                        if (m > w)
                            urect.push_back(GemmJob{m - w, w, nb, -1.0, (char*)Xb + te * w, ldx, A21, H.ld, 1.0,
                                                    at(H, H.A, e + w, e), H.ld});
                    }
                }
                add_diag(seg, diag, f32);
                add_gemm(seg, panel, f32, false, true, false, false);
                add_elem(seg, scale);
                add_gemm(seg, ulow, f32, false, true, true, false);
                add_gemm(seg, urect, f32, false, true, false, false);
            }
            std::vector<Split4Desc> sd = split4_group(q);
            add_split4(seg, sd);
        }
        return;
    }
    for (int64_t s = 0; s < max_nblk_; ++s) {
        const int64_t k = s * FACT_NB;
        std::vector<DiagDesc> diag;
        std::vector<GemmJob> panel, trail;
        std::vector<ElemDesc> scale;
        std::vector<SplitDesc> split;
        for (const Head& H : heads_) {
            if (k >= H.n) continue;
            const int nb = (int)std::min<int64_t>(FACT_NB, H.n - k);
            void* Li = (char*)H.Dinv + te * (size_t)s * FACT_NB * FACT_NB;
            diag.push_back(DiagDesc{at(H, H.A, k, k), H.ld, nb, Li, H.sgn, H.info, k});
            const int64_t e = k + nb, m = H.n - e;
            if (m <= 0) continue;
            const int64_t ldx = round_up(m, 4);
            void* A21 = at(H, H.A, e, k);
            panel.push_back(GemmJob{m, nb, nb, 1.0, A21, H.ld, Li, FACT_NB, 0.0, H.X, ldx});
            scale.push_back(ElemDesc{f32 ? ELEM_SCALE_COLS_F32 : ELEM_SCALE_COLS_F64, H.X, ldx, A21, H.ld, m, nb,
                                     (char*)H.sgn + te * k});
            if (f32 && trail_split_)
                split.push_back(SplitDesc{(const float*)H.X, ldx, (const float*)H.sgn + k, m, round_up(m, 128), nb,
                                          H.sp, (float*)at(H, H.A, e, e), H.ld, 0, 0});
            else
                trail.push_back(GemmJob{m, m, nb, -1.0, H.X, ldx, A21, H.ld, 1.0, at(H, H.A, e, e), H.ld});
        }
        add_diag(seg, diag, f32);
        add_gemm(seg, panel, f32, false, true, false, false);
        add_elem(seg, scale);
        add_gemm(seg, trail, f32, false, true, true, true);
        add_split(seg, split);
    }
}

void BatchPlan::solve_ops(Segment& seg) {
    const bool f32 = precision_ == 1;
    const size_t te = f32 ? sizeof(float) : sizeof(double);
    const int copy = f32 ? ELEM_COPY_F32 : ELEM_COPY_F64;
    auto at = [&](const void* base, int64_t ld, int64_t i, int64_t j) {
        return (void*)((char*)base + te * (size_t)(i + j * ld));
    };
    for (int64_t s = 0; s < max_nblk_; ++s) {   // forward: W <- L^-1 W
        const int64_t k = s * FACT_NB;
        std::vector<GemmJob> g1, g2;
        std::vector<ElemDesc> cp;
        for (const Head& H : heads_) {
            if (k >= H.n) continue;
            const int nb = (int)std::min<int64_t>(FACT_NB, H.n - k);
            const void* Li = (char*)H.Dinv + te * (size_t)s * FACT_NB * FACT_NB;
            void* Wk = at(H.W, H.ld, k, 0);
            g1.push_back(GemmJob{nb, H.ne, nb, 1.0, Li, FACT_NB, Wk, H.ld, 0.0, H.tmp, FACT_NB});
            cp.push_back(ElemDesc{copy, H.tmp, FACT_NB, Wk, H.ld, nb, H.ne, nullptr});
            const int64_t e = k + nb;
            if (e < H.n)
                g2.push_back(GemmJob{H.n - e, H.ne, nb, -1.0, at(H.A, H.ld, e, k), H.ld, Wk, H.ld, 1.0,
                                     at(H.W, H.ld, e, 0), H.ld});
        }
        add_gemm(seg, g1, f32, false, false, false, false);
        add_elem(seg, cp);
        add_gemm(seg, g2, f32, false, false, false, false);
    }
    std::vector<ElemDesc> sg;   // W <- S W
    for (const Head& H : heads_)
        sg.push_back(ElemDesc{f32 ? ELEM_SCALE_ROWS_F32 : ELEM_SCALE_ROWS_F64, nullptr, H.ld, H.W, H.ld, H.n, H.ne,
                              H.sgn});
    add_elem(seg, sg);
    for (int64_t s = max_nblk_ - 1; s >= 0; --s) {   // backward: W <- L^-T W
        const int64_t k = s * FACT_NB;
        std::vector<GemmJob> g3, g4;
        std::vector<ElemDesc> cp;
        for (const Head& H : heads_) {
            if (k >= H.n) continue;
            const int nb = (int)std::min<int64_t>(FACT_NB, H.n - k);
            const void* Li = (char*)H.Dinv + te * (size_t)s * FACT_NB * FACT_NB;
            void* Wk = at(H.W, H.ld, k, 0);
            g3.push_back(GemmJob{nb, H.ne, nb, 1.0, Li, FACT_NB, Wk, H.ld, 0.0, H.tmp, FACT_NB});
            cp.push_back(ElemDesc{copy, H.tmp, FACT_NB, Wk, H.ld, nb, H.ne, nullptr});
            if (k > 0)
                g4.push_back(GemmJob{k, H.ne, nb, -1.0, at(H.A, H.ld, k, 0), H.ld, Wk, H.ld, 1.0, H.W, H.ld});
        }
        add_gemm(seg, g3, f32, true, false, false, false);
        add_elem(seg, cp);
        add_gemm(seg, g4, f32, true, false, false, false);
    }
}

void BatchPlan::build_schedule() {
    gemm_descs_.clear();
    diag_descs_.clear();
    elem_descs_.clear();
    split_descs_.clear();
    split4_descs_.clear();
    prep_.clear();
    factor_.clear();
    solve_.clear();
    ir_init_.clear();
    ir_pre_.clear();
    ir_post_.clear();
    ir_resid_.clear();
    gain_.clear();
    const bool f32 = precision_ == 1;
    if (f32) {   // FP32 copy of the lower triangle of H
        std::vector<ElemDesc> cv;
        for (const Head& H : heads_)
            cv.push_back(ElemDesc{ELEM_F64_TO_F32_LOWER, H.H, H.ld, H.A, H.ld, H.n, H.n, nullptr});
        add_elem(prep_, cv);
    }
    factor_ops(factor_);
    solve_ops(solve_);
    std::vector<ElemDesc> init, pre, post, resid_copy;
    std::vector<GemmJob> resid;
    for (const Head& H : heads_) {
        if (f32) {
            init.push_back(ElemDesc{ELEM_ZERO_F64, nullptr, H.ld, H.Y, H.ld, H.n, H.ne, nullptr});
            init.push_back(ElemDesc{ELEM_COPY_F64, H.B, H.ld, H.R, H.ld, H.n, H.ne, nullptr});
            pre.push_back(ElemDesc{ELEM_F64_TO_F32, H.R, H.ld, H.W, H.ld, H.n, H.ne, nullptr});
            post.push_back(ElemDesc{ELEM_ADD_F32_TO_F64, H.W, H.ld, H.Y, H.ld, H.n, H.ne, nullptr});
            resid_copy.push_back(ElemDesc{ELEM_COPY_F64, H.B, H.ld, H.R, H.ld, H.n, H.ne, nullptr});
            resid.push_back(GemmJob{H.n, H.ne, H.n, -1.0, H.H, H.ld, H.Y, H.ld, 1.0, H.R, H.ld});
        } else {
            init.push_back(ElemDesc{ELEM_COPY_F64, H.B, H.ld, H.Y, H.ld, H.n, H.ne, nullptr});
        }
    }
    add_elem(ir_init_, init);
    add_elem(ir_pre_, pre);
    add_elem(ir_post_, post);
    add_elem(ir_resid_, resid_copy);
    if (!resid.empty()) add_gemm(ir_resid_, resid, false, false, false, false, false);
    std::vector<GemmJob> gain;
    for (const Head& H : heads_) gain.push_back(GemmJob{H.ne, H.nd, H.n, 1.0, H.Y, H.ld, H.D, H.ld, 0.0, H.G, H.ne});
    add_gemm(gain_, gain, false, true, false, false, false);
    upload_tables();
    // a new schedule invalidates a captured graph
    if (exec_) { (void)hipGraphExecDestroy(exec_); exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
}

template <typename D>
static void upload_table(const std::vector<D>& v, D*& d, size_t& cap) {
    if (v.size() > cap) {
        if (d) OMH_HIP(hipFree(d));
        d = nullptr;
        OMH_HIP(hipMalloc((void**)&d, std::max<size_t>(1, v.size()) * sizeof(D)));
        cap = v.size();
    }
    if (!v.empty()) OMH_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(D), hipMemcpyHostToDevice));
}

void BatchPlan::upload_tables() {
    OMH_HIP(hipStreamSynchronize(stream_));
    upload_table(gemm_descs_, d_gemm_, cap_gemm_);
    upload_table(diag_descs_, d_diag_, cap_diag_);
    upload_table(elem_descs_, d_elem_, cap_elem_);
    upload_table(split_descs_, d_split_, cap_split_);
    upload_table(split4_descs_, d_split4_, cap_split4_);
}

void BatchPlan::launch(const Segment& seg, hipStream_t st) {
    for (const Op& op : seg) {
        switch (op.type) {
            case Op::GEMM: launch_grouped_gemm(op.gemm, d_gemm_, op.trailing ? trail_cap_ : 0, st); break;
            case Op::DIAG: launch_grouped_diag(op.f32, d_diag_ + op.off, op.ng, st); break;
            case Op::ELEM: launch_grouped_elem(d_elem_ + op.off, op.ng, op.max_elems, st); break;
            case Op::SPLIT: launch_grouped_split(d_split_ + op.off, op.ng, op.blocks, st); break;
            case Op::SPLIT_UPDATE: launch_grouped_split_update(d_split_ + op.off, op.ng, op.blocks, st); break;
            case Op::SPLIT4: launch_grouped_split4(d_split4_ + op.off, op.ng, op.blocks, st); break;
            case Op::SPLIT4_UPDATE: launch_grouped_split4_update(d_split4_ + op.off, op.ng, op.blocks, st); break;
        }
    }
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
// Each head's HeadMat and source matrix on its assembly stream (forked from
// st and joined back: under capture, graph dependencies), so that the small
// heads' launches run beside each other instead of one after another.
void BatchPlan::enqueue_assembly(hipStream_t st) {
    OMH_HIP(hipEventRecord(asm_fork_, st));
    for (int s = 0; s < ASM_STREAMS; ++s) OMH_HIP(hipStreamWaitEvent(asm_stream_[s], asm_fork_, 0));
    for (size_t h = 0; h < heads_.size(); ++h) {
        Head& H = heads_[h];
        const int s = asm_of_[h];
        ArenaScope scope(&asm_arena_[s]);
        const Layout L{H.ld, H.nv, H.nt, 1};
        assemble_headmat(*H.g, 3, H.H, L, asm_stream_[s], precision_ == 1);   // far-field Gauss for the mixed path
        dsm_launch(*H.g, H.dip, H.nd, H.cD, H.cP, H.used, 3, 10, 0.001, H.D, H.ld, L, asm_stream_[s]);
    }
    for (int s = 0; s < ASM_STREAMS; ++s) {
        OMH_HIP(hipEventRecord(asm_join_[s], asm_stream_[s]));
        OMH_HIP(hipStreamWaitEvent(st, asm_join_[s], 0));
    }
}

void BatchPlan::enqueue(hipStream_t st, int refine_steps) {
    enqueue_assembly(st);
    launch(prep_, st);
    launch(factor_, st);
    launch(ir_init_, st);
    if (precision_ == 1) {
        for (int it = 0; it < refine_steps; ++it) {
            launch(ir_pre_, st);
            launch(solve_, st);
            launch(ir_post_, st);
            if (it + 1 < refine_steps) launch(ir_resid_, st);
        }
    } else {
        launch(solve_, st);
    }
    launch(gain_, st);
}

// Eager pass: refinement stops when every head's relative update is below
// tol (or stagnates); the count is kept for graph capture.
void BatchPlan::run_eager(hipStream_t st) {
    enqueue_assembly(st);
    launch(prep_, st);
    launch(factor_, st);
    launch(ir_init_, st);
    if (precision_ == 1) {
        double prev = 2.0;
        int it = 0;
        RefineState state = REFINE_CONTINUE;
        for (; it < max_refine_; ++it) {
            launch(ir_pre_, st);
            launch(solve_, st);
            launch(ir_post_, st);
            double worst = 0.0;
            for (Head& H : heads_) {
                if (it == 0) {   // Y was zero: the first update is Y itself
                    H.rel_update = 1.0;
                } else {
                    const double dn = frob_norm_f32((const float*)H.W, H.ld, H.n, H.ne, st);
                    const double xn = frob_norm_f64(H.Y, H.ld, H.n, H.ne, st);
                    H.rel_update = xn > 0 ? dn / xn : 0.0;
                }
                worst = std::max(worst, H.rel_update);
            }
            state = refine_state(it, worst, prev, refine_tol_);
            if (state != REFINE_CONTINUE) { ++it; break; }
            prev = worst;
            launch(ir_resid_, st);
        }
        refine_steps_ = std::max(1, std::min(it, max_refine_));
        refine_ok_ = state == REFINE_CONVERGED;
    } else {
        launch(solve_, st);
        refine_steps_ = 0;
    }
    launch(gain_, st);
    calibrated_ = true;
}

void BatchPlan::capture() {
    if (!calibrated_) throw std::runtime_error("run the batch once without the graph before capturing it");
    arena_.frozen = true;
    for (int s = 0; s < ASM_STREAMS; ++s) asm_arena_[s].frozen = true;
    OMH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue(stream_, refine_steps_);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OMH_HIP(hipStreamEndCapture(stream_, &graph_));
    size_t nodes = 0;
    OMH_HIP(hipGraphGetNodes(graph_, nullptr, &nodes));
    graph_nodes_ = (int64_t)nodes;
    OMH_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

bool BatchPlan::run(hipStream_t caller, bool use_graph) {
    OMH_HIP(hipEventRecord(ev_in_, caller));
    OMH_HIP(hipStreamWaitEvent(stream_, ev_in_, 0));
    const bool graph = use_graph && calibrated_;
    if (graph) {
        if (!exec_) capture();
        OMH_HIP(hipGraphLaunch(exec_, stream_));
    } else {
        ArenaScope scope(&arena_);
        run_eager(stream_);
    }
    OMH_HIP(hipEventRecord(ev_out_, stream_));
    OMH_HIP(hipStreamWaitEvent(caller, ev_out_, 0));
    if (!refine_ok_)
        throw NotConverged("iterative refinement of the batch did not reach its tolerance: the FP32 factor of a "
                           "head is too coarse for its matrix (use FP64)");
    return graph;
}

// ---------------------------------------------------------------------------
// tuning of the trailing update
// ---------------------------------------------------------------------------
std::vector<BatchTuneResult> BatchPlan::tune(hipStream_t caller) {
    OMH_HIP(hipEventRecord(ev_in_, caller));
    OMH_HIP(hipStreamWaitEvent(stream_, ev_in_, 0));
    const bool f32 = precision_ == 1;
    const size_t te = f32 ? sizeof(float) : sizeof(double);
    // zero matrices: the timing runs write into them before the next pass
    // assembles them again
    for (Head& H : heads_) {
        std::vector<ElemDesc> z;
        if (f32) {
            const int64_t words = (int64_t)((size_t)H.ld * H.n / 2);   // two floats per zeroed double
            z.push_back(ElemDesc{ELEM_ZERO_F64, nullptr, words, H.A, words, words, 1, nullptr});
        } else {
            z.push_back(ElemDesc{ELEM_ZERO_F64, nullptr, H.ld, H.A, H.ld, H.ld, H.n, nullptr});
        }
        ElemDesc* dz = nullptr;
        OMH_HIP(hipMalloc((void**)&dz, sizeof(ElemDesc)));
        OMH_HIP(hipMemcpy(dz, z.data(), sizeof(ElemDesc), hipMemcpyHostToDevice));
        launch_grouped_elem(dz, 1, z[0].m * z[0].n, stream_);
        OMH_HIP(hipStreamSynchronize(stream_));
        OMH_HIP(hipFree(dz));
    }
    // sample steps of the factorisation
    std::vector<int64_t> steps{0, max_nblk_ / 4, max_nblk_ / 2};
    steps.erase(std::unique(steps.begin(), steps.end()), steps.end());
    struct Cand { int wb; int64_t cap; };
    std::vector<Cand> cands;
    for (int wb : (f32 ? std::vector<int>{2, 4} : std::vector<int>{2}))
        for (int64_t cap : {int64_t(0), int64_t(512), int64_t(1024), int64_t(2048)}) cands.push_back({wb, cap});
    hipEvent_t a, b;
    OMH_HIP(hipEventCreate(&a));
    OMH_HIP(hipEventCreate(&b));
    std::vector<BatchTuneResult> res;
    double best = 1e300;
    for (const Cand& c : cands) {
        std::vector<GemmDesc> descs;
        std::vector<GemmLaunch> ls;
        for (int64_t s : steps) {
            const int64_t k = s * FACT_NB;
            std::vector<GemmJob> trail;
            for (const Head& H : heads_) {
                const int64_t e = k + FACT_NB, m = H.n - e;
                if (m <= 0) continue;
                auto at = [&](int64_t i, int64_t j) { return (void*)((char*)H.A + te * (size_t)(i + j * H.ld)); };
                trail.push_back(GemmJob{m, m, FACT_NB, -1.0, H.X, round_up(m, 4), at(e, k), H.ld, 1.0, at(e, e),
                                        H.ld});
            }
            build_grouped_gemm(trail, f32, false, true, true, c.wb, descs, ls);
        }
        if (ls.empty()) continue;
        GemmDesc* dd = nullptr;
        OMH_HIP(hipMalloc((void**)&dd, descs.size() * sizeof(GemmDesc)));
        OMH_HIP(hipMemcpy(dd, descs.data(), descs.size() * sizeof(GemmDesc), hipMemcpyHostToDevice));
        for (const GemmLaunch& l : ls) launch_grouped_gemm(l, dd, c.cap, stream_);   // warm-up
        OMH_HIP(hipEventRecord(a, stream_));
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
            for (const GemmLaunch& l : ls) launch_grouped_gemm(l, dd, c.cap, stream_);
        OMH_HIP(hipEventRecord(b, stream_));
        OMH_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        OMH_HIP(hipEventElapsedTime(&ms, a, b));
        OMH_HIP(hipFree(dd));
        const double t = ms * 1e-3 / reps;
        res.push_back(BatchTuneResult{c.wb * 32, c.cap, false, t});
        if (t < best) {
            best = t;
            trail_wb_ = c.wb;
            trail_cap_ = c.cap;
            trail_split_ = false;
        }
    }
    if (f32) {
        // the split-FP16 update in rank-1024 super-steps: the split and
        // update of the block groups that start at the sampled steps, per
        // FACT_NB columns
        std::vector<std::vector<Split4Desc>> per_group;
        std::vector<int64_t> groups;
        for (int64_t s : steps) groups.push_back(s / S4_PANELS);
        groups.erase(std::unique(groups.begin(), groups.end()), groups.end());
        for (int64_t q : groups) {
            std::vector<Split4Desc> sd = split4_group(q);
            int64_t blocks = 0, tiles = 0;
            for (Split4Desc& d : sd) {
                d.blk0 = blocks;
                d.tile0 = tiles;
                blocks += d.mp / 64;
                const int64_t t = cdiv(d.m, 128);
                tiles += t * (t + 1) / 2;
            }
            if (!sd.empty()) per_group.push_back(sd);
        }
        if (!per_group.empty()) {
            std::vector<Split4Desc*> dd(per_group.size());
            std::vector<int64_t> nb(per_group.size()), nt(per_group.size());
            for (size_t i = 0; i < per_group.size(); ++i) {
                const auto& sd = per_group[i];
                OMH_HIP(hipMalloc((void**)&dd[i], sd.size() * sizeof(Split4Desc)));
                OMH_HIP(hipMemcpy(dd[i], sd.data(), sd.size() * sizeof(Split4Desc), hipMemcpyHostToDevice));
                const Split4Desc& last = sd.back();
                const int64_t t = cdiv(last.m, 128);
                nb[i] = last.blk0 + last.mp / 64;
                nt[i] = last.tile0 + t * (t + 1) / 2;
            }
            auto pass = [&]() {
                for (size_t i = 0; i < per_group.size(); ++i) {
                    launch_grouped_split4(dd[i], (int)per_group[i].size(), nb[i], stream_);
                    launch_grouped_split4_update(dd[i], (int)per_group[i].size(), nt[i], stream_);
                }
            };
            pass();   // warm-up
            OMH_HIP(hipEventRecord(a, stream_));
            const int reps = 3;
            for (int r = 0; r < reps; ++r) pass();
            OMH_HIP(hipEventRecord(b, stream_));
            OMH_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            OMH_HIP(hipEventElapsedTime(&ms, a, b));
            for (Split4Desc* p : dd) OMH_HIP(hipFree(p));
            // per FACT_NB columns, comparable with the rank-FACT_NB candidates
            const double t = ms * 1e-3 / reps / S4_PANELS * (double)steps.size() / (double)per_group.size();
            res.push_back(BatchTuneResult{128, 0, true, t});
            if (t < best) {
                best = t;
                trail_split_ = true;
            }
        }
    }
    OMH_HIP(hipEventDestroy(a));
    OMH_HIP(hipEventDestroy(b));
    build_schedule();
    calibrated_ = false;   // the next eager pass re-assembles over the zeroed matrices
    OMH_HIP(hipEventRecord(ev_out_, stream_));
    OMH_HIP(hipStreamWaitEvent(caller, ev_out_, 0));
    return res;
}

// ---------------------------------------------------------------------------
// results
// ---------------------------------------------------------------------------
void BatchPlan::stats(int h, int& factor_info, double& rel_update, hipStream_t caller) {
    const Head& H = heads_.at(h);
    OMH_HIP(hipEventRecord(ev_in_, caller));
    OMH_HIP(hipStreamWaitEvent(stream_, ev_in_, 0));
    int v = 0;
    OMH_HIP(hipMemcpyAsync(&v, H.info, sizeof(int), hipMemcpyDeviceToHost, stream_));
    OMH_HIP(hipStreamSynchronize(stream_));
    factor_info = v;
    if (precision_ == 1) {
        ArenaScope scope(&arena_);
        const double dn = frob_norm_f32((const float*)H.W, H.ld, H.n, H.ne, stream_);
        const double xn = frob_norm_f64(H.Y, H.ld, H.n, H.ne, stream_);
        rel_update = xn > 0 ? dn / xn : 0.0;
    } else {
        rel_update = 0.0;
    }
}

int64_t BatchPlan::launches() const {
    return (int64_t)(prep_.size() + factor_.size() + ir_init_.size() + gain_.size()) +
           (precision_ == 1 ? refine_steps_ * (int64_t)(ir_pre_.size() + solve_.size() + ir_post_.size()) +
                                  std::max(0, refine_steps_ - 1) * (int64_t)ir_resid_.size()
                            : (int64_t)solve_.size());
}

}  // namespace omh
