// HeadMat assembly on the GPU: S, N, D/D* blocks and deflation.
//
// Follows Details::HeadMatrix (assembleHeadMat.cpp:74-115) pair by pair:
// HeadMatrixBlocks::set_blocks (operators.h:517-525) with the block rules of
// DiagonalBlock (operators.h:180-197) / NonDiagonalBlock (:363-384), then
// deflate (assembleHeadMat.cpp:28-60).
//
// Work decomposition (FP64 VALU bound: every pair evaluation is a few
// dozen FP64 operations on per-triangle constants that stay in cache):
//   S  : one thread per triangle pair, 16 analytic OperatorS evaluations
//        (the analytic triangle is the row mesh, the Gauss triangle the
//        column mesh, as the reference).  Diagonal blocks run over the upper
//        triangle of tiles only.
//   D  : two passes.  Pass 1: one thread per (row triangle, column triangle)
//        pair computes the three P1 integrals into a scratch tile.  Pass 2:
//        one thread per (row triangle, column vertex) gathers them over the
//        vertex's triangles in mesh order -- the accumulation order of the
//        reference, without atomics.
//   N  : one thread per vertex pair gathering the ~6x6 incident triangle
//        pairs' S values (operators.h:134-148).
//   Far field (the mixed-precision path): S, D and D* of a mesh pair in one
//   pass over its triangle pairs, far pairs from one set of Gauss-point
//   distances, the other pairs through a queue (k_pairs_far / k_pairs_fix).
// Every entry is written to both halves of the symmetric matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "adaptive.hpp"
#include "assembly.hpp"
#include "bem_device.hpp"
#include "common.hpp"
#include "device_geometry.hpp"

namespace omh {

// ---------------------------------------------------------------------------
// scratch arena and fill
// ---------------------------------------------------------------------------
namespace {
struct Arena {
    void* p[SCR_NSLOTS] = {};
    size_t sz[SCR_NSLOTS] = {};
    hipStream_t last[SCR_NSLOTS] = {};
};
std::mutex g_arena_mu;
std::map<int, Arena> g_arenas;
}  // namespace

thread_local ScratchArena* tl_arena = nullptr;   // ArenaScope override of this thread

// Hands out slot `slot` for work on stream st.  Growth is stream ordered:
// the old buffer is freed and the new one allocated on st, so that the
// caller's other streams do not wait.  Only when the slot was last handed
// out on another stream (whose kernels may still use it, and which may no
// longer exist) does the device finish its work first.  A frozen arena (a
// batch plan's, used on the plan's own stream, possibly under graph
// capture) neither grows nor waits.
static void* take(void*& p, size_t& sz, hipStream_t& last, size_t bytes, hipStream_t st, bool frozen) {
    if (p && last != st && !frozen) OMH_HIP(hipDeviceSynchronize());
    last = st;
    if (sz >= bytes && p) return p;
    if (frozen) throw std::runtime_error("frozen scratch arena is too small (plan sizes changed)");
    if (p) OMH_HIP(hipFreeAsync(p, st));
    p = nullptr;
    sz = 0;
    const size_t want = (std::max<size_t>(bytes, 1) + 4095) & ~size_t(4095);
    OMH_HIP(hipMallocAsync(&p, want, st));
    sz = want;
    return p;
}

// Frees stream-ordered allocations once the device has finished with them.
static void free_slots(void** p, size_t* sz) {
    bool any = false;
    for (int s = 0; s < SCR_NSLOTS; ++s) any = any || p[s];
    if (!any) return;
    (void)hipDeviceSynchronize();
    for (int s = 0; s < SCR_NSLOTS; ++s)
        if (p[s]) {
            (void)hipFreeAsync(p[s], nullptr);
            p[s] = nullptr;
            sz[s] = 0;
        }
    (void)hipStreamSynchronize(nullptr);
}

void ScratchArena::release() { free_slots(p, sz); }

ArenaScope::ArenaScope(ScratchArena* a) : prev_(tl_arena) { tl_arena = a; }
ArenaScope::~ArenaScope() { tl_arena = prev_; }

void* scratch(ScratchSlot slot, size_t bytes, hipStream_t st) {
    if (tl_arena) {
        ScratchArena& a = *tl_arena;
        return take(a.p[slot], a.sz[slot], a.last[slot], bytes, st, a.frozen);
    }
    int dev = 0;
    OMH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_arena_mu);
    Arena& a = g_arenas[dev];
    return take(a.p[slot], a.sz[slot], a.last[slot], bytes, st, false);
}

void release_scratch() {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& kv : g_arenas) {
        (void)hipSetDevice(kv.first);
        free_slots(kv.second.p, kv.second.sz);
    }
    g_arenas.clear();
    (void)hipSetDevice(cur);
}

// Columns over grid.y (strided), two consecutive rows per thread over
// grid.x; 16-byte stores where the pair is aligned.
__global__ void k_fill(double* p, int64_t rows, int64_t cols, int64_t ld, double v) {
    for (int64_t c = blockIdx.y; c < cols; c += gridDim.y) {
        double* col = p + c * ld;
        for (int64_t r = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x); r < rows;
             r += 2 * (int64_t)gridDim.x * blockDim.x) {
            if (r + 1 < rows && (reinterpret_cast<uintptr_t>(col + r) & 15) == 0) {
                *reinterpret_cast<double2*>(col + r) = make_double2(v, v);
            } else {
                col[r] = v;
                if (r + 1 < rows) col[r + 1] = v;
            }
        }
    }
}

void fill_f64(double* p, int64_t rows, int64_t cols, int64_t ld, double v, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    const dim3 grid((unsigned)std::min<int64_t>(cdiv(rows, 512), 64), (unsigned)std::min<int64_t>(cols, 8192));
    k_fill<<<grid, 256, 0, s>>>(p, rows, cols, ld, v);
    OMH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// device geometry tables
// ---------------------------------------------------------------------------
DeviceGeometry::~DeviceGeometry() {
    if (device < 0) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(tri);   // hipFree waits for the kernels that may still read the tables
    (void)hipFree(tri_vid);
    (void)hipFree(vert_unk);
    (void)hipFree(slot_vid);
    (void)hipFree(adj_off);
    (void)hipFree(adj);
    (void)hipFree(gp);
    (void)hipSetDevice(cur);
}

// Far-field Gauss points of every triangle (GP_STRIDE doubles each: the
// order-3 then the order-2 rule, by the multadd sequence of gauss_point).
__global__ void k_gauss_table(const TriRec* __restrict__ tri, int64_t nt, double* __restrict__ gp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const d3 P[3] = {ld3(tri[t].p[0]), ld3(tri[t].p[1]), ld3(tri[t].p[2])};
    double* o = gp + t * GP_STRIDE;
    for (int k = 0; k < 16; ++k) {
        const d3 x = gauss_point(3, k, P);
        o[GP16 + 3 * k] = x.x; o[GP16 + 3 * k + 1] = x.y; o[GP16 + 3 * k + 2] = x.z;
    }
    for (int k = 0; k < 7; ++k) {
        const d3 x = gauss_point(2, k, P);
        o[GP7 + 3 * k] = x.x; o[GP7 + 3 * k + 1] = x.y; o[GP7 + 3 * k + 2] = x.z;
    }
}

template <typename T>
static T* upload(const std::vector<T>& v) {
    T* d = nullptr;
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    OMH_HIP(hipMalloc(&d, bytes));
    if (!v.empty()) OMH_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

DeviceGeometry& device_geometry(Geometry& g) {
    int dev = 0;
    OMH_HIP(hipGetDevice(&dev));
    if (g.dev && g.dev->device == dev) return *g.dev;
    auto dg = std::make_shared<DeviceGeometry>();
    std::vector<TriRec> recs;
    std::vector<int32_t> tvid;
    std::vector<int32_t> slots;
    std::vector<int64_t> adj_off{0};
    std::vector<AdjRec> adj;
    for (const auto& m : g.meshes) {
        MeshDev md;
        md.tri_begin = (int64_t)recs.size();
        md.slot_begin = (int64_t)slots.size();
        md.tri_unk0 = m.tri_first_index;
        md.barrier = m.current_barrier ? 1 : 0;
        for (const auto& t : m.tris) {
            TriRec r{};
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) r.p[k][c] = g.vertices[t[k]][c];
            // Triangle ctor, triangle.h:57-62
            double e1[3], e2[3], on[3];
            for (int c = 0; c < 3; ++c) { e1[c] = r.p[1][c] - r.p[0][c]; e2[c] = r.p[2][c] - r.p[0][c]; }
            on[0] = e1[1] * e2[2] - e1[2] * e2[1];
            on[1] = e1[2] * e2[0] - e1[0] * e2[2];
            on[2] = e1[0] * e2[1] - e1[1] * e2[0];
            const double nr = std::sqrt(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
            r.area = 0.5 * nr;
            r.inv_area = 1.0 / r.area;
            for (int c = 0; c < 3; ++c) r.n[c] = on[c] / nr;
            // OperatorS ctor, analytics.h:22-30
            for (int i = 0; i < 3; ++i) {
                const int nx = (i + 1) % 3;
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = r.p[i][c] - r.p[nx][c];
                const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                for (int c = 0; c < 3; ++c) v[c] /= len;
                r.len[i] = len;
                r.enrm[i][0] = r.n[1] * v[2] - r.n[2] * v[1];
                r.enrm[i][1] = r.n[2] * v[0] - r.n[0] * v[2];
                r.enrm[i][2] = r.n[0] * v[1] - r.n[1] * v[0];
                // P1SingleLayerPotential ctor, analytics.h:81-85: D_i = p_{i+1}-p_i
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = r.p[nx][c] - r.p[i][c];
                const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int c = 0; c < 3; ++c) r.U[i][c] = d[c] / dl;
            }
            r.diam = std::max(r.len[0], std::max(r.len[1], r.len[2]));
            recs.push_back(r);
            for (int k = 0; k < 3; ++k) tvid.push_back((int32_t)t[k]);
        }
        md.tri_end = (int64_t)recs.size();
        // vertex -> incident triangles in mesh order (mesh.h:252-257)
        std::map<int64_t, std::vector<std::pair<int64_t, int>>> inc;
        for (size_t ti = 0; ti < m.tris.size(); ++ti)
            for (int k = 0; k < 3; ++k) inc[m.tris[ti][k]].push_back({md.tri_begin + (int64_t)ti, k});
        static const int opp[3][2] = {{1, 2}, {2, 0}, {0, 1}};
        for (int64_t v : m.verts) {
            slots.push_back((int32_t)v);
            auto it = inc.find(v);
            if (it != inc.end())
                for (const auto& tk : it->second) {
                    AdjRec a{};
                    a.tri = (int32_t)tk.first;
                    a.k = tk.second;
                    const TriRec& r = recs[tk.first];
                    for (int c = 0; c < 3; ++c) a.cb[c] = r.p[opp[a.k][0]][c] - r.p[opp[a.k][1]][c];
                    adj.push_back(a);
                }
            adj_off.push_back((int64_t)adj.size());
        }
        md.slot_end = (int64_t)slots.size();
        dg->meshes.push_back(md);
    }
    dg->n_tris = (int64_t)recs.size();
    dg->n_verts = (int64_t)g.vertices.size();
    dg->n_slots = (int64_t)slots.size();
    dg->tri = upload(recs);
    dg->tri_vid = upload(tvid);
    dg->vert_unk = upload(g.vert_index);
    dg->slot_vid = upload(slots);
    dg->adj_off = upload(adj_off);
    dg->adj = upload(adj);
    OMH_HIP(hipMalloc(&dg->gp, sizeof(double) * GP_STRIDE * (size_t)std::max<int64_t>(1, dg->n_tris)));
    if (dg->n_tris > 0) {
        k_gauss_table<<<(unsigned)cdiv(dg->n_tris, 256), 256>>>(dg->tri, dg->n_tris, dg->gp);
        OMH_LAUNCH_CHECK();
        OMH_HIP(hipStreamSynchronize(nullptr));
    }
    dg->device = dev;
    g.dev = dg;
    return *g.dev;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

// Minimum waves per SIMD requested for the FP64 pair kernels (register cap).
#ifndef OMH_PAIR_MIN_WAVES
#define OMH_PAIR_MIN_WAVES 2
#endif

constexpr int TILE = 16;
constexpr double kK = 1.0 / (4.0 * 3.14159265358979323846);   // K, constants.h:23

__device__ __forceinline__ int64_t cdiv_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Linear tile id -> (lo, hi) with lo <= hi over the upper triangle of tiles.
__device__ __forceinline__ void upper_tile(int64_t t, int64_t& lo, int64_t& hi) {
    int64_t h = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (h * (h + 1) / 2 > t) --h;
    while ((h + 1) * (h + 2) / 2 <= t) ++h;
    hi = h;
    lo = t - h * (h + 1) / 2;
}

// S block: value(t1,t2) = coeff * integral over t2 (Gauss) of OperatorS(t1).
// TEMP: writes the dense temporary block tmp[i + j*n1] (and its mirror for
// diagonal blocks) -- the SymBloc/Bloc of operators.h:159-173,337-352.
// Row triangles come from (tri, gp), column triangles from (triB, gpB): the
// same tables for a HeadMat pair, a separate source mesh's for SurfSourceMat.
template <bool TEMP>
__global__ void __launch_bounds__(256, OMH_PAIR_MIN_WAVES) k_s_block(const TriRec* __restrict__ tri, int64_t a0, int64_t n1,
                                                 int64_t b0, int64_t n2, int diag, int64_t ntx, int order,
                                                 double coeff, double* __restrict__ H, Layout L,
                                                 int64_t ua0, int64_t ub0, double* __restrict__ tmp,
                                                 const double* __restrict__ gp, int far,
                                                 const TriRec* __restrict__ triB, const double* __restrict__ gpB) {
    int64_t ti, tj;
    if (diag) {
        upper_tile(blockIdx.x, ti, tj);
    } else {
        ti = blockIdx.x / ntx;
        tj = blockIdx.x - ti * ntx;
    }
    const int64_t i = ti * TILE + threadIdx.x / TILE;
    const int64_t j = tj * TILE + threadIdx.x % TILE;
    if (i >= n1 || j >= n2) return;
    if (diag && j < i) return;
    const TriRec& t1 = tri[a0 + i];
    const TriRec& t2 = triB[b0 + j];
    int fc = 0;
    if (far && order == 3) {
        fc = far_class(gp + (a0 + i) * GP_STRIDE, gpB + (b0 + j) * GP_STRIDE, fmax(t1.diam, t2.diam));
    }
    double res = 0.0;
    if (fc) {   // Gauss on both triangles (outer: column triangle, inner: row triangle)
        res = t1.area * s_far(gpB + (b0 + j) * GP_STRIDE, gp + (a0 + i) * GP_STRIDE, fc);
    } else {
        const OpS o = load_ops(t1);
        const d3 P[3] = {ld3(t2.p[0]), ld3(t2.p[1]), ld3(t2.p[2])};
        const int np = c_nb_pts[order];
        for (int k = 0; k < np; ++k) res += c_rules[order][k][3] * ops_eval(o, gauss_point(order, k, P));
    }
    const double v = (t2.area * res) * coeff;
    if (TEMP) {
        tmp[i + j * n1] = v;
        if (diag) tmp[j + i * n1] = v;
    } else {
        const int64_t r = dev_index(L, ua0 + i), c = dev_index(L, ub0 + j);
        H[r + c * L.ld] = v;
        H[c + r * L.ld] = v;
    }
}

// D pass 1: W[(k*n2 + j)*n1c + i] = integral over row triangle (a0+i)
// (Gauss) of P1SLP(column triangle b0+j)_k  (operators.h:103-106).  Row
// triangles from (tri, gp), column triangles from (triB, gpB).
__global__ void __launch_bounds__(256, OMH_PAIR_MIN_WAVES) k_d_pairs(const TriRec* __restrict__ tri, int64_t a0, int64_t n1c,
                                                 int64_t b0, int64_t n2, int order, double* __restrict__ W,
                                                 const double* __restrict__ gp, int far,
                                                 const TriRec* __restrict__ triB, const double* __restrict__ gpB) {
    const int64_t ni = cdiv_dev(n1c, 64);
    const int64_t tj = blockIdx.x / ni, tb = blockIdx.x - tj * ni;
    const int64_t i = tb * 64 + threadIdx.x % 64;
    const int64_t j = tj * 4 + threadIdx.x / 64;
    if (i >= n1c || j >= n2) return;
    const TriRec& t1 = tri[a0 + i];
    const TriRec& t2 = triB[b0 + j];
    int fc = 0;
    if (far && order == 3) {
        fc = far_class(gp + (a0 + i) * GP_STRIDE, gpB + (b0 + j) * GP_STRIDE, fmax(t1.diam, t2.diam));
    }
    double rx = 0.0, ry = 0.0, rz = 0.0;
    if (fc) {   // Gauss on both triangles (outer: row triangle, inner: column triangle)
        const d3 v = d_far(gp + (a0 + i) * GP_STRIDE, gpB + (b0 + j) * GP_STRIDE, ld3(t2.n), fc);
        rx = t2.area * v.x;
        ry = t2.area * v.y;
        rz = t2.area * v.z;
    } else {
        const P1S o = load_p1s(t2);
        const d3 P[3] = {ld3(t1.p[0]), ld3(t1.p[1]), ld3(t1.p[2])};
        const int np = c_nb_pts[order];
        for (int k = 0; k < np; ++k) {
            const double w = c_rules[order][k][3];
            const d3 v = p1s_eval(o, gauss_point(order, k, P));
            rx += w * v.x;
            ry += w * v.y;
            rz += w * v.z;
        }
    }
    const double A = t1.area;
    W[(0 * n2 + j) * n1c + i] = A * rx;
    W[(1 * n2 + j) * n1c + i] = A * ry;
    W[(2 * n2 + j) * n1c + i] = A * rz;
}

// Adaptive integration of the pair integrals (Integrator(order, levels, tol)
// with levels > 0; integrator.h:69-93): the S entry integrates OperatorS of
// the row triangle over the column triangle, the D pass the P1 double layer
// of the column triangle over the row triangle, each through the stack of
// adaptive.hpp.  Same pair indexing and outputs as k_s_block / k_d_pairs.
struct OpsIntegrand {
    OpS o;
    __device__ __forceinline__ void operator()(d3 x, double* out) const { out[0] = ops_eval(o, x); }
};
struct P1Integrand {
    P1S o;
    __device__ __forceinline__ void operator()(d3 x, double* out) const {
        const d3 v = p1s_eval(o, x);
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    }
};

template <bool TEMP>
__global__ void __launch_bounds__(256) k_s_block_adapt(const TriRec* __restrict__ tri, int64_t a0, int64_t n1,
                                                       int64_t b0, int64_t n2, int diag, int64_t ntx, int order,
                                                       int levels, double tol, double coeff, double* __restrict__ H,
                                                       Layout L, int64_t ua0, int64_t ub0, double* __restrict__ tmp) {
    int64_t ti, tj;
    if (diag) {
        upper_tile(blockIdx.x, ti, tj);
    } else {
        ti = blockIdx.x / ntx;
        tj = blockIdx.x - ti * ntx;
    }
    const int64_t i = ti * TILE + threadIdx.x / TILE;
    const int64_t j = tj * TILE + threadIdx.x % TILE;
    if (i >= n1 || j >= n2) return;
    if (diag && j < i) return;
    const TriRec& t1 = tri[a0 + i];
    const TriRec& t2 = tri[b0 + j];
    OpsIntegrand f{load_ops(t1)};
    const d3 P[3] = {ld3(t2.p[0]), ld3(t2.p[1]), ld3(t2.p[2])};
    double res = 0.0;
    integrate<1>(f, order, levels, tol, P, t2.area, &res);
    const double v = res * coeff;
    if (TEMP) {
        tmp[i + j * n1] = v;
        if (diag) tmp[j + i * n1] = v;
    } else {
        const int64_t r = dev_index(L, ua0 + i), c = dev_index(L, ub0 + j);
        H[r + c * L.ld] = v;
        H[c + r * L.ld] = v;
    }
}

__global__ void __launch_bounds__(256) k_d_pairs_adapt(const TriRec* __restrict__ tri, int64_t a0, int64_t n1c,
                                                       int64_t b0, int64_t n2, int order, int levels, double tol,
                                                       double* __restrict__ W) {
    const int64_t ni = cdiv_dev(n1c, 64);
    const int64_t tj = blockIdx.x / ni, tb = blockIdx.x - tj * ni;
    const int64_t i = tb * 64 + threadIdx.x % 64;
    const int64_t j = tj * 4 + threadIdx.x / 64;
    if (i >= n1c || j >= n2) return;
    const TriRec& t1 = tri[a0 + i];
    const TriRec& t2 = tri[b0 + j];
    P1Integrand f{load_p1s(t2)};
    const d3 P[3] = {ld3(t1.p[0]), ld3(t1.p[1]), ld3(t1.p[2])};
    double r[3];
    integrate<3>(f, order, levels, tol, P, t1.area, r);
    W[(0 * n2 + j) * n1c + i] = r[0];
    W[(1 * n2 + j) * n1c + i] = r[1];
    W[(2 * n2 + j) * n1c + i] = r[2];
}

// D pass 2: H(row t1, vertex v) += sum over the triangles t2 of v (mesh
// order) of W_k(t1,t2) * coeff.  rect: the column is the vertex's own index
// (a source mesh's local vertex, not an unknown of L) and the mirror entry
// is not written (SurfSourceMat's n x nsrc matrix).
__global__ void __launch_bounds__(256) k_d_gather(const double* __restrict__ W, int64_t n1c, int64_t n2,
                                                  int64_t b0, const int64_t* __restrict__ adj_off,
                                                  const AdjRec* __restrict__ adj, const int32_t* __restrict__ slot_vid,
                                                  const int64_t* __restrict__ vert_unk, int64_t s0, int64_t ns,
                                                  double coeff, double* __restrict__ H, Layout L, int64_t urow0,
                                                  int rect) {
    const int64_t ni = cdiv_dev(n1c, 64);
    const int64_t ts = blockIdx.x / ni, tb = blockIdx.x - ts * ni;
    const int64_t i = tb * 64 + threadIdx.x % 64;
    const int64_t s = ts * 4 + threadIdx.x / 64;
    if (i >= n1c || s >= ns) return;
    const int64_t r = dev_index(L, urow0 + i);
    const int64_t vu = vert_unk[slot_vid[s0 + s]];
    const int64_t c = rect ? vu : dev_index(L, vu);
    double acc = H[r + c * L.ld];
    for (int64_t p = adj_off[s0 + s]; p < adj_off[s0 + s + 1]; ++p) {
        const AdjRec e = adj[p];
        acc += W[((int64_t)e.k * n2 + (e.tri - b0)) * n1c + i] * coeff;
    }
    H[r + c * L.ld] = acc;
    if (!rect) H[c + r * L.ld] = acc;
}

// ---------------------------------------------------------------------------
// Far-field assembly of a mesh pair (A rows, B columns), one stripe of A's
// triangles at a time: S(t, u), D(t, u) and D(u, t) of each triangle pair
// come out of one pass over the pairs, the far pairs from one set of
// Gauss-point distances (pair_far).  Pairs that are not in the 7-point far
// class go to a queue (near pairs from the front, 16-point pairs from the
// back) for a second kernel, so that the waves of the first one do not
// diverge over the classes.  The double-layer values go to two scratch
// arrays, W12[(k n2 + u) nc + t - c0] = D_k(t, u) (row fastest) and
// W21[(k nc + t - c0) n2 + u] = D_k(u, t) (column fastest), so that both
// per-vertex gathers read them along the rows of the HeadMat they write.
// When A == B only pairs u >= t are evaluated (the gathers take D(t, u) for
// u >= t from W12 and for u < t from W21).  The gathers write the (triangle
// row, vertex column) half of the D blocks only; k_mirror_tv copies it to
// the other half once all pairs are done.
// ---------------------------------------------------------------------------
struct PairStripe {
    int64_t a0, c0, nc, n1;   // A's first triangle, stripe start and size, A's triangle count
    int64_t b0, n2;           // B's first triangle, B's triangle count
    int diag;                 // A == B
    int temp;                 // S into tmp (n1 x n2) instead of the HeadMat
    double coeffS;
    int64_t ua0, ub0;         // reference unknowns of A's and B's first triangles
    double* tmp;              // S(i, j) at tmp[i + j n1]
    double* W12;
    double* W21;
    int64_t* q;               // queue, capacity qcap
    unsigned* qn;             // [near pairs, 16-point pairs]
    int64_t qcap;
};

// Writes one pair's values: sv = S(t, u) with its coefficient, a = D(t, u),
// b = D(u, t) (each already scaled by the areas).
template <bool DA, bool DB>
__device__ __forceinline__ void pair_store(const PairStripe& ps, double* __restrict__ H, const Layout& L, int64_t i,
                                           int64_t il, int64_t j, double sv, d3 a, d3 b) {
    if (ps.temp) {
        ps.tmp[i + j * ps.n1] = sv;
        if (ps.diag) ps.tmp[j + i * ps.n1] = sv;
    } else {
        const int64_t r = dev_index(L, ps.ua0 + i), c = dev_index(L, ps.ub0 + j);
        H[r + c * L.ld] = sv;
        H[c + r * L.ld] = sv;
    }
    const int64_t cs = ps.nc * ps.n2;
    if (DA) {
        const int64_t o = j * ps.nc + il;
        ps.W12[o] = a.x;
        ps.W12[o + cs] = a.y;
        ps.W12[o + 2 * cs] = a.z;
    }
    if (DB) {
        const int64_t o = il * ps.n2 + j;
        ps.W21[o] = b.x;
        ps.W21[o + cs] = b.y;
        ps.W21[o + 2 * cs] = b.z;
    }
}

// Far pair values with the area factors of s_far / d_far (operators.h:103-106).
template <bool DA, bool DB>
__device__ __forceinline__ void pair_far_values(const TriRec& T, const TriRec& U, const double* gt, const double* gu,
                                                int fc, double coeffS, double& sv, d3& a, d3& b) {
    double s;
    pair_far<DA, DB>(gt, gu, ld3(T.n), ld3(U.n), fc, s, a, b);
    sv = (U.area * (T.area * s)) * coeffS;
    a = mk(T.area * (U.area * a.x), T.area * (U.area * a.y), T.area * (U.area * a.z));
    b = mk(U.area * (T.area * b.x), U.area * (T.area * b.y), U.area * (T.area * b.z));
}

// Pass 1: one thread per pair, 16 x 16 pairs per block (a wave covers 8 x 8).
#ifndef OMH_PAIR_FAR_WAVES
#define OMH_PAIR_FAR_WAVES 3
#endif
template <bool DA, bool DB>
__global__ void __launch_bounds__(256, OMH_PAIR_FAR_WAVES) k_pairs_far(const TriRec* __restrict__ tri, const double* __restrict__ gp,
                                                   PairStripe ps, double* __restrict__ H, Layout L) {
    const int64_t ntj = cdiv_dev(ps.n2, 16);
    const int64_t ti = blockIdx.x / ntj, tj = blockIdx.x - ti * ntj;
    if (ps.diag && tj * 16 + 15 < ps.c0 + ti * 16) return;   // the whole tile lies below the diagonal
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t il = ti * 16 + (lane >> 3) + 8 * (w >> 1);
    const int64_t j = tj * 16 + (lane & 7) + 8 * (w & 1);
    const int64_t i = ps.c0 + il;
    const bool valid = il < ps.nc && j < ps.n2 && !(ps.diag && j < i);
    int fc = 2;
    if (valid)
        fc = (ps.diag && i == j) ? 0
                                 : far_class(gp + (ps.a0 + i) * GP_STRIDE, gp + (ps.b0 + j) * GP_STRIDE,
                                             fmax(tri[ps.a0 + i].diam, tri[ps.b0 + j].diam));
    // queue the other classes: one atomic per wave and class
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool mine = valid && fc == k;
        const unsigned long long m = __ballot(mine);
        if (m) {
            const int leader = __builtin_ctzll(m);
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&ps.qn[k], (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (mine) {
                const int64_t pos = (int64_t)base + __popcll(m & below);
                ps.q[k == 0 ? pos : ps.qcap - 1 - pos] = il * ps.n2 + j;
            }
        }
    }
    if (!valid || fc != 2) return;
    const TriRec& T = tri[ps.a0 + i];
    const TriRec& U = tri[ps.b0 + j];
    double sv;
    d3 a, b;
    pair_far_values<DA, DB>(T, U, gp + (ps.a0 + i) * GP_STRIDE, gp + (ps.b0 + j) * GP_STRIDE, 2, ps.coeffS, sv, a, b);
    pair_store<DA, DB>(ps, H, L, i, il, j, sv, a, b);
}

// Pass 2 over the queue: near pairs with the analytic integrals of k_s_block
// and k_d_pairs (the same arithmetic), 16-point pairs with pair_far.
template <bool DA, bool DB>
__global__ void __launch_bounds__(256, OMH_PAIR_MIN_WAVES) k_pairs_fix(const TriRec* __restrict__ tri, const double* __restrict__ gp,
                                                   PairStripe ps, double* __restrict__ H, Layout L) {
    const int64_t nn = ps.qn[0], nf = ps.qn[1];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn + nf;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t code = e < nn ? ps.q[e] : ps.q[ps.qcap - 1 - (e - nn)];
        const int64_t il = code / ps.n2, j = code - il * ps.n2, i = ps.c0 + il;
        const TriRec& T = tri[ps.a0 + i];
        const TriRec& U = tri[ps.b0 + j];
        double sv;
        d3 a = mk(0.0, 0.0, 0.0), b = mk(0.0, 0.0, 0.0);
        if (e >= nn) {
            pair_far_values<DA, DB>(T, U, gp + (ps.a0 + i) * GP_STRIDE, gp + (ps.b0 + j) * GP_STRIDE, 1, ps.coeffS,
                                    sv, a, b);
        } else {
            constexpr int order = 3;
            const int np = c_nb_pts[order];
            {   // S: OperatorS of t at the Gauss points of u
                const OpS o = load_ops(T);
                const d3 P[3] = {ld3(U.p[0]), ld3(U.p[1]), ld3(U.p[2])};
                double res = 0.0;
                for (int k = 0; k < np; ++k) res += c_rules[order][k][3] * ops_eval(o, gauss_point(order, k, P));
                sv = (U.area * res) * ps.coeffS;
            }
            if (DA) {   // D(t, u): P1 double layer of u at the Gauss points of t
                const P1S o = load_p1s(U);
                const d3 P[3] = {ld3(T.p[0]), ld3(T.p[1]), ld3(T.p[2])};
                double rx = 0.0, ry = 0.0, rz = 0.0;
                for (int k = 0; k < np; ++k) {
                    const double wk = c_rules[order][k][3];
                    const d3 v = p1s_eval(o, gauss_point(order, k, P));
                    rx += wk * v.x;
                    ry += wk * v.y;
                    rz += wk * v.z;
                }
                a = mk(T.area * rx, T.area * ry, T.area * rz);
            }
            if (DB && !(ps.diag && i == j)) {   // D(u, t)
                const P1S o = load_p1s(T);
                const d3 P[3] = {ld3(U.p[0]), ld3(U.p[1]), ld3(U.p[2])};
                double rx = 0.0, ry = 0.0, rz = 0.0;
                for (int k = 0; k < np; ++k) {
                    const double wk = c_rules[order][k][3];
                    const d3 v = p1s_eval(o, gauss_point(order, k, P));
                    rx += wk * v.x;
                    ry += wk * v.y;
                    rz += wk * v.z;
                }
                b = mk(U.area * rx, U.area * ry, U.area * rz);
            }
        }
        pair_store<DA, DB>(ps, H, L, i, il, j, sv, a, b);
    }
}

// Gather of D(t, u) for the stripe's rows t: H(t, vertex of slot s of B)
// += coeff sum over the triangles u of s (mesh order; u >= t when A == B)
// of W12(k, t - c0, u).  Thread = row (64 per block), slot (4 per block).
__global__ void __launch_bounds__(256) k_gather12(const double* __restrict__ W12, PairStripe ps,
                                                  const int64_t* __restrict__ adj_off, const AdjRec* __restrict__ adj,
                                                  const int32_t* __restrict__ slot_vid,
                                                  const int64_t* __restrict__ vert_unk, int64_t s0, int64_t ns,
                                                  double coeff, double* __restrict__ H, Layout L) {
    const int64_t ni = cdiv_dev(ps.nc, 64);
    const int64_t ts = blockIdx.x / ni, tb = blockIdx.x - ts * ni;
    const int64_t il = tb * 64 + threadIdx.x % 64;
    const int64_t s = ts * 4 + threadIdx.x / 64;
    if (il >= ps.nc || s >= ns) return;
    const int64_t i = ps.c0 + il;
    const int64_t r = dev_index(L, ps.ua0 + i), c = dev_index(L, vert_unk[slot_vid[s0 + s]]);
    const int64_t cs = ps.nc * ps.n2;
    double acc = H[r + c * L.ld];
    for (int64_t p = adj_off[s0 + s]; p < adj_off[s0 + s + 1]; ++p) {
        const AdjRec e = adj[p];
        const int64_t u = e.tri - ps.b0;
        if (ps.diag && u < i) continue;
        acc += W12[e.k * cs + u * ps.nc + il] * coeff;
    }
    H[r + c * L.ld] = acc;
}

// Gather of D(u, t) for the stripe's columns t: H(row triangle u of B,
// vertex of slot s of A) += coeff sum over the triangles t of s inside the
// stripe (t < u when A == B) of W21(k, t - c0, u).  Block = slot of A,
// threads over the rows u; slots without a triangle in the stripe return.
__global__ void __launch_bounds__(256) k_gather21(const double* __restrict__ W21, PairStripe ps,
                                                  const int64_t* __restrict__ adj_off, const AdjRec* __restrict__ adj,
                                                  const int32_t* __restrict__ slot_vid,
                                                  const int64_t* __restrict__ vert_unk, int64_t s0, double coeff,
                                                  double* __restrict__ H, Layout L) {
    const int64_t s = blockIdx.x;
    const int64_t p0 = adj_off[s0 + s], p1 = adj_off[s0 + s + 1];
    bool any = false;
    for (int64_t p = p0; p < p1; ++p) {
        const int64_t tl = adj[p].tri - ps.a0 - ps.c0;
        any |= tl >= 0 && tl < ps.nc;
    }
    if (!any) return;
    const int64_t c = dev_index(L, vert_unk[slot_vid[s0 + s]]);
    const int64_t cs = ps.nc * ps.n2;
    for (int64_t u = threadIdx.x; u < ps.n2; u += blockDim.x) {
        const int64_t r = dev_index(L, ps.ub0 + u);
        double acc = H[r + c * L.ld];
        for (int64_t p = p0; p < p1; ++p) {
            const AdjRec e = adj[p];
            const int64_t tl = e.tri - ps.a0 - ps.c0;
            if (tl < 0 || tl >= ps.nc || (ps.diag && ps.c0 + tl >= u)) continue;
            acc += W21[e.k * cs + tl * ps.n2 + u] * coeff;
        }
        H[r + c * L.ld] = acc;
    }
}

// H(v, t) = H(t, v) for the device rows [t0, t0 + nt) and columns
// [v0, v0 + nv): 64 x 64 tiles through LDS, both sides coalesced.
__global__ void __launch_bounds__(256) k_mirror_tv(double* __restrict__ H, int64_t ld, int64_t t0, int64_t nt,
                                                   int64_t v0, int64_t nv) {
    __shared__ double tile[64][65];
    const int64_t ntt = cdiv_dev(nt, 64);
    const int64_t bv = blockIdx.x / ntt, bt = blockIdx.x - bv * ntt;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = ty; k < 64; k += 4) {   // rows t along tx, columns v along k
        const int64_t t = bt * 64 + tx, v = bv * 64 + k;
        if (t < nt && v < nv) tile[k][tx] = H[(t0 + t) + (v0 + v) * ld];
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {   // rows v along tx, columns t along k
        const int64_t v = bv * 64 + tx, t = bt * 64 + k;
        if (t < nt && v < nv) H[(v0 + v) + (t0 + t) * ld] = tile[tx][k];
    }
}

__global__ void k_zero_counts(unsigned* qn) {
    if (threadIdx.x < 2) qn[threadIdx.x] = 0u;
}

// N block (operators.h:134-148 and the loops of :272-296 / :455-480):
//   N(v1, v2) = -f coeff sum_{t1 ~ v1} sum_{t2 ~ v2} (cb1 . cb2) S(t1, t2) / (A1 A2)
// (f = 1/4, or 1/2 for a vertex shared by two different meshes), evaluated
// as two contractions, both reading and writing along contiguous memory:
//   stage 1: U_d(t1, v2) = sum_{t2 ~ v2} cb2[d] / A2 S(t1, t2)   (d = 0..2)
//   stage 2: N(v1, v2) = -f coeff sum_{t1 ~ v1} (cb1 . U(t1, v2)) / A1.
// S(t1, t2) = S[t1 + t2 lds] for local triangle indices of the two meshes
// (the S block inside the HeadMat, or the dense temporary of a barrier
// pair).  U_d(t1, jj) = U[jj + t1 ldu + d nt1 ldu] for the vertex columns
// j0 + jj of a chunk (column fastest, for stage 2).
// Stage 1: a block covers 64 triangles t1 (lanes, reading S down its
// columns) x 16 vertex columns (4 per wave, incidences read once per wave);
// the 3 x 16 x 64 results are transposed in LDS and written column fastest.
constexpr int N1_ROWS = 64, N1_COLS = 16;
__global__ void __launch_bounds__(256) k_n_stage1(const TriRec* __restrict__ tri, const int64_t* __restrict__ adj_off,
                                                  const AdjRec* __restrict__ adj, int64_t slot0, int64_t nj,
                                                  const double* __restrict__ S, int64_t lds, int64_t tb0,
                                                  double* __restrict__ U, int64_t ldu, int64_t nt1) {
    __shared__ double T[3][N1_COLS][N1_ROWS + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t t1 = (int64_t)blockIdx.x * N1_ROWS + lane;
    const int64_t jb = (int64_t)blockIdx.y * N1_COLS;
#pragma unroll
    for (int k = 0; k < N1_COLS / 4; ++k) {
        const int jl = wave * (N1_COLS / 4) + k;
        const int64_t jj = jb + jl;
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
        if (jj < nj && t1 < nt1) {
            for (int64_t p = adj_off[slot0 + jj]; p < adj_off[slot0 + jj + 1]; ++p) {
                const AdjRec e = adj[p];
                const double sv = S[t1 + (e.tri - tb0) * lds] / tri[e.tri].area;
                u0 += e.cb[0] * sv;
                u1 += e.cb[1] * sv;
                u2 += e.cb[2] * sv;
            }
        }
        T[0][jl][lane] = u0;
        T[1][jl][lane] = u1;
        T[2][jl][lane] = u2;
    }
    __syncthreads();
    const int64_t cs = nt1 * ldu;
    const int jl = threadIdx.x % N1_COLS;
    for (int r = threadIdx.x / N1_COLS; r < N1_ROWS; r += 256 / N1_COLS) {
        const int64_t row = (int64_t)blockIdx.x * N1_ROWS + r, jj = jb + jl;
        if (row < nt1 && jj < nj) {
            double* u = U + jj + row * ldu;
            u[0] = T[0][jl][r];
            u[cs] = T[1][jl][r];
            u[2 * cs] = T[2][jl][r];
        }
    }
}

// Stage 2: threads over the chunk's columns (vertex slots j0 + jj of mesh
// B), block row y = vertex slot a of mesh A; the rows U(t1, .) of a's
// triangles are read along the columns.  A block of one mesh computes the
// columns b >= a.  Only the lower-triangle entry of each vertex pair is
// accumulated (k_mirror_lower fills the upper one once the V x V block is
// complete).
// rect: mesh B is a separate source mesh (its slot and vertex tables
// slot_vid_b / vert_unk_b): the column is B's local vertex index, no vertex
// is shared, and each entry is written once, at (row, column).
__global__ void __launch_bounds__(256) k_n_stage2(const TriRec* __restrict__ tri, const int64_t* __restrict__ adj_off,
                                                  const AdjRec* __restrict__ adj, const int32_t* __restrict__ slot_vid,
                                                  const int64_t* __restrict__ vert_unk, int64_t sa0, int64_t na,
                                                  int64_t sb0, int64_t j0, int64_t nj, int same, int64_t ta0,
                                                  double coeff, const double* __restrict__ U, int64_t ldu,
                                                  int64_t nt1, double* __restrict__ H, Layout L, int use_atomic,
                                                  const int32_t* __restrict__ slot_vid_b,
                                                  const int64_t* __restrict__ vert_unk_b, int rect) {
    const int64_t jj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= nj) return;
    const int64_t b = j0 + jj;
    const int64_t cs = nt1 * ldu;
    const int32_t v2 = slot_vid_b[sb0 + b];
    const int64_t c = rect ? vert_unk_b[v2] : dev_index(L, vert_unk_b[v2]);
    for (int64_t a = blockIdx.y; a < na; a += gridDim.y) {
        if (same && b < a) continue;
        double acc = 0.0;
        for (int64_t p = adj_off[sa0 + a]; p < adj_off[sa0 + a + 1]; ++p) {
            const AdjRec e = adj[p];
            const double ia = 1.0 / tri[e.tri].area;
            const double w0 = e.cb[0] * ia, w1 = e.cb[1] * ia, w2 = e.cb[2] * ia;
            const double* u = U + jj + (e.tri - ta0) * ldu;
            acc += w0 * u[0] + w1 * u[cs] + w2 * u[2 * cs];
        }
        const int32_t v1 = slot_vid[sa0 + a];
        const double factor = (same || rect) ? 0.25 : ((v1 == v2) ? 0.5 : 0.25);
        const double val = -factor * acc * coeff;
        const int64_t r = dev_index(L, vert_unk[v1]);
        double* h = (rect || r > c) ? &H[r + c * L.ld] : &H[c + r * L.ld];
        if (use_atomic) atomicAdd(h, val);
        else *h += val;
    }
}

// H(j, i) = H(i, j) for i > j inside the square block [o, o + nb)^2:
// 64 x 64 tiles on or below the diagonal through LDS, both sides coalesced.
__global__ void __launch_bounds__(256) k_mirror_lower(double* __restrict__ H, int64_t ld, int64_t o, int64_t nb) {
    __shared__ double tile[64][65];
    int64_t bj, bi;
    upper_tile(blockIdx.x, bj, bi);   // bj <= bi
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = ty; k < 64; k += 4) {   // rows i along tx, columns j along k
        const int64_t i = bi * 64 + tx, j = bj * 64 + k;
        if (i < nb && j < nb && i > j) tile[k][tx] = H[(o + i) + (o + j) * ld];
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {   // rows j along tx, columns i along k
        const int64_t j = bj * 64 + tx, i = bi * 64 + k;
        if (i < nb && j < nb && i > j) H[(o + j) + (o + i) * ld] = tile[tx][k];
    }
}

__global__ void k_deflate_coef(const double* __restrict__ H, Layout L, int64_t u, double nbv, double* coef) {
    const int64_t d = dev_index(L, u);
    coef[0] = H[d + d * L.ld] / nbv;
}

// deflate (assembleHeadMat.cpp:40-57): every (v1, v2) vertex pair of the
// mesh with v2 at or after v1 in mesh order gets += coef.
__global__ void __launch_bounds__(256) k_deflate(const int32_t* __restrict__ slot_vid, const int64_t* __restrict__ vert_unk,
                                                 int64_t s0, int64_t ns, const double* __restrict__ coef,
                                                 double* __restrict__ H, Layout L) {
    int64_t ta, tb;
    upper_tile(blockIdx.x, ta, tb);   // ta <= tb
    const int64_t a = ta * TILE + threadIdx.x / TILE;
    const int64_t b = tb * TILE + threadIdx.x % TILE;
    if (a >= ns || b >= ns || b < a) return;
    const double c0 = coef[0];
    const int64_t r = dev_index(L, vert_unk[slot_vid[s0 + a]]), c = dev_index(L, vert_unk[slot_vid[s0 + b]]);
    H[r + c * L.ld] += c0;
    if (r != c) H[c + r * L.ld] += c0;
}

// ---------------------------------------------------------------------------
// Triangle-triangle intersection (the overlap check of SurfSourceMat,
// Geometry::check, geometry.cpp:329-343, with Mesh::intersection,
// mesh.cpp:246-252, over every pair of a geometry triangle and a source
// triangle).  Closed test from orientation determinants: two triangles
// that are not coplanar meet iff an edge of one meets the other (the ends
// of their common segment lie on the triangles' boundaries); exactly
// coplanar triangles are tested in the plane (edge crossings, then one
// vertex inside the other triangle).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double orient3(d3 a, d3 b, d3 c, d3 d) {   // det[b - a, c - a, d - a]
    return dot(cross(sub(b, a), sub(c, a)), sub(d, a));
}

// Segment [a, b] against triangle (p, q, r), the segment not in its plane.
__device__ __forceinline__ bool seg_meets_tri(d3 a, d3 b, d3 p, d3 q, d3 r) {
    const double sa = orient3(p, q, r, a), sb = orient3(p, q, r, b);
    if ((sa > 0.0 && sb > 0.0) || (sa < 0.0 && sb < 0.0) || (sa == 0.0 && sb == 0.0)) return false;
    const double v0 = orient3(a, b, p, q), v1 = orient3(a, b, q, r), v2 = orient3(a, b, r, p);
    return (v0 >= 0.0 && v1 >= 0.0 && v2 >= 0.0) || (v0 <= 0.0 && v1 <= 0.0 && v2 <= 0.0);
}

struct p2 { double u, v; };
__device__ __forceinline__ double orient2(p2 a, p2 b, p2 c) { return (b.u - a.u) * (c.v - a.v) - (b.v - a.v) * (c.u - a.u); }
__device__ __forceinline__ bool on_seg2(p2 a, p2 b, p2 c) {   // c collinear with [a, b]: inside its box
    return fmin(a.u, b.u) <= c.u && c.u <= fmax(a.u, b.u) && fmin(a.v, b.v) <= c.v && c.v <= fmax(a.v, b.v);
}
__device__ __forceinline__ bool segs_meet2(p2 a, p2 b, p2 c, p2 d) {
    const double d1 = orient2(c, d, a), d2 = orient2(c, d, b), d3v = orient2(a, b, c), d4 = orient2(a, b, d);
    if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3v > 0 && d4 < 0) || (d3v < 0 && d4 > 0))) return true;
    return (d1 == 0 && on_seg2(c, d, a)) || (d2 == 0 && on_seg2(c, d, b)) || (d3v == 0 && on_seg2(a, b, c)) ||
           (d4 == 0 && on_seg2(a, b, d));
}
__device__ __forceinline__ bool inside_tri2(p2 x, const p2 T[3]) {
    const double o0 = orient2(T[0], T[1], x), o1 = orient2(T[1], T[2], x), o2 = orient2(T[2], T[0], x);
    return (o0 >= 0 && o1 >= 0 && o2 >= 0) || (o0 <= 0 && o1 <= 0 && o2 <= 0);
}

__device__ bool tris_meet(const d3 P[3], const d3 Q[3]) {
    const double s0 = orient3(Q[0], Q[1], Q[2], P[0]), s1 = orient3(Q[0], Q[1], Q[2], P[1]),
                 s2 = orient3(Q[0], Q[1], Q[2], P[2]);
    if ((s0 > 0 && s1 > 0 && s2 > 0) || (s0 < 0 && s1 < 0 && s2 < 0)) return false;
    if (s0 == 0.0 && s1 == 0.0 && s2 == 0.0) {   // coplanar: drop the normal's largest axis
        const d3 nq = cross(sub(Q[1], Q[0]), sub(Q[2], Q[0]));
        const double ax = fabs(nq.x), ay = fabs(nq.y), az = fabs(nq.z);
        p2 A[3], B[3];
        for (int k = 0; k < 3; ++k) {
            if (ax >= ay && ax >= az) { A[k] = {P[k].y, P[k].z}; B[k] = {Q[k].y, Q[k].z}; }
            else if (ay >= az) { A[k] = {P[k].x, P[k].z}; B[k] = {Q[k].x, Q[k].z}; }
            else { A[k] = {P[k].x, P[k].y}; B[k] = {Q[k].x, Q[k].y}; }
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (segs_meet2(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3])) return true;
        return inside_tri2(A[0], B) || inside_tri2(B[0], A);
    }
    for (int k = 0; k < 3; ++k) {
        if (seg_meets_tri(P[k], P[(k + 1) % 3], Q[0], Q[1], Q[2])) return true;
        if (seg_meets_tri(Q[k], Q[(k + 1) % 3], P[0], P[1], P[2])) return true;
    }
    return false;
}

// hit[0] |= 1 when a triangle of A (na records from a0) meets a triangle
// of B (nb records).  Thread = B triangle, block row = A triangles (looped).
__global__ void __launch_bounds__(256) k_tri_overlap(const TriRec* __restrict__ triA, int64_t a0, int64_t na,
                                                     const TriRec* __restrict__ triB, int64_t nb,
                                                     unsigned* __restrict__ hit) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const d3 Q[3] = {ld3(triB[j].p[0]), ld3(triB[j].p[1]), ld3(triB[j].p[2])};
    const d3 qlo = mk(fmin(Q[0].x, fmin(Q[1].x, Q[2].x)), fmin(Q[0].y, fmin(Q[1].y, Q[2].y)),
                      fmin(Q[0].z, fmin(Q[1].z, Q[2].z)));
    const d3 qhi = mk(fmax(Q[0].x, fmax(Q[1].x, Q[2].x)), fmax(Q[0].y, fmax(Q[1].y, Q[2].y)),
                      fmax(Q[0].z, fmax(Q[1].z, Q[2].z)));
    for (int64_t i = blockIdx.y; i < na; i += gridDim.y) {
        const TriRec& t = triA[a0 + i];
        const d3 P[3] = {ld3(t.p[0]), ld3(t.p[1]), ld3(t.p[2])};
        if (fmax(P[0].x, fmax(P[1].x, P[2].x)) < qlo.x || fmin(P[0].x, fmin(P[1].x, P[2].x)) > qhi.x) continue;
        if (fmax(P[0].y, fmax(P[1].y, P[2].y)) < qlo.y || fmin(P[0].y, fmin(P[1].y, P[2].y)) > qhi.y) continue;
        if (fmax(P[0].z, fmax(P[1].z, P[2].z)) < qlo.z || fmin(P[0].z, fmin(P[1].z, P[2].z)) > qhi.z) continue;
        if (tris_meet(P, Q)) {
            atomicOr(hit, 1u);
            return;
        }
    }
}

int64_t n_upper_tiles(int64_t n) {
    const int64_t t = cdiv(n, TILE);
    return t * (t + 1) / 2;
}

}  // namespace

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static void launch_s(DeviceGeometry& dg, int m1, int m2, int order, double coeff, double* H, const Layout& L,
                     double* tmp, int far, hipStream_t st, int levels = 0, double tol = 0.0) {
    const MeshDev& A = dg.meshes[m1];
    const MeshDev& B = dg.meshes[m2];
    const int64_t n1 = A.tri_end - A.tri_begin, n2 = B.tri_end - B.tri_begin;
    const int diag = (m1 == m2);
    const int64_t ntx = cdiv(n2, TILE);
    const int64_t nblk = diag ? n_upper_tiles(n1) : cdiv(n1, TILE) * ntx;
    // algorithmic work: OperatorS evaluations (pairs x Gauss points)
    static const int npts[4] = {3, 6, 7, 16};
    const double pairs = diag ? 0.5 * (double)n1 * (double)(n1 + 1) : (double)n1 * (double)n2;
    if (levels > 0) {
        if (tmp)
            k_s_block_adapt<true><<<(unsigned)nblk, 256, 0, st>>>(dg.tri, A.tri_begin, n1, B.tri_begin, n2, diag, ntx,
                                                                 order, levels, tol, coeff, H, L, A.tri_unk0,
                                                                 B.tri_unk0, tmp);
        else
            k_s_block_adapt<false><<<(unsigned)nblk, 256, 0, st>>>(dg.tri, A.tri_begin, n1, B.tri_begin, n2, diag, ntx,
                                                                  order, levels, tol, coeff, H, L, A.tri_unk0,
                                                                  B.tri_unk0, tmp);
        OMH_LAUNCH_CHECK();
        return;
    }
    ProfScope ps("s_block", pairs * npts[order], st);
    if (tmp)
        k_s_block<true><<<(unsigned)nblk, 256, 0, st>>>(dg.tri, A.tri_begin, n1, B.tri_begin, n2, diag, ntx, order,
                                                        coeff, H, L, A.tri_unk0, B.tri_unk0, tmp, dg.gp, far, dg.tri,
                                                        dg.gp);
    else
        k_s_block<false><<<(unsigned)nblk, 256, 0, st>>>(dg.tri, A.tri_begin, n1, B.tri_begin, n2, diag, ntx, order,
                                                         coeff, H, L, A.tri_unk0, B.tri_unk0, tmp, dg.gp, far, dg.tri,
                                                         dg.gp);
    OMH_LAUNCH_CHECK();
}

static void launch_n(DeviceGeometry& dg, const Geometry& g, int m1, int m2, double coeff, const double* tmp,
                     double* H, const Layout& L, hipStream_t st) {
    const MeshDev& A = dg.meshes[m1];
    const MeshDev& B = dg.meshes[m2];
    const int64_t na = A.slot_end - A.slot_begin, nb = B.slot_end - B.slot_begin;
    const int64_t nt1 = A.tri_end - A.tri_begin;
    if (na <= 0 || nb <= 0 || nt1 <= 0) return;
    const int same = (m1 == m2);
    int use_atomic = 0;
    if (!same) {   // vertices shared by the two meshes can alias entries
        const auto& va = g.meshes[m1].verts;
        std::vector<int64_t> sa(va.begin(), va.end());
        std::sort(sa.begin(), sa.end());
        for (int64_t v : g.meshes[m2].verts)
            if (std::binary_search(sa.begin(), sa.end(), v)) { use_atomic = 1; break; }
    }
    // the S block: inside the HeadMat (triangle unknowns of a mesh are
    // contiguous in either ordering) or the temporary of a barrier pair
    const double* S;
    int64_t lds;
    if (tmp) {
        S = tmp;
        lds = nt1;
    } else {
        S = H + dev_index(L, A.tri_unk0) + dev_index(L, B.tri_unk0) * L.ld;
        lds = L.ld;
    }
    const size_t budget = size_t(1) << 30;   // 1 GiB of U per chunk (shares the D pass-1 scratch)
    int64_t chunk = std::max<int64_t>(32, (int64_t)(budget / (3 * sizeof(double) * (size_t)nt1)) / 32 * 32);
    chunk = std::min<int64_t>(chunk, nb);
    // OMH_N_CHUNK_COLS caps the columns per chunk (tests run the chunked
    // path on small meshes with it)
    if (const char* env = std::getenv("OMH_N_CHUNK_COLS")) {
        const long long v = std::atoll(env);
        if (v > 0) chunk = std::min<int64_t>(chunk, (int64_t)v);
    }
    const int64_t ldu = (chunk + 31) / 32 * 32;
    double* U = (double*)scratch(SCR_DTMP, 3 * sizeof(double) * (size_t)ldu * (size_t)nt1, st);
    for (int64_t j0 = 0; j0 < nb; j0 += chunk) {
        const int64_t nj = std::min<int64_t>(chunk, nb - j0);
        const dim3 g1((unsigned)cdiv(nt1, N1_ROWS), (unsigned)cdiv(nj, N1_COLS));
        k_n_stage1<<<g1, 256, 0, st>>>(dg.tri, dg.adj_off, dg.adj, B.slot_begin + j0, nj, S, lds, B.tri_begin, U, ldu,
                                       nt1);
        OMH_LAUNCH_CHECK();
        const dim3 g2((unsigned)cdiv(nj, 256), (unsigned)std::min<int64_t>(na, 65535));
        k_n_stage2<<<g2, 256, 0, st>>>(dg.tri, dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk, A.slot_begin, na,
                                       B.slot_begin, j0, nj, same, A.tri_begin, coeff, U, ldu, nt1, H, L, use_atomic,
                                       dg.slot_vid, dg.vert_unk, 0);
        OMH_LAUNCH_CHECK();
    }
}

// D(rows = triangles of mesh mr, columns = vertices of mesh mc), operators.h:84-115.
static void launch_d(DeviceGeometry& dg, int mr, int mc, int order, double coeff, double* H, const Layout& L,
                     int far, hipStream_t st, int levels = 0, double tol = 0.0) {
    const MeshDev& R = dg.meshes[mr];
    const MeshDev& Cm = dg.meshes[mc];
    const int64_t n1 = R.tri_end - R.tri_begin, n2 = Cm.tri_end - Cm.tri_begin;
    const int64_t ns = Cm.slot_end - Cm.slot_begin;
    const size_t budget = size_t(1) << 30;   // 1 GiB of pass-1 scratch
    int64_t chunk = std::max<int64_t>(64, (int64_t)(budget / (3 * sizeof(double) * (size_t)n2)));
    chunk = std::min(chunk, n1);
    double* W = (double*)scratch(SCR_DTMP, 3 * sizeof(double) * (size_t)n2 * (size_t)chunk, st);
    for (int64_t c0 = 0; c0 < n1; c0 += chunk) {
        const int64_t nc = std::min(chunk, n1 - c0);
        const int64_t blk1 = cdiv(nc, 64) * cdiv(n2, 4);
        static const int npts[4] = {3, 6, 7, 16};
        if (levels > 0) {
            k_d_pairs_adapt<<<(unsigned)blk1, 256, 0, st>>>(dg.tri, R.tri_begin + c0, nc, Cm.tri_begin, n2, order,
                                                            levels, tol, W);
        } else {
            ProfScope ps("d_pairs", (double)nc * (double)n2 * npts[order], st);
            k_d_pairs<<<(unsigned)blk1, 256, 0, st>>>(dg.tri, R.tri_begin + c0, nc, Cm.tri_begin, n2, order, W, dg.gp,
                                                      far, dg.tri, dg.gp);
        }
        OMH_LAUNCH_CHECK();
        const int64_t blk2 = cdiv(nc, 64) * cdiv(ns, 4);
        k_d_gather<<<(unsigned)blk2, 256, 0, st>>>(W, nc, n2, Cm.tri_begin, dg.adj_off, dg.adj, dg.slot_vid,
                                                   dg.vert_unk, Cm.slot_begin, ns, coeff, H, L, R.tri_unk0 + c0, 0);
        OMH_LAUNCH_CHECK();
    }
}

// ---------------------------------------------------------------------------
// SurfSourceMat
// ---------------------------------------------------------------------------
bool meshes_intersect(Geometry& g, Geometry& src, hipStream_t st) {
    DeviceGeometry& dg = device_geometry(g);
    DeviceGeometry& ds = device_geometry(src);
    if (dg.n_tris <= 0 || ds.n_tris <= 0) return false;
    unsigned* hit = (unsigned*)scratch(SCR_SMALL, 64, st);
    k_zero_counts<<<1, 64, 0, st>>>(hit);
    OMH_LAUNCH_CHECK();
    const dim3 grid((unsigned)cdiv(ds.n_tris, 256), (unsigned)std::min<int64_t>(dg.n_tris, 1024));
    k_tri_overlap<<<grid, 256, 0, st>>>(dg.tri, 0, dg.n_tris, ds.tri, ds.n_tris, hit);
    OMH_LAUNCH_CHECK();
    unsigned h = 0;
    OMH_HIP(hipMemcpyAsync(&h, hit, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    return h != 0;
}

void surf_source_mat(Geometry& g, Geometry& src, int order, double* out, const Layout& L, hipStream_t st,
                     bool far_field) {
    if (order < 0 || order > 3) throw std::invalid_argument("Gauss order must be 0..3");
    if (src.meshes.size() != 1) throw std::invalid_argument("a source mesh is one mesh");
    const int64_t n = g.size(), nsrc = (int64_t)src.vertices.size();
    if (meshes_intersect(g, src, st)) throw OverlappingSourceMesh();
    const int dom = g.domain_of_point(src.vertices[src.meshes[0].verts.front()].data());
    if (dom < 0) throw std::invalid_argument("the source mesh lies in no domain of the geometry");
    const double sigma = g.domains[dom].cond;
    if (sigma == 0.0) throw std::invalid_argument("the source mesh lies in a domain of zero conductivity");
    DeviceGeometry& dg = device_geometry(g);
    DeviceGeometry& ds = device_geometry(src);
    fill_f64(out, n, nsrc, L.ld, 0.0, st);
    const int far = (far_field && order == 3) ? 1 : 0;
    const double Lc = -1.0 / sigma;
    const MeshDev& B = ds.meshes[0];
    const int64_t n2 = B.tri_end - B.tri_begin, nb = B.slot_end - B.slot_begin;
    static const int npts[4] = {3, 6, 7, 16};
    for (const auto& bd : g.domains[dom].boundaries) {
        const double factorN = bd.second ? kK : -kK;
        for (const auto& om : bd.first.omeshes) {
            const double cN = factorN * om.second;
            const MeshDev& A = dg.meshes[om.first];
            const int64_t n1 = A.tri_end - A.tri_begin, na = A.slot_end - A.slot_begin;
            if (n1 <= 0 || na <= 0) continue;
            // N block from an S block of coefficient 1 in a temporary
            // (operators.h:428-437)
            double* tmp = (double*)scratch(SCR_ASSEMBLY, sizeof(double) * (size_t)n1 * (size_t)n2, st);
            {
                const int64_t ntx = cdiv(n2, TILE);
                ProfScope ps("s_block", (double)n1 * (double)n2 * npts[order], st);
                k_s_block<true><<<(unsigned)(cdiv(n1, TILE) * ntx), 256, 0, st>>>(
                    dg.tri, A.tri_begin, n1, B.tri_begin, n2, 0, ntx, order, 1.0, nullptr, L, 0, 0, tmp, dg.gp, far,
                    ds.tri, ds.gp);
                OMH_LAUNCH_CHECK();
            }
            const size_t budget = size_t(1) << 30;
            const int64_t chunk = std::min<int64_t>(
                nb, std::max<int64_t>(32, (int64_t)(budget / (3 * sizeof(double) * (size_t)n1)) / 32 * 32));
            const int64_t ldu = (chunk + 31) / 32 * 32;
            double* U = (double*)scratch(SCR_DTMP, 3 * sizeof(double) * (size_t)ldu * (size_t)n1, st);
            for (int64_t j0 = 0; j0 < nb; j0 += chunk) {
                const int64_t nj = std::min<int64_t>(chunk, nb - j0);
                const dim3 g1((unsigned)cdiv(n1, N1_ROWS), (unsigned)cdiv(nj, N1_COLS));
                k_n_stage1<<<g1, 256, 0, st>>>(ds.tri, ds.adj_off, ds.adj, B.slot_begin + j0, nj, tmp, n1,
                                               B.tri_begin, U, ldu, n1);
                OMH_LAUNCH_CHECK();
                const dim3 g2((unsigned)cdiv(nj, 256), (unsigned)std::min<int64_t>(na, 65535));
                k_n_stage2<<<g2, 256, 0, st>>>(dg.tri, dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk, A.slot_begin, na,
                                               B.slot_begin, j0, nj, 0, A.tri_begin, cN, U, ldu, n1, out, L, 0,
                                               ds.slot_vid, ds.vert_unk, 1);
                OMH_LAUNCH_CHECK();
            }
            // D block (operators.h:439-443), rows = the head mesh's triangle
            // unknowns; a current barrier's triangles have no rows
            if (A.barrier) continue;
            const int64_t dchunk = std::min<int64_t>(n1, std::max<int64_t>(64, (int64_t)(budget / (3 * sizeof(double) * (size_t)n2))));
            double* W = (double*)scratch(SCR_DTMP, 3 * sizeof(double) * (size_t)n2 * (size_t)dchunk, st);
            for (int64_t c0 = 0; c0 < n1; c0 += dchunk) {
                const int64_t nc = std::min(dchunk, n1 - c0);
                ProfScope ps("d_pairs", (double)nc * (double)n2 * npts[order], st);
                k_d_pairs<<<(unsigned)(cdiv(nc, 64) * cdiv(n2, 4)), 256, 0, st>>>(
                    dg.tri, A.tri_begin + c0, nc, B.tri_begin, n2, order, W, dg.gp, far, ds.tri, ds.gp);
                OMH_LAUNCH_CHECK();
                k_d_gather<<<(unsigned)(cdiv(nc, 64) * cdiv(nb, 4)), 256, 0, st>>>(
                    W, nc, n2, B.tri_begin, ds.adj_off, ds.adj, ds.slot_vid, ds.vert_unk, B.slot_begin, nb, cN * Lc,
                    out, L, A.tri_unk0 + c0, 1);
                OMH_LAUNCH_CHECK();
            }
        }
    }
}

// Far-field assembly of the S, D and D* blocks of one mesh pair (m1 rows,
// m2 columns): S with coefficient coeffS into the HeadMat, or into tmp
// (n1 x n2) for a barrier pair; D(m1, m2) (da) and D(m2, m1) (db) with
// coefficient cD.  m1 == m2: da and db together make the diagonal D block.
template <bool DA, bool DB>
static void launch_pairs_far_t(DeviceGeometry& dg, int m1, int m2, double coeffS, double* tmp, double cD, double* H,
                               const Layout& L, hipStream_t st) {
    const MeshDev& A = dg.meshes[m1];
    const MeshDev& B = dg.meshes[m2];
    const int64_t n1 = A.tri_end - A.tri_begin, n2 = B.tri_end - B.tri_begin;
    if (n1 <= 0 || n2 <= 0) return;
    const int diag = (m1 == m2);
    const size_t budget = size_t(1) << 30;   // bytes per W array
    int64_t nc = (int64_t)(budget / (3 * sizeof(double) * (size_t)n2));
    nc = std::max<int64_t>(16, std::min<int64_t>({nc, n1, 32768}));
    if (const char* env = std::getenv("OMH_PAIR_STRIPE")) {   // tests run the striped path on small meshes
        const long long v = std::atoll(env);
        if (v > 0) nc = std::min<int64_t>(nc, (int64_t)v);
    }
    const size_t wsz = 3 * (size_t)nc * (size_t)n2;
    double* W = (DA || DB) ? (double*)scratch(SCR_DTMP, sizeof(double) * 2 * wsz, st) : nullptr;
    const int64_t qcap = nc * n2;
    char* qb = (char*)scratch(SCR_PAIRQ, sizeof(int64_t) * (size_t)qcap + 256, st);
    PairStripe ps{};
    ps.a0 = A.tri_begin;
    ps.n1 = n1;
    ps.b0 = B.tri_begin;
    ps.n2 = n2;
    ps.diag = diag;
    ps.temp = tmp != nullptr;
    ps.coeffS = coeffS;
    ps.ua0 = A.tri_unk0;
    ps.ub0 = B.tri_unk0;
    ps.tmp = tmp;
    ps.W12 = W;
    ps.W21 = W ? W + wsz : nullptr;
    ps.qn = (unsigned*)qb;
    ps.q = (int64_t*)(qb + 256);
    ps.qcap = qcap;
    const int64_t ns1 = A.slot_end - A.slot_begin, ns2 = B.slot_end - B.slot_begin;
    for (int64_t c0 = 0; c0 < n1; c0 += nc) {
        ps.c0 = c0;
        ps.nc = std::min(nc, n1 - c0);
        const double pairs = diag ? (double)ps.nc * (double)(n2 - c0) - 0.5 * (double)ps.nc * (double)(ps.nc - 1)
                                  : (double)ps.nc * (double)n2;
        ProfScope prof("pairs_far", pairs, st);
        k_zero_counts<<<1, 64, 0, st>>>(ps.qn);
        OMH_LAUNCH_CHECK();
        k_pairs_far<DA, DB><<<(unsigned)(cdiv(ps.nc, 16) * cdiv(n2, 16)), 256, 0, st>>>(dg.tri, dg.gp, ps, H, L);
        OMH_LAUNCH_CHECK();
        k_pairs_fix<DA, DB><<<2048, 256, 0, st>>>(dg.tri, dg.gp, ps, H, L);
        OMH_LAUNCH_CHECK();
        if (DA && ns2 > 0) {
            k_gather12<<<(unsigned)(cdiv(ps.nc, 64) * cdiv(ns2, 4)), 256, 0, st>>>(
                ps.W12, ps, dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk, B.slot_begin, ns2, cD, H, L);
            OMH_LAUNCH_CHECK();
        }
        if (DB && ns1 > 0) {
            k_gather21<<<(unsigned)ns1, 256, 0, st>>>(ps.W21, ps, dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk,
                                                       A.slot_begin, cD, H, L);
            OMH_LAUNCH_CHECK();
        }
    }
}

static void launch_pairs_far(DeviceGeometry& dg, int m1, int m2, double coeffS, double* tmp, bool da, bool db,
                             double cD, double* H, const Layout& L, hipStream_t st) {
    if (da && db) launch_pairs_far_t<true, true>(dg, m1, m2, coeffS, tmp, cD, H, L, st);
    else if (da) launch_pairs_far_t<true, false>(dg, m1, m2, coeffS, tmp, cD, H, L, st);
    else if (db) launch_pairs_far_t<false, true>(dg, m1, m2, coeffS, tmp, cD, H, L, st);
    else launch_pairs_far_t<false, false>(dg, m1, m2, coeffS, tmp, cD, H, L, st);
}

void assemble_headmat(Geometry& g, int order, double* H, const Layout& L, hipStream_t st, bool far_field, int levels,
                      double tol) {
    if (order < 0 || order > 3) throw std::invalid_argument("Gauss order must be 0..3");
    if (levels < 0 || levels >= ADAPT_MAX_FRAMES) throw std::invalid_argument("adaptive levels must be 0..11");
    const int far = (far_field && levels == 0) ? 1 : 0;   // adaptive integration: analytic pairs only
    DeviceGeometry& dg = device_geometry(g);
    const int64_t n = g.size();
    fill_f64(H, n, n, L.ld, 0.0, st);
    // OMH_FAR_SEPARATE=1: the far field through the per-operator kernels
    // (tests compare the fused pass against them)
    const char* sep = std::getenv("OMH_FAR_SEPARATE");
    const bool fused = far && order == 3 && !(sep && std::atoi(sep) != 0);
    for (const MeshPair& mp : g.pairs) {
        const int m1 = mp.m1, m2 = mp.m2;
        const bool b1 = g.meshes[m1].current_barrier, b2 = g.meshes[m2].current_barrier;
        const bool same = (m1 == m2);
        if (fused) {   // S, D and D* of the pair in one pass, then N from S
            const bool inH = (same ? !b1 : (!b1 && !b2)) && mp.cS != 0.0;
            double* tmp = nullptr;
            if (!inH) {
                const int64_t n1 = (int64_t)g.meshes[m1].tris.size(), n2 = (int64_t)g.meshes[m2].tris.size();
                tmp = (double*)scratch(SCR_ASSEMBLY, sizeof(double) * (size_t)n1 * (size_t)n2, st);
            }
            launch_pairs_far(dg, m1, m2, inH ? mp.cS : 1.0, tmp, !b1, same ? !b1 : !b2, mp.cD, H, L, st);
            launch_n(dg, g, m1, m2, inH ? mp.cN / mp.cS : mp.cN, tmp, H, L, st);
            continue;
        }
        double Scoeff = 0.0;
        if (same ? !b1 : (!b1 && !b2)) {   // set_S_block
            launch_s(dg, m1, m2, order, mp.cS, H, L, nullptr, far, st, levels, tol);
            Scoeff = mp.cS;
        }
        if (Scoeff != 0.0) {               // set_N_block
            launch_n(dg, g, m1, m2, mp.cN / Scoeff, nullptr, H, L, st);
        } else {
            const int64_t n1 = (int64_t)g.meshes[m1].tris.size(), n2 = (int64_t)g.meshes[m2].tris.size();
            double* tmp = (double*)scratch(SCR_ASSEMBLY, sizeof(double) * (size_t)n1 * (size_t)n2, st);
            launch_s(dg, m1, m2, order, 1.0, H, L, tmp, far, st, levels, tol);
            launch_n(dg, g, m1, m2, mp.cN, tmp, H, L, st);
        }
        if (!b1) launch_d(dg, m1, m2, order, mp.cD, H, L, far, st, levels, tol);              // set_D_block
        if (!same && !b2) launch_d(dg, m2, m1, order, mp.cD, H, L, far, st, levels, tol);     // set_Dstar_block
    }
    if (fused) {   // the D blocks' other half
        const int64_t nt = L.nt, nv = L.nv;
        const int64_t t0 = L.tri_first ? 0 : nv, v0 = L.tri_first ? nt : 0;
        if (nt > 0 && nv > 0) {
            k_mirror_tv<<<(unsigned)(cdiv(nt, 64) * cdiv(nv, 64)), 256, 0, st>>>(H, L.ld, t0, nt, v0, nv);
            OMH_LAUNCH_CHECK();
        }
    }
    // deflate all current barriers, one constant per isolated part
    double* coef = (double*)scratch(SCR_SMALL, 64, st);
    for (const auto& part : g.parts) {
        int64_t nbv = 0, i_first = 0;
        for (int mid : part) {
            const Mesh& m = g.meshes[mid];
            if (m.outermost) {
                nbv += (int64_t)m.verts.size();
                if (i_first == 0) i_first = g.vert_index[m.verts.front()];
            }
        }
        if (nbv == 0) continue;
        k_deflate_coef<<<1, 1, 0, st>>>(H, L, i_first, (double)nbv, coef);
        OMH_LAUNCH_CHECK();
        for (int mid : part) {
            if (!g.meshes[mid].outermost) continue;
            const MeshDev& md = dg.meshes[mid];
            const int64_t ns = md.slot_end - md.slot_begin;
            k_deflate<<<(unsigned)n_upper_tiles(ns), 256, 0, st>>>(dg.slot_vid, dg.vert_unk, md.slot_begin, ns, coef,
                                                                   H, L);
            OMH_LAUNCH_CHECK();
        }
    }
    // the N blocks were accumulated in the lower triangle of the V x V block
    if (L.nv > 0) {
        const int64_t v0 = L.tri_first ? L.nt : 0;
        const int64_t t = cdiv(L.nv, 64);
        k_mirror_lower<<<(unsigned)(t * (t + 1) / 2), 256, 0, st>>>(H, L.ld, v0, L.nv);
        OMH_LAUNCH_CHECK();
    }
}

}  // namespace omh
