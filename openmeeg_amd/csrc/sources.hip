// Right-hand sides and sensors on the GPU: DipSourceMat (adaptive
// quadrature), point location by solid angles and the Danielsson
// closest-point projection of Head2EEGMat.
//
// Reference behaviour (OpenMEEG 2.5.16):
//   SourceMatrix<Dipole>          assembleSourceMat.cpp:66-104
//   operatorPotential             operators.cpp:41-60
//   operatorPotentialDerivative   operators.cpp:63-88
//   PotentialDerivative           analytics.h:132-158
//   adaptive integration          integrator.h:44-93
//   Geometry::domain(point)       geometry.cpp:146-154, domain.cpp:5-10,
//                                 interface.cpp:18-45, mesh.cpp:239-244
//   Head2EEGMat                   assembleSensors.cpp:25-39
//   dpc / dist_point_*            danielsson.cpp:26-129
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <stack>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "bem_device.hpp"
#include "common.hpp"
#include "device_geometry.hpp"
#include "sources.hpp"
#include "adaptive.hpp"

namespace omh {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kK = 1.0 / (4.0 * kPi);
constexpr int MAX_FRAMES = ADAPT_MAX_FRAMES;

__device__ __forceinline__ int64_t cdiv_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The two integrands of the dipole right-hand side.
struct PotentialFn {            // Dipole::potential (dipole.h:29-35)
    d3 r0, q;
    __device__ __forceinline__ void operator()(d3 r, double* out) const { out[0] = dip_potential(r0, q, r); }
};

struct DerivativeFn {           // PotentialDerivative (analytics.h:140-150)
    d3 r0, q, n, p0, e10, e20;
    double d11, d12, d22;
    __device__ void init(const TriRec& t, d3 pos, d3 mom) {
        r0 = pos; q = mom;
        n = ld3(t.n);
        p0 = ld3(t.p[0]);
        e10 = sub(ld3(t.p[1]), p0);
        e20 = sub(ld3(t.p[2]), p0);
        const double c11 = dot(e10, e10), c12 = dot(e10, e20), c22 = dot(e20, e20);
        const double inv = 1.0 / (c11 * c22 - c12 * c12);   // BarycentricCoordinates.h:33-51
        d11 = c22 * inv; d12 = -c12 * inv; d22 = c11 * inv;
    }
    __device__ __forceinline__ void operator()(d3 r, double* out) const {
        const d3 v = sub(r, p0);                            // BarycentricCoordinates.h:53-69
        const double a = dot(v, e10), b = dot(v, e20);
        const double l1 = d11 * a + d12 * b;
        const double l2 = d12 * a + d22 * b;
        const double l0 = 1.0 - l1 - l2;
        const double g = dip_gradient_n(r0, q, r, n);
        out[0] = g * l0; out[1] = g * l1; out[2] = g * l2;
    }
};

// One (triangle, dipole) pair of one mesh: the P1 derivative integral into
// W[(k*nd + s)*nt + t] and, for non-barrier meshes, the P0 potential
// integral added to out(row(t), s).
__global__ void __launch_bounds__(256) k_dsm_pairs(const TriRec* __restrict__ tri, int64_t t0, int64_t nt,
                                                   const double* __restrict__ dip, int64_t s0, int64_t nd,
                                                   const double* __restrict__ cD, const double* __restrict__ cP,
                                                   int64_t nm, int mesh, int with_pot, int order, int levels,
                                                   double tol, double* __restrict__ W, double* __restrict__ out,
                                                   int64_t ld_out, Layout L, int64_t tri_unk0) {
    const int64_t nti = cdiv_dev(nt, 64);
    const int64_t ts = blockIdx.x / nti, tb = blockIdx.x - ts * nti;
    const int64_t t = tb * 64 + threadIdx.x % 64;
    const int64_t s = ts * 4 + threadIdx.x / 64;
    if (t >= nt || s >= nd) return;
    const int64_t sg = s0 + s;
    const double coefD = cD[sg * nm + mesh];
    double v3v[3] = {0.0, 0.0, 0.0};
    if (coefD != 0.0) {
        const TriRec& T = tri[t0 + t];
        const d3 P[3] = {ld3(T.p[0]), ld3(T.p[1]), ld3(T.p[2])};
        const d3 pos = ld3(dip + 6 * sg), mom = ld3(dip + 6 * sg + 3);
        DerivativeFn fd;
        fd.init(T, pos, mom);
        integrate<3>(fd, order, levels, tol, P, T.area, v3v);
        if (with_pot) {
            PotentialFn fp{pos, mom};
            double pv;
            integrate<1>(fp, order, levels, tol, P, T.area, &pv);
            const int64_t r = dev_index(L, tri_unk0 + t);
            out[r + sg * ld_out] += pv * cP[sg * nm + mesh];
        }
    }
    W[(0 * nd + s) * nt + t] = v3v[0];
    W[(1 * nd + s) * nt + t] = v3v[1];
    W[(2 * nd + s) * nt + t] = v3v[2];
}

// out(row(v), s) += sum over the triangles of v (mesh order) of W * coeffD.
__global__ void __launch_bounds__(256) k_dsm_gather(const double* __restrict__ W, int64_t t0, int64_t nt,
                                                    int64_t s0, int64_t nd, const double* __restrict__ cD,
                                                    int64_t nm, int mesh, const int64_t* __restrict__ adj_off,
                                                    const AdjRec* __restrict__ adj, const int32_t* __restrict__ slot_vid,
                                                    const int64_t* __restrict__ vert_unk, int64_t sl0, int64_t nsl,
                                                    double* __restrict__ out, int64_t ld_out, Layout L) {
    const int64_t nvi = cdiv_dev(nsl, 64);
    const int64_t ts = blockIdx.x / nvi, tb = blockIdx.x - ts * nvi;
    const int64_t v = tb * 64 + threadIdx.x % 64;
    const int64_t s = ts * 4 + threadIdx.x / 64;
    if (v >= nsl || s >= nd) return;
    const int64_t sg = s0 + s;
    const double coefD = cD[sg * nm + mesh];
    if (coefD == 0.0) return;
    const int64_t r = dev_index(L, vert_unk[slot_vid[sl0 + v]]);
    double acc = out[r + sg * ld_out];
    for (int64_t p = adj_off[sl0 + v]; p < adj_off[sl0 + v + 1]; ++p) {
        const AdjRec e = adj[p];
        acc += W[((int64_t)e.k * nd + s) * nt + (e.tri - t0)] * coefD;
    }
    out[r + sg * ld_out] = acc;
}

// Solid angle of mesh m seen from point s: block reduction over triangles.
__global__ void __launch_bounds__(256) k_mesh_solid_angle(const TriRec* __restrict__ tri, const int64_t* __restrict__ mt,
                                                          int nm, const double* __restrict__ pts, int64_t stride,
                                                          double* __restrict__ out) {
    const int64_t s = blockIdx.x / nm;
    const int m = (int)(blockIdx.x - s * nm);
    const d3 x = ld3(pts + stride * s);
    double acc = 0.0;
    for (int64_t t = mt[m] + threadIdx.x; t < mt[m + 1]; t += blockDim.x) {
        const TriRec& T = tri[t];
        const d3 r0 = sub(ld3(T.p[0]), x), r1 = sub(ld3(T.p[1]), x), r2 = sub(ld3(T.p[2]), x);
        const double d0 = norm(r0), d1 = norm(r1), d2 = norm(r2);
        const double det = 2 * T.area * dot(r0, ld3(T.n));
        if (fabs(det) >= 1e-10)
            acc += 2 * atan2(det, d0 * d1 * d2 + (d0 * dot(r1, r2) + d1 * dot(r2, r0) + d2 * dot(r0, r1)));
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s * nm + m] = red[0];
}

// dpc of danielsson.cpp:26-74 (recursive form unrolled: at most 3 levels).
__device__ double dpc(d3 p, const d3 Q[3], double al[3]) {
    int idx[3] = {0, 1, 2};
    int nb = 3;
    while (true) {
        if (nb == 1) {
            al[idx[0]] = 1.0;
            return norm(sub(p, Q[idx[0]]));
        }
        d3 A0Ai[3];
        for (int i = 1; i < nb; ++i) A0Ai[i] = sub(Q[idx[i]], Q[idx[0]]);
        const d3 A0M = sub(p, Q[idx[0]]);
        if (nb == 2) {
            al[idx[1]] = dot(A0M, A0Ai[1]) / dot(A0Ai[1], A0Ai[1]);
            al[idx[0]] = 1.0 - al[idx[1]];
        } else {
            const double a00 = dot(A0Ai[1], A0Ai[1]), a10 = dot(A0Ai[1], A0Ai[2]), a11 = dot(A0Ai[2], A0Ai[2]);
            const double b0 = dot(A0M, A0Ai[1]), b1 = dot(A0M, A0Ai[2]);
            const double d = a00 * a11 - a10 * a10;
            al[idx[1]] = (b0 * a11 - b1 * a10) / d;
            al[idx[2]] = (a00 * b1 - a10 * b0) / d;
            al[idx[0]] = 1.0 - al[idx[1]] - al[idx[2]];
        }
        int neg = -1;
        for (int i = 0; i < nb; ++i)
            if (al[idx[i]] < 0) { neg = i; break; }
        if (neg >= 0) {
            al[idx[neg]] = 0;
            const int tmp = idx[neg]; idx[neg] = idx[nb - 1]; idx[nb - 1] = tmp;
            --nb;
            continue;
        }
        d3 MH = scl(-1.0, A0M);
        for (int i = 1; i < nb; ++i) MH = add(MH, scl(al[idx[i]], A0Ai[i]));
        return norm(MH);
    }
}

// Closest triangle of mesh m to electrode e: the first strict minimum in
// mesh order (dist_point_interface, danielsson.cpp:85-102).
__global__ void __launch_bounds__(256) k_closest(const TriRec* __restrict__ tri, const int64_t* __restrict__ mt,
                                                 int nm, const double* __restrict__ el, double* __restrict__ dist,
                                                 int64_t* __restrict__ tri_out, double* __restrict__ alpha_out) {
    const int64_t e = blockIdx.x / nm;
    const int m = (int)(blockIdx.x - e * nm);
    const d3 p = ld3(el + 3 * e);
    double best = DBL_MAX;
    int64_t bt = INT64_MAX;
    double ba[3] = {0, 0, 0};
    for (int64_t t = mt[m] + threadIdx.x; t < mt[m + 1]; t += blockDim.x) {
        const TriRec& T = tri[t];
        const d3 Q[3] = {ld3(T.p[0]), ld3(T.p[1]), ld3(T.p[2])};
        double al[3] = {0, 0, 0};
        const double d = dpc(p, Q, al);
        if (d < best) { best = d; bt = t; ba[0] = al[0]; ba[1] = al[1]; ba[2] = al[2]; }
    }
    __shared__ double sd[256];
    __shared__ int64_t stt[256];
    __shared__ int sid[256];
    sd[threadIdx.x] = best;
    stt[threadIdx.x] = bt;
    sid[threadIdx.x] = threadIdx.x;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const int o = threadIdx.x + w;
            if (sd[o] < sd[threadIdx.x] || (sd[o] == sd[threadIdx.x] && stt[o] < stt[threadIdx.x])) {
                sd[threadIdx.x] = sd[o];
                stt[threadIdx.x] = stt[o];
                sid[threadIdx.x] = sid[o];
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x == sid[0]) {
        dist[e * nm + m] = best;
        tri_out[e * nm + m] = bt;
        alpha_out[3 * (e * nm + m) + 0] = ba[0];
        alpha_out[3 * (e * nm + m) + 1] = ba[1];
        alpha_out[3 * (e * nm + m) + 2] = ba[2];
    }
}

// Adjoint of the dipole right-hand side with respect to the dipole
// parameters.  For one (triangle t of a mesh, dipole s) pair, the
// non-adaptive rule integrates the derivatives of the two integrands
// (assembleSourceMat.cpp:84-99 with levels = 0),
//   V(r)   = q.x/|x|^3,                    x = r - r0   (dipole.h:29-35)
//   g(r)   = (q.n)/|x|^3 - 3 (q.x)(x.n)/|x|^5           (dipole.h:38-46)
// with respect to q and r0, weights them with the adjoint field W (the
// potential row of t and the three vertex rows, times the right-hand-side
// coefficients), and reduces the six sums over the triangles of the wave
// into grad[s*6 + 0..2] (d/dr0) and grad[s*6 + 3..5] (d/dq).
__global__ void __launch_bounds__(256) k_dsm_adjoint(const TriRec* __restrict__ tri, const int32_t* __restrict__ tri_vid,
                                                     const int64_t* __restrict__ vert_unk, int64_t t0, int64_t nt,
                                                     const double* __restrict__ dip, int64_t nd,
                                                     const double* __restrict__ cD, const double* __restrict__ cP,
                                                     int64_t nm, int mesh, int with_pot, int order,
                                                     const double* __restrict__ W, int64_t ldw, Layout L,
                                                     int64_t tri_unk0, double* __restrict__ grad) {
    const int64_t nti = cdiv_dev(nt, 64);
    const int64_t ts = blockIdx.x / nti, tb = blockIdx.x - ts * nti;
    const int64_t t = tb * 64 + threadIdx.x % 64;
    const int64_t s = ts * 4 + threadIdx.x / 64;
    double g[6] = {0, 0, 0, 0, 0, 0};
    if (t < nt && s < nd) {
        const double coefD = cD[s * nm + mesh], coefP = with_pot ? cP[s * nm + mesh] : 0.0;
        if (coefD != 0.0 || coefP != 0.0) {
            const TriRec& T = tri[t0 + t];
            const d3 P[3] = {ld3(T.p[0]), ld3(T.p[1]), ld3(T.p[2])};
            const d3 n = ld3(T.n);
            const d3 r0 = ld3(dip + 6 * s), q = ld3(dip + 6 * s + 3);
            const double wp = with_pot ? W[dev_index(L, tri_unk0 + t) + s * ldw] * coefP : 0.0;
            double wv[3];
            for (int k = 0; k < 3; ++k) {
                const int64_t u = vert_unk[tri_vid[3 * (t0 + t) + k]];
                wv[k] = u < 0 ? 0.0 : W[dev_index(L, u) + s * ldw] * coefD;
            }
            // barycentric coordinates (BarycentricCoordinates.h:33-69)
            const d3 e10 = sub(P[1], P[0]), e20 = sub(P[2], P[0]);
            const double c11 = dot(e10, e10), c12 = dot(e10, e20), c22 = dot(e20, e20);
            const double inv = 1.0 / (c11 * c22 - c12 * c12);
            const double d11 = c22 * inv, d12 = -c12 * inv, d22 = c11 * inv;
            const int np = c_nb_pts[order];
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int i = 0; i < np; ++i) {
                const d3 r = gauss_point(order, i, P);
                const d3 x = sub(r, r0);
                const double i2 = 1.0 / dot(x, x);
                const double i3 = i2 * sqrt(i2), i5 = i3 * i2, i7 = i5 * i2;
                const double qx = dot(q, x), xn = dot(x, n), qn = dot(q, n);
                const d3 v = sub(r, P[0]);
                const double a = dot(v, e10), b = dot(v, e20);
                const double l1 = d11 * a + d12 * b, l2 = d12 * a + d22 * b, l0 = 1.0 - l1 - l2;
                const double wsum = wv[0] * l0 + wv[1] * l1 + wv[2] * l2;   // P1-weighted adjoint
                // dV/dq = x/|x|^3 ; dV/dr0 = -(q/|x|^3 - 3 (q.x) x/|x|^5)
                // dg/dq = n/|x|^3 - 3 (x.n) x/|x|^5
                // dg/dr0 = 3 (q.n) x/|x|^5 + 3 ((x.n) q + (q.x) n)/|x|^5 - 15 (q.x)(x.n) x/|x|^7
                const double cx = 3 * qx * i5;
                const double cg = 3 * qn * i5 - 15 * qx * xn * i7;
                const double w = c_rules[order][i][3];
                const double xs[3] = {x.x, x.y, x.z}, qs[3] = {q.x, q.y, q.z}, ns[3] = {n.x, n.y, n.z};
                for (int c = 0; c < 3; ++c) {
                    const double dV_dr0 = -(qs[c] * i3 - cx * xs[c]);
                    const double dg_dr0 = cg * xs[c] + 3 * (xn * qs[c] + qx * ns[c]) * i5;
                    const double dV_dq = xs[c] * i3;
                    const double dg_dq = ns[c] * i3 - 3 * xn * xs[c] * i5;
                    acc[c] += w * (wp * dV_dr0 + wsum * dg_dr0);
                    acc[3 + c] += w * (wp * dV_dq + wsum * dg_dq);
                }
            }
            for (int c = 0; c < 6; ++c) g[c] = T.area * acc[c];
        }
    }
    // wave reduction (one dipole per wave), then one atomic per component
    for (int c = 0; c < 6; ++c) {
        double v = g[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        g[c] = v;
    }
    if ((threadIdx.x & 63) == 0 && s < nd)
        for (int c = 0; c < 6; ++c) atomicAdd(&grad[s * 6 + c], g[c]);
}

// MagFactor (constants.h:22): mu0 / (4 pi).
constexpr double MAG_FACTOR = 1e-7;

// Head2MEGMat before the sensor weights (assembleSensors.cpp:69-99) for one
// mesh: assemble_ferguson (assembleFerguson.cpp:18-30) with
// Details::operatorFerguson (operators.h:40-59),
//   F(x, v) = sum over the mesh's triangles T around v of (B - A) S_T(x) / (2 |T|)
// (A, B the edge of T opposite v, triangle.h:87-90,154; S_T = OperatorS),
// out(i, v) += coeff F(x_i, v) . d_i / |d_i|.  Thread = vertex slot
// (x), integration point (block y, looped).
__global__ void __launch_bounds__(256) k_ferguson(const TriRec* __restrict__ tri, const int64_t* __restrict__ adj_off,
                                                  const AdjRec* __restrict__ adj, const int32_t* __restrict__ slot_vid,
                                                  const int64_t* __restrict__ vert_unk, int64_t s0, int64_t ns,
                                                  const double* __restrict__ pts, const double* __restrict__ dirs,
                                                  int64_t npts, double coeff, double* __restrict__ out, int64_t ld,
                                                  Layout L) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int64_t vi = vert_unk[slot_vid[s0 + s]];
    if (vi < 0) return;
    const int64_t col = dev_index(L, vi);
    for (int64_t i = blockIdx.y; i < npts; i += gridDim.y) {
        const d3 x = ld3(pts + 3 * i), d = ld3(dirs + 3 * i);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int64_t p = adj_off[s0 + s]; p < adj_off[s0 + s + 1]; ++p) {
            const AdjRec e = adj[p];
            const TriRec& T = tri[e.tri];
            const double sv = ops_eval(load_ops(T), x) / (2 * T.area);
            fx -= e.cb[0] * sv;   // cb = A - B
            fy -= e.cb[1] * sv;
            fz -= e.cb[2] * sv;
        }
        out[i + col * ld] += coeff * (fx * d.x + fy * d.y + fz * d.z) / norm(d);
    }
}

// DipSource2MEGMat before the sensor weights (assembleSensors.cpp:133-156):
// out(i, j) = MagFactor (q_j x (x_i - r_j)) . d_i / (|x_i - r_j|^3 |d_i|).
// Thread = integration point (x), dipole (block y, looped).
__global__ void __launch_bounds__(256) k_dip2meg(const double* __restrict__ dip, int64_t n_dip,
                                                 const double* __restrict__ pts, const double* __restrict__ dirs,
                                                 int64_t npts, double* __restrict__ out, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const d3 x = ld3(pts + 3 * i), d = ld3(dirs + 3 * i);
    for (int64_t j = blockIdx.y; j < n_dip; j += gridDim.y) {
        const d3 r = ld3(dip + 6 * j), q = ld3(dip + 6 * j + 3);
        const d3 diff = sub(x, r);
        const double nd = norm(diff);
        const d3 f = scl(1.0 / (nd * nd * nd), cross(q, diff));
        out[i + j * ld] = dot(f, d) * MAG_FACTOR / norm(d);
    }
}

// Surf2VolMat (assembleHeadMat.cpp:325-363, PartialBlock operators.h:302-
// 331) for one mesh, over a chunk of the kept points [p0, p0 + pc).  Kept
// point r sits at pts[3 kept[r]]; cD[r] (cS[r]) is its coefficient for
// this mesh's vertex (triangle) columns, zero when the mesh does not bound
// the point's domain.  Threads run along the points, so that the writes to
// out (points x unknowns, column-major) are coalesced.
// Pass 1: W[(k nt + t) pc + p] = P1SingleLayerPotential_t(x_p)[k], one
// evaluation per (triangle, point); block = 256 points of one triangle.
__global__ void __launch_bounds__(256) k_surf2vol_p1(const TriRec* __restrict__ tri, int64_t t0, int64_t nt,
                                                     const double* __restrict__ pts, const int64_t* __restrict__ kept,
                                                     int64_t p0, int64_t pc, double* __restrict__ W) {
    const int64_t nb = cdiv_dev(pc, 256);
    const int64_t t = blockIdx.x / nb, p = (blockIdx.x - t * nb) * 256 + threadIdx.x;
    if (p >= pc) return;
    const d3 v = p1s_eval(load_p1s(tri[t0 + t]), ld3(pts + 3 * kept[p0 + p]));
    W[(0 * nt + t) * pc + p] = v.x;
    W[(1 * nt + t) * pc + p] = v.y;
    W[(2 * nt + t) * pc + p] = v.z;
}

// Pass 2, thread = point (x), vertex slot of the mesh (block y, looped):
// out(r, v) += cD[r] sum over the triangles T of v (mesh order) of W(T, r)
// at v's position in T.
__global__ void __launch_bounds__(256) k_surf2vol_d(const int64_t* __restrict__ adj_off, const AdjRec* __restrict__ adj,
                                                    const int32_t* __restrict__ slot_vid,
                                                    const int64_t* __restrict__ vert_unk, int64_t s0, int64_t ns,
                                                    int64_t t0, int64_t nt, const double* __restrict__ W, int64_t p0,
                                                    int64_t pc, const double* __restrict__ cD, double* __restrict__ out,
                                                    Layout L) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pc) return;
    const double c = cD[p0 + p];
    if (c == 0.0) return;
    for (int64_t s = blockIdx.y; s < ns; s += gridDim.y) {
        const int64_t vi = vert_unk[slot_vid[s0 + s]];
        if (vi < 0) continue;
        double acc = 0.0;
        for (int64_t q = adj_off[s0 + s]; q < adj_off[s0 + s + 1]; ++q) {
            const AdjRec e = adj[q];
            acc += W[((int64_t)e.k * nt + (e.tri - t0)) * pc + p];
        }
        out[(p0 + p) + dev_index(L, vi) * L.ld] += c * acc;
    }
}

// S part, thread = point (x), triangle of the mesh (block y, looped):
// out(r, t) = cS[r] OperatorS_t(x_r) (the reference assigns these entries).
__global__ void __launch_bounds__(256) k_surf2vol_s(const TriRec* __restrict__ tri, int64_t t0, int64_t nt,
                                                    int64_t unk0, const double* __restrict__ pts,
                                                    const int64_t* __restrict__ kept, int64_t nk,
                                                    const double* __restrict__ cS, double* __restrict__ out, Layout L) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nk) return;
    const double c = cS[r];
    if (c == 0.0) return;
    const d3 x = ld3(pts + 3 * kept[r]);
    for (int64_t t = blockIdx.y; t < nt; t += gridDim.y)
        out[r + dev_index(L, unk0 + t) * L.ld] = c * ops_eval(load_ops(tri[t0 + t]), x);
}

// DipSource2InternalPotMat (assembleSourceMat.cpp:154-186): out(r, j) =
// coef[j] q_j . (x - r_j) / |x - r_j|^3 when kept point r lies in dipole
// j's domain (pdom[r] == ddom[j]), else 0.  Thread = kept point (x),
// dipole (block y, looped).
__global__ void __launch_bounds__(256) k_dip2ip(const double* __restrict__ dip, int64_t n_dip,
                                                const int32_t* __restrict__ ddom, const double* __restrict__ coef,
                                                const double* __restrict__ pts, const int64_t* __restrict__ kept,
                                                const int32_t* __restrict__ pdom, int64_t nk, double* __restrict__ out,
                                                int64_t ld) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nk) return;
    const d3 x = ld3(pts + 3 * kept[r]);
    for (int64_t j = blockIdx.y; j < n_dip; j += gridDim.y) {
        double v = 0.0;
        if (pdom[r] == ddom[j]) {
            const d3 diff = sub(x, ld3(dip + 6 * j));
            const double nd = norm(diff);
            v = coef[j] * dot(ld3(dip + 6 * j + 3), diff) / (nd * nd * nd);
        }
        out[r + j * ld] = v;
    }
}

// EITSourceMat (assembleSourceMat.cpp:109-152) for one injection triangle t
// of a current-barrier mesh and one mesh pair (t's mesh, m2): column `out`
// (one electrode) += w x row t of the transfer matrix.
// D and identity part, thread = vertex slot of m2: cD x D(t, v) (the P1
// double layer of v's triangles integrated over t by Gauss, as the D
// blocks) plus cI x |t| / 3 when v is a vertex of t (addIdentity,
// operators.h:200-206,267-269).
__global__ void __launch_bounds__(256) k_eit_d(const TriRec* __restrict__ tri, int64_t t, const int32_t* __restrict__ tvid,
                                               const int64_t* __restrict__ adj_off, const AdjRec* __restrict__ adj,
                                               const int32_t* __restrict__ slot_vid, const int64_t* __restrict__ vert_unk,
                                               int64_t s0, int64_t ns, int order, double cD, double cI,
                                               double* __restrict__ out, Layout L) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int32_t vid = slot_vid[s0 + s];
    const int64_t vi = vert_unk[vid];
    if (vi < 0) return;
    const TriRec& T = tri[t];
    const d3 P[3] = {ld3(T.p[0]), ld3(T.p[1]), ld3(T.p[2])};
    const int np = c_nb_pts[order];
    double acc = 0.0;
    for (int64_t p = adj_off[s0 + s]; p < adj_off[s0 + s + 1]; ++p) {
        const AdjRec e = adj[p];
        const P1S o = load_p1s(tri[e.tri]);
        double r = 0.0;
        for (int k = 0; k < np; ++k) {
            const d3 v = p1s_eval(o, gauss_point(order, k, P));
            r += c_rules[order][k][3] * (e.k == 0 ? v.x : (e.k == 1 ? v.y : v.z));
        }
        acc += T.area * r;
    }
    double val = cD * acc;
    if (cI != 0.0 && (vid == tvid[3 * t] || vid == tvid[3 * t + 1] || vid == tvid[3 * t + 2])) val += cI * T.area / 3.0;
    out[dev_index(L, vi)] += val;
}

// S part, thread = triangle u of m2: cS x S(t, u) (OperatorS of t integrated
// over u by Gauss, as the S blocks).
__global__ void __launch_bounds__(256) k_eit_s(const TriRec* __restrict__ tri, int64_t t, int64_t u0, int64_t nu,
                                               int64_t unk0, int order, double cS, double* __restrict__ out, Layout L) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const TriRec& U = tri[u0 + u];
    const OpS o = load_ops(tri[t]);
    const d3 P[3] = {ld3(U.p[0]), ld3(U.p[1]), ld3(U.p[2])};
    const int np = c_nb_pts[order];
    double r = 0.0;
    for (int k = 0; k < np; ++k) r += c_rules[order][k][3] * ops_eval(o, gauss_point(order, k, P));
    out[dev_index(L, unk0 + u)] += cS * (U.area * r);
}

}  // namespace

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static std::vector<double> mesh_solid_angles(Geometry& g, const double* d_pts, int64_t stride, int64_t npts,
                                             hipStream_t st) {
    DeviceGeometry& dg = device_geometry(g);
    const int nm = (int)g.meshes.size();
    std::vector<int64_t> mt(nm + 1);
    for (int m = 0; m < nm; ++m) mt[m] = dg.meshes[m].tri_begin;
    mt[nm] = dg.n_tris;
    // the source-matrix slot (not SCR_SMALL): a source matrix on a side stream
    // then shares no slot with a factorisation or solve on the caller's
    int64_t* d_mt = (int64_t*)scratch(SCR_SRCCOEF, sizeof(int64_t) * (nm + 1) + sizeof(double) * npts * nm, st);
    double* d_sa = (double*)(d_mt + nm + 1);
    OMH_HIP(hipMemcpyAsync(d_mt, mt.data(), sizeof(int64_t) * (nm + 1), hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));
    k_mesh_solid_angle<<<(unsigned)(npts * nm), 256, 0, st>>>(dg.tri, d_mt, nm, d_pts, stride, d_sa);
    OMH_LAUNCH_CHECK();
    std::vector<double> sa((size_t)npts * nm);
    OMH_HIP(hipMemcpyAsync(sa.data(), d_sa, sizeof(double) * sa.size(), hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    return sa;
}

std::vector<int> locate_domains(Geometry& g, const double* d_pts, int64_t stride, int64_t npts, hipStream_t st) {
    const int nm = (int)g.meshes.size();
    const std::vector<double> sa = mesh_solid_angles(g, d_pts, stride, npts, st);
    std::vector<int> dom(npts, -1);
    for (int64_t s = 0; s < npts; ++s) {
        for (size_t k = 0; k < g.domains.size() && dom[s] < 0; ++k) {
            bool inside = true;
            for (const auto& b : g.domains[k].boundaries) {
                double a = 0.0;
                for (const auto& om : b.first.omeshes) a += om.second * sa[s * nm + om.first];
                inside = inside && (g.interface_contains_sa(a) == b.second);
            }
            if (inside) dom[s] = (int)k;
        }
        if (dom[s] < 0) throw std::runtime_error("a dipole lies in no domain");
    }
    return dom;
}

// Per (dipole, mesh) coefficients of assembleSourceMat.cpp:84-99: cD of the
// P1 derivative rows, cP of the P0 potential rows (zero on current barriers).
void dsm_coefficients(const Geometry& g, int64_t n_dip, const std::vector<int>& domain, std::vector<double>& cD,
                             std::vector<double>& cP, std::vector<char>& used) {
    const int nm = (int)g.meshes.size();
    cD.assign((size_t)n_dip * nm, 0.0);
    cP.assign((size_t)n_dip * nm, 0.0);
    used.assign(nm, 0);
    for (int64_t s = 0; s < n_dip; ++s) {
        const Domain& d = g.domains.at(domain[s]);
        if (d.cond == 0.0) continue;
        for (const auto& b : d.boundaries) {
            const double fD = b.second ? -kK : kK;
            for (const auto& om : b.first.omeshes) {
                const double c = fD * om.second;
                cD[s * nm + om.first] += c;
                if (!g.meshes[om.first].current_barrier) cP[s * nm + om.first] += c / d.cond;
                used[om.first] = 1;
            }
        }
    }
}

void dsm_launch(Geometry& g, const double* d_dip, int64_t n_dip, const double* d_cD, const double* d_cP,
                const std::vector<char>& used, int order, int levels, double tol, double* d_out, int64_t ld_out,
                const Layout& L, hipStream_t st) {
    if (order < 0 || order > 3) throw std::invalid_argument("Gauss order must be 0..3");
    if (levels < 0 || levels >= MAX_FRAMES) throw std::invalid_argument("adaptive levels must be 0..11");
    DeviceGeometry& dg = device_geometry(g);
    const int64_t n = g.size();
    const int nm = (int)g.meshes.size();
    fill_f64(d_out, n, n_dip, ld_out, 0.0, st);
    if (n_dip == 0) return;
    for (int m = 0; m < nm; ++m) {
        if (!used[m]) continue;
        const MeshDev& md = dg.meshes[m];
        const int64_t nt = md.tri_end - md.tri_begin, nsl = md.slot_end - md.slot_begin;
        const size_t budget = size_t(1) << 29;
        int64_t chunk = std::max<int64_t>(4, (int64_t)(budget / (3 * sizeof(double) * (size_t)nt)));
        chunk = std::min(chunk, n_dip);
        double* W = (double*)scratch(SCR_SOURCES, 3 * sizeof(double) * (size_t)nt * (size_t)chunk, st);
        for (int64_t s0 = 0; s0 < n_dip; s0 += chunk) {
            const int64_t nd = std::min(chunk, n_dip - s0);
            const int64_t b1 = cdiv(nt, 64) * cdiv(nd, 4);
            k_dsm_pairs<<<(unsigned)b1, 256, 0, st>>>(dg.tri, md.tri_begin, nt, d_dip, s0, nd, d_cD, d_cP, nm, m,
                                                      md.barrier ? 0 : 1, order, levels, tol, W, d_out, ld_out, L,
                                                      md.tri_unk0);
            OMH_LAUNCH_CHECK();
            const int64_t b2 = cdiv(nsl, 64) * cdiv(nd, 4);
            k_dsm_gather<<<(unsigned)b2, 256, 0, st>>>(W, md.tri_begin, nt, s0, nd, d_cD, nm, m, dg.adj_off, dg.adj,
                                                       dg.slot_vid, dg.vert_unk, md.slot_begin, nsl, d_out, ld_out, L);
            OMH_LAUNCH_CHECK();
        }
    }
}

void dip_source_mat(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain, int order,
                    int levels, double tol, double* d_out, int64_t ld_out, const Layout& L, hipStream_t st) {
    std::vector<double> cD, cP;
    std::vector<char> used;
    dsm_coefficients(g, n_dip, domain, cD, cP, used);
    double* d_c = (double*)scratch(SCR_SRCCOEF, sizeof(double) * std::max<size_t>(1, 2 * cD.size()), st);
    if (!cD.empty()) {
        OMH_HIP(hipMemcpyAsync(d_c, cD.data(), sizeof(double) * cD.size(), hipMemcpyHostToDevice, st));
        OMH_HIP(hipMemcpyAsync(d_c + cD.size(), cP.data(), sizeof(double) * cP.size(), hipMemcpyHostToDevice, st));
        OMH_HIP(hipStreamSynchronize(st));   // the host vectors die with this frame
    }
    dsm_launch(g, d_dip, n_dip, d_c, d_c + cD.size(), used, order, levels, tol, d_out, ld_out, L, st);
}

void dip_source_adjoint(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain, int order,
                        const double* d_W, int64_t ldw, const Layout& L, double* d_grad, hipStream_t st) {
    if (order < 0 || order > 3) throw std::invalid_argument("Gauss order must be 0..3");
    DeviceGeometry& dg = device_geometry(g);
    const int nm = (int)g.meshes.size();
    fill_f64(d_grad, 6, n_dip, 6, 0.0, st);
    if (n_dip == 0) return;
    std::vector<double> cD, cP;
    std::vector<char> used;
    dsm_coefficients(g, n_dip, domain, cD, cP, used);
    double* d_c = (double*)scratch(SCR_SMALL, sizeof(double) * 2 * cD.size(), st);
    OMH_HIP(hipMemcpyAsync(d_c, cD.data(), sizeof(double) * cD.size(), hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_c + cD.size(), cP.data(), sizeof(double) * cP.size(), hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));
    for (int m = 0; m < nm; ++m) {
        if (!used[m]) continue;
        const MeshDev& md = dg.meshes[m];
        const int64_t nt = md.tri_end - md.tri_begin;
        const int64_t blocks = cdiv(nt, 64) * cdiv(n_dip, 4);
        k_dsm_adjoint<<<(unsigned)blocks, 256, 0, st>>>(dg.tri, dg.tri_vid, dg.vert_unk, md.tri_begin, nt, d_dip, n_dip,
                                                        d_c, d_c + cD.size(), nm, m, md.barrier ? 0 : 1, order, d_W,
                                                        ldw, L, md.tri_unk0, d_grad);
        OMH_LAUNCH_CHECK();
    }
}

// dist_point_geom (danielsson.cpp:107-129) for n device points: the
// closest triangle (global id) on the interfaces of zero-conductivity
// domains, that interface, and `alphas` as the reference leaves them (the
// barycentric weights from the last interface scanned).
struct Closest {
    std::vector<int64_t> tri;
    std::vector<const Interface*> itf;
    std::vector<std::array<double, 3>> alphas;
};

static Closest closest_outer(Geometry& g, const double* d_pts, int64_t n, hipStream_t st) {
    DeviceGeometry& dg = device_geometry(g);
    const int nm = (int)g.meshes.size();
    std::vector<int64_t> mt(nm + 1);
    for (int m = 0; m < nm; ++m) mt[m] = dg.meshes[m].tri_begin;
    mt[nm] = dg.n_tris;
    const size_t cnt = (size_t)n * nm;
    char* buf = (char*)scratch(SCR_SMALL, sizeof(int64_t) * (nm + 1) + cnt * (sizeof(double) * 4 + sizeof(int64_t)), st);
    int64_t* d_mt = (int64_t*)buf;
    double* d_dist = (double*)(d_mt + nm + 1);
    int64_t* d_tri = (int64_t*)(d_dist + cnt);
    double* d_al = (double*)(d_tri + cnt);
    OMH_HIP(hipMemcpyAsync(d_mt, mt.data(), sizeof(int64_t) * (nm + 1), hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));
    if (n > 0) {
        k_closest<<<(unsigned)(n * nm), 256, 0, st>>>(dg.tri, d_mt, nm, d_pts, d_dist, d_tri, d_al);
        OMH_LAUNCH_CHECK();
    }
    std::vector<double> dist(cnt), al(3 * cnt);
    std::vector<int64_t> tri(cnt);
    OMH_HIP(hipMemcpyAsync(dist.data(), d_dist, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    OMH_HIP(hipMemcpyAsync(tri.data(), d_tri, sizeof(int64_t) * cnt, hipMemcpyDeviceToHost, st));
    OMH_HIP(hipMemcpyAsync(al.data(), d_al, sizeof(double) * 3 * cnt, hipMemcpyDeviceToHost, st));
    OMH_HIP(hipStreamSynchronize(st));
    Closest c;
    c.tri.assign(n, -1);
    c.itf.assign(n, nullptr);
    c.alphas.assign(n, {0, 0, 0});
    for (int64_t e = 0; e < n; ++e) {
        double distmin = DBL_MAX;
        for (const auto& d : g.domains) {
            if (d.cond != 0.0) continue;
            for (const auto& b : d.boundaries) {
                double dm = DBL_MAX;
                int64_t bt = -1;
                const double* ba = nullptr;
                for (const auto& om : b.first.omeshes) {
                    const size_t k = (size_t)e * nm + om.first;
                    if (dist[k] < dm) { dm = dist[k]; bt = tri[k]; ba = &al[3 * k]; }
                }
                if (ba) c.alphas[e] = {ba[0], ba[1], ba[2]};
                if (dm < distmin) {
                    distmin = dm;
                    c.tri[e] = bt;
                    c.itf[e] = &b.first;
                }
            }
        }
        if (c.tri[e] < 0) throw std::runtime_error("the geometry has no zero-conductivity domain");
    }
    return c;
}

void head2eeg(Geometry& g, const double* d_el, int64_t n_el, int64_t* cols, double* vals, hipStream_t st) {
    const Closest c = closest_outer(g, d_el, n_el, st);
    // global triangle id -> vertex ids
    std::vector<std::array<i64, 3>> tv;
    for (const auto& m : g.meshes) tv.insert(tv.end(), m.tris.begin(), m.tris.end());
    for (int64_t e = 0; e < n_el; ++e)
        for (int k = 0; k < 3; ++k) {
            cols[3 * e + k] = g.vert_index[tv[c.tri[e]][k]];
            vals[3 * e + k] = c.alphas[e][k];
        }
}

void head2meg(Geometry& g, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out, int64_t ld_out,
              const Layout& L, hipStream_t st) {
    DeviceGeometry& dg = device_geometry(g);
    fill_f64(d_out, npts, g.size(), ld_out, 0.0, st);
    if (npts <= 0) return;
    for (int m = 0; m < (int)g.meshes.size(); ++m) {
        // Geometry::conductivity_jump (geometry.h:188-194)
        double jump = 0.0;
        for (int d : g.domains_of(m)) jump += g.domains[d].cond * g.domains[d].mesh_orientation(m);
        const MeshDev& md = dg.meshes[m];
        const int64_t ns = md.slot_end - md.slot_begin;
        if (ns <= 0 || jump == 0.0) continue;
        const dim3 grid((unsigned)cdiv(ns, 256), (unsigned)std::min<int64_t>(npts, 65535));
        k_ferguson<<<grid, 256, 0, st>>>(dg.tri, dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk, md.slot_begin, ns,
                                         d_pts, d_dirs, npts, MAG_FACTOR * jump, d_out, ld_out, L);
        OMH_LAUNCH_CHECK();
    }
}

void surf2meg(Geometry& src, const double* d_pts, const double* d_dirs, int64_t npts, double* d_out, int64_t ld_out,
              hipStream_t st) {
    if (src.meshes.size() != 1) throw std::invalid_argument("a source mesh is one mesh");
    DeviceGeometry& ds = device_geometry(src);
    const int64_t nv = (int64_t)src.vertices.size();
    fill_f64(d_out, npts, nv, ld_out, 0.0, st);
    const MeshDev& md = ds.meshes[0];
    const int64_t ns = md.slot_end - md.slot_begin;
    if (npts <= 0 || ns <= 0) return;
    // columns: the source mesh's local vertex indices (reference order of a
    // one-mesh numbering, no triangle columns)
    Layout L;
    L.ld = ld_out;
    L.nv = nv;
    L.nt = 0;
    L.tri_first = 0;
    const dim3 grid((unsigned)cdiv(ns, 256), (unsigned)std::min<int64_t>(npts, 65535));
    k_ferguson<<<grid, 256, 0, st>>>(ds.tri, ds.adj_off, ds.adj, ds.slot_vid, ds.vert_unk, md.slot_begin, ns, d_pts,
                                     d_dirs, npts, 1.0, d_out, ld_out, L);
    OMH_LAUNCH_CHECK();
}

void dip2meg(const double* d_dip, int64_t n_dip, const double* d_pts, const double* d_dirs, int64_t npts,
             double* d_out, int64_t ld_out, hipStream_t st) {
    if (npts <= 0 || n_dip <= 0) return;
    const dim3 grid((unsigned)cdiv(npts, 256), (unsigned)std::min<int64_t>(n_dip, 65535));
    k_dip2meg<<<grid, 256, 0, st>>>(d_dip, n_dip, d_pts, d_dirs, npts, d_out, ld_out);
    OMH_LAUNCH_CHECK();
}

// Kept points (those of a domain with non-zero conductivity, in input
// order) with their domains; their indices go to the device in `slot`.
static std::vector<int64_t> kept_points(Geometry& g, const double* d_pts, int64_t npts, hipStream_t st,
                                        std::vector<int>& kdom) {
    const std::vector<int> dom = locate_domains(g, d_pts, 3, npts, st);
    std::vector<int64_t> kept;
    kdom.clear();
    for (int64_t i = 0; i < npts; ++i)
        if (g.domains[dom[i]].cond != 0.0) {
            kept.push_back(i);
            kdom.push_back(dom[i]);
        }
    return kept;
}

std::vector<int64_t> surf2vol(Geometry& g, const double* d_pts, int64_t npts, double* d_out, const Layout& L,
                              hipStream_t st) {
    std::vector<int> kdom;
    const std::vector<int64_t> kept = npts > 0 ? kept_points(g, d_pts, npts, st, kdom) : std::vector<int64_t>();
    const int64_t nk = (int64_t)kept.size(), n = g.size();
    if (nk == 0) return kept;
    fill_f64(d_out, nk, n, L.ld, 0.0, st);
    const int nm = (int)g.meshes.size();
    // coefficients per (mesh, kept point): D = -K x the boundary orientation,
    // S = K x orientation / sigma (not on a current barrier)
    std::vector<double> cD((size_t)nm * nk, 0.0), cS((size_t)nm * nk, 0.0);
    std::vector<char> used(nm, 0);
    for (int64_t r = 0; r < nk; ++r) {
        const Domain& d = g.domains[kdom[r]];
        for (const auto& b : d.boundaries)
            for (const auto& om : b.first.omeshes) {
                const double coeff = (b.second ? om.second : -om.second) * kK;
                cD[(size_t)om.first * nk + r] += -coeff;
                if (!g.meshes[om.first].current_barrier) cS[(size_t)om.first * nk + r] = coeff / d.cond;
                used[om.first] = 1;
            }
    }
    DeviceGeometry& dg = device_geometry(g);
    char* buf = (char*)scratch(SCR_SRCCOEF, sizeof(int64_t) * nk + 2 * sizeof(double) * cD.size(), st);
    int64_t* d_kept = (int64_t*)buf;
    double* d_cD = (double*)(buf + sizeof(int64_t) * nk);
    double* d_cS = d_cD + cD.size();
    OMH_HIP(hipMemcpyAsync(d_kept, kept.data(), sizeof(int64_t) * nk, hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_cD, cD.data(), sizeof(double) * cD.size(), hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_cS, cS.data(), sizeof(double) * cS.size(), hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));   // the host vectors die with this frame
    // pass-1 scratch: 3 x triangles x points of a chunk, at most 1 GiB
    int64_t ntmax = 1;
    for (int m = 0; m < nm; ++m) ntmax = std::max<int64_t>(ntmax, dg.meshes[m].tri_end - dg.meshes[m].tri_begin);
    const int64_t pc_max = std::max<int64_t>(256, std::min<int64_t>(nk, ((int64_t(1) << 30) / (24 * ntmax)) / 256 * 256));
    double* W = (double*)scratch(SCR_DTMP, sizeof(double) * 3 * (size_t)ntmax * (size_t)std::min<int64_t>(pc_max, nk), st);
    for (int m = 0; m < nm; ++m) {
        if (!used[m]) continue;
        const MeshDev& md = dg.meshes[m];
        const int64_t ns = md.slot_end - md.slot_begin, nt = md.tri_end - md.tri_begin;
        if (ns > 0 && nt > 0) {
            for (int64_t p0 = 0; p0 < nk; p0 += pc_max) {
                const int64_t pc = std::min<int64_t>(pc_max, nk - p0);
                k_surf2vol_p1<<<(unsigned)(nt * cdiv(pc, 256)), 256, 0, st>>>(dg.tri, md.tri_begin, nt, d_pts, d_kept,
                                                                               p0, pc, W);
                OMH_LAUNCH_CHECK();
                k_surf2vol_d<<<dim3((unsigned)cdiv(pc, 256), (unsigned)std::min<int64_t>(ns, 65535)), 256, 0, st>>>(
                    dg.adj_off, dg.adj, dg.slot_vid, dg.vert_unk, md.slot_begin, ns, md.tri_begin, nt, W, p0, pc,
                    d_cD + (size_t)m * nk, d_out, L);
                OMH_LAUNCH_CHECK();
            }
        }
        if (nt > 0 && !md.barrier) {
            k_surf2vol_s<<<dim3((unsigned)cdiv(nk, 256), (unsigned)std::min<int64_t>(nt, 65535)), 256, 0, st>>>(
                dg.tri, md.tri_begin, nt, md.tri_unk0, d_pts, d_kept, nk, d_cS + (size_t)m * nk, d_out, L);
            OMH_LAUNCH_CHECK();
        }
    }
    return kept;
}

std::vector<int64_t> dip2internal_pot(Geometry& g, const double* d_dip, int64_t n_dip, const std::vector<int>& domain,
                                      const double* d_pts, int64_t npts, double* d_out, int64_t ld_out,
                                      hipStream_t st) {
    std::vector<int> kdom;
    const std::vector<int64_t> kept = npts > 0 ? kept_points(g, d_pts, npts, st, kdom) : std::vector<int64_t>();
    const int64_t nk = (int64_t)kept.size();
    if (nk == 0 || n_dip <= 0) return kept;
    std::vector<int32_t> pdom(kdom.begin(), kdom.end()), ddom(n_dip);
    std::vector<double> coef(n_dip);
    for (int64_t j = 0; j < n_dip; ++j) {
        ddom[j] = domain[j];
        const double c = g.domains.at(domain[j]).cond;
        // K / sigma (a dipole in a zero-conductivity domain has no point in it)
        coef[j] = c != 0.0 ? kK / c : 0.0;
    }
    char* buf = (char*)scratch(SCR_SRCCOEF, sizeof(int64_t) * nk + sizeof(double) * n_dip +
                                              sizeof(int32_t) * (nk + n_dip), st);
    int64_t* d_kept = (int64_t*)buf;
    double* d_coef = (double*)(buf + sizeof(int64_t) * nk);
    int32_t* d_pdom = (int32_t*)(d_coef + n_dip);
    int32_t* d_ddom = d_pdom + nk;
    OMH_HIP(hipMemcpyAsync(d_kept, kept.data(), sizeof(int64_t) * nk, hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_coef, coef.data(), sizeof(double) * n_dip, hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_pdom, pdom.data(), sizeof(int32_t) * nk, hipMemcpyHostToDevice, st));
    OMH_HIP(hipMemcpyAsync(d_ddom, ddom.data(), sizeof(int32_t) * n_dip, hipMemcpyHostToDevice, st));
    OMH_HIP(hipStreamSynchronize(st));
    k_dip2ip<<<dim3((unsigned)cdiv(nk, 256), (unsigned)std::min<int64_t>(n_dip, 65535)), 256, 0, st>>>(
        d_dip, n_dip, d_ddom, d_coef, d_pts, d_kept, d_pdom, nk, d_out, ld_out);
    OMH_LAUNCH_CHECK();
    return kept;
}

// almost_equal(r, 0) of om_common.h:24-26.
static bool almost_zero(double r) { return std::fabs(r) < DBL_MIN; }

// Sensors::findInjectionTriangles (sensors.cpp:167-220): the closest
// triangle of each electrode and, when its radius is not zero and the
// triangle is smaller than 4 pi r^2, the triangles of the same interface
// reached through shared edges whose centres lie within r; the weight
// pi r^2 / (their area) (1 when r = 0).
static void injection_triangles(Geometry& g, const double* d_pos, const double* radii, int64_t ne, hipStream_t st,
                                std::vector<std::vector<int64_t>>& tris, std::vector<double>& weight) {
    const Closest c = closest_outer(g, d_pos, ne, st);
    DeviceGeometry& dg = device_geometry(g);
    std::vector<std::array<i64, 3>> tv;
    for (const auto& m : g.meshes) tv.insert(tv.end(), m.tris.begin(), m.tris.end());
    auto area = [&](int64_t t) {
        const auto& a = g.vertices[tv[t][0]];
        const auto& b = g.vertices[tv[t][1]];
        const auto& q = g.vertices[tv[t][2]];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
        const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
        return 0.5 * std::sqrt(x * x + y * y + z * z);
    };
    std::vector<double> pos(3 * ne);
    OMH_HIP(hipMemcpy(pos.data(), d_pos, sizeof(double) * 3 * ne, hipMemcpyDeviceToHost));
    tris.assign(ne, {});
    weight.assign(ne, 1.0);
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t t0 = c.tri[e];
        tris[e].push_back(t0);
        const double r = radii ? radii[e] : 0.0;
        if (almost_zero(r)) continue;
        if (area(t0) < 4 * kPi * r * r) {
            // edge -> triangles of the interface (Mesh::adjacent_triangles over its meshes)
            std::map<std::pair<i64, i64>, std::vector<int64_t>> edges;
            for (const auto& om : c.itf[e]->omeshes)
                for (int64_t t = dg.meshes[om.first].tri_begin; t < dg.meshes[om.first].tri_end; ++t)
                    for (int k = 0; k < 3; ++k) {
                        const i64 a = tv[t][k], b = tv[t][(k + 1) % 3];
                        edges[{std::min(a, b), std::max(a, b)}].push_back(t);
                    }
            std::set<int64_t> seen{t0};
            std::stack<int64_t> stk;
            stk.push(t0);
            const double px = pos[3 * e], py = pos[3 * e + 1], pz = pos[3 * e + 2];
            while (!stk.empty()) {
                const int64_t t = stk.top();
                stk.pop();
                double cx = 0, cy = 0, cz = 0;
                for (int k = 0; k < 3; ++k) {
                    cx += g.vertices[tv[t][k]][0] / 3;
                    cy += g.vertices[tv[t][k]][1] / 3;
                    cz += g.vertices[tv[t][k]][2] / 3;
                }
                if (std::sqrt((cx - px) * (cx - px) + (cy - py) * (cy - py) + (cz - pz) * (cz - pz)) < r) {
                    if (t != t0) tris[e].push_back(t);
                    for (int k = 0; k < 3; ++k) {
                        const i64 a = tv[t][k], b = tv[t][(k + 1) % 3];
                        for (int64_t u : edges[{std::min(a, b), std::max(a, b)}])
                            if (u != t && seen.insert(u).second) stk.push(u);
                    }
                }
            }
        }
        double sum = 0.0;
        for (int64_t t : tris[e]) sum += area(t);
        weight[e] = kPi * r * r / sum;
    }
}

void eit_source_mat(Geometry& g, const double* d_pos, const double* radii, int64_t ne, int order, double* d_out,
                    const Layout& L, hipStream_t st) {
    if (order < 0 || order > 3) throw std::invalid_argument("Gauss order must be 0..3");
    const int64_t n = g.size();
    if (ne <= 0) return;
    fill_f64(d_out, n, ne, L.ld, 0.0, st);
    std::vector<std::vector<int64_t>> tris;
    std::vector<double> weight;
    injection_triangles(g, d_pos, radii, ne, st, tris, weight);
    DeviceGeometry& dg = device_geometry(g);
    std::vector<int> tmesh;
    for (int m = 0; m < (int)g.meshes.size(); ++m) tmesh.insert(tmesh.end(), g.meshes[m].tris.size(), m);
    for (int64_t e = 0; e < ne; ++e) {
        double* col = d_out + e * L.ld;
        const bool intensity = radii == nullptr || almost_zero(radii[e]);
        for (int64_t t : tris[e]) {
            const int m1 = tmesh[t];
            if (!g.meshes[m1].current_barrier) continue;   // only barrier rows of the transfer matrix
            double w = weight[e];
            if (intensity) {   // 1 / |t|
                const auto& a = g.vertices[g.meshes[m1].tris[t - dg.meshes[m1].tri_begin][0]];
                const auto& b = g.vertices[g.meshes[m1].tris[t - dg.meshes[m1].tri_begin][1]];
                const auto& q = g.vertices[g.meshes[m1].tris[t - dg.meshes[m1].tri_begin][2]];
                const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const double e2[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
                const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2],
                             z = e1[0] * e2[1] - e1[1] * e2[0];
                w = 1.0 / (0.5 * std::sqrt(x * x + y * y + z * z));
            }
            for (const MeshPair& mp : g.pairs) {
                if (mp.m1 != m1) continue;
                const MeshDev& B = dg.meshes[mp.m2];
                const double o = mp.orientation;
                const int64_t ns = B.slot_end - B.slot_begin, nu = B.tri_end - B.tri_begin;
                const double cI = (mp.m2 == m1) ? -0.5 * w : 0.0;
                if (ns > 0) {
                    k_eit_d<<<(unsigned)cdiv(ns, 256), 256, 0, st>>>(dg.tri, t, dg.tri_vid, dg.adj_off, dg.adj,
                                                                     dg.slot_vid, dg.vert_unk, B.slot_begin, ns, order,
                                                                     kK * o * w, cI, col, L);
                    OMH_LAUNCH_CHECK();
                }
                if (mp.m2 != m1 && !B.barrier && nu > 0) {
                    const double cS = -kK * o * g.sigma_inv(m1, mp.m2) * w;
                    k_eit_s<<<(unsigned)cdiv(nu, 256), 256, 0, st>>>(dg.tri, t, B.tri_begin, nu, B.tri_unk0, order,
                                                                     cS, col, L);
                    OMH_LAUNCH_CHECK();
                }
            }
        }
    }
}

}  // namespace omh
