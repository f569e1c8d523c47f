// Device-side integration kernels of the symmetric BEM, FP64.
//
// Restated from the mathematics the reference implements (file:line in
// OpenMEEG 2.5.16):
//   Gauss rules (Bonnet, App. B.3)       integrator.h:95,108-159
//   Gauss point / weighted sum           integrator.h:55-66
//   OperatorS (Ferguson-Zhang-Stroink)   analytics.h:19-67
//   solid angle (De Munck)               triangle.h:123-136
//   P1SingleLayerPotential (Graglia)     analytics.h:75-128
//   Dipole potential / gradient          dipole.h:29-46
//   BarycentricCoordinates               BarycentricCoordinates.h:30-70
//
// Per-triangle constants are precomputed once per geometry into TriRec so
// that every pair evaluation only reads them.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace omh {

struct TriRec {            // 36 doubles = 288 B
    double p[3][3];        // vertices
    double n[3];           // unit normal (p1-p0)x(p2-p0)/|.|
    double area;
    double inv_area;
    double len[3];         // |p_i - p_{i+1}| (OperatorS edge lengths)
    double enrm[3][3];     // n x unit(p_i - p_{i+1}) (OperatorS in-plane edge normals)
    double U[3][3];        // unit(p_{i+1} - p_i) (P1SLP edge directions)
    double diam;           // longest edge (far-field test)
};

struct AdjRec {            // one (vertex slot, triangle) incidence of a mesh
    int32_t tri;           // global triangle id
    int32_t k;             // position of the vertex in the triangle (0..2)
    double cb[3];          // opposite edge A-B of triangle.h:87-90,154
};

static __device__ __constant__ double c_rules[4][16][4] = {
    {{0.166666666666667, 0.166666666666667, 0.666666666666667, 0.333333333333334},
     {0.166666666666667, 0.666666666666667, 0.166666666666667, 0.333333333333334},
     {0.666666666666667, 0.166666666666667, 0.166666666666667, 0.333333333333334}},
    {{0.445948490915965, 0.445948490915965, 0.108103018168070, 0.223381589678010},
     {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.223381589678010},
     {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.223381589678010},
     {0.091576213509771, 0.091576213509771, 0.816847572980458, 0.109951743655322},
     {0.091576213509771, 0.816847572980458, 0.091576213509771, 0.109951743655322},
     {0.816847572980458, 0.091576213509771, 0.091576213509771, 0.109951743655322}},
    {{0.333333333333333, 0.333333333333333, 0.333333333333333, 0.2250},
     {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
     {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
     {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
     {0.101286507323456, 0.101286507323456, 0.797426985353088, 0.125939180544828},
     {0.101286507323456, 0.797426985353088, 0.101286507323456, 0.125939180544828},
     {0.797426985353088, 0.101286507323456, 0.101286507323456, 0.125939180544828}},
    {{0.333333333333333, 0.333333333333333, 0.333333333333333, 0.144315607677786},
     {0.081414823414554, 0.459292588292722, 0.459292588292722, 0.095091634267284},
     {0.459292588292722, 0.081414823414554, 0.459292588292722, 0.095091634267284},
     {0.459292588292722, 0.459292588292722, 0.081414823414554, 0.095091634267284},
     {0.898905543365937, 0.050547228317031, 0.050547228317031, 0.032458497623198},
     {0.050547228317031, 0.898905543365937, 0.050547228317031, 0.032458497623198},
     {0.050547228317031, 0.050547228317031, 0.898905543365937, 0.032458497623198},
     {0.658861384496479, 0.170569307751760, 0.170569307751761, 0.103217370534718},
     {0.170569307751760, 0.658861384496479, 0.170569307751761, 0.103217370534718},
     {0.170569307751760, 0.170569307751761, 0.658861384496479, 0.103217370534718},
     {0.008394777409957, 0.728492392955404, 0.263112829634639, 0.027230314174434},
     {0.728492392955404, 0.008394777409957, 0.263112829634639, 0.027230314174434},
     {0.728492392955404, 0.263112829634639, 0.008394777409957, 0.027230314174434},
     {0.008394777409957, 0.263112829634639, 0.728492392955404, 0.027230314174434},
     {0.263112829634639, 0.008394777409957, 0.728492392955404, 0.027230314174434},
     {0.263112829634639, 0.728492392955404, 0.008394777409957, 0.027230314174434}}};

static __device__ __constant__ int c_nb_pts[4] = {3, 6, 7, 16};

struct d3 { double x, y, z; };

__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 scl(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double norm(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

// ---------------------------------------------------------------------------
// Far field.  For a triangle pair whose centroid distance d is at least
// FAR_NEAR times the longer of the two longest edges, the inner analytic
// integral of a pair integral is replaced by Gauss quadrature on the inner
// triangle: the 16-point rule on both triangles, or from FAR_COARSE on the
// 7-point rule on both.  Against the analytic OperatorS the relative error is
// below 1.2e-11 at d = 4 diameters (16 x 16) and 8e-11 at 10 (7 x 7); the
// analytic P1 double layer itself loses digits to cancellation at such
// distances.
// Per-triangle Gauss points are tabulated: GP_STRIDE doubles per triangle,
// the 16 points of the order-3 rule then the 7 of the order-2 rule.
// ---------------------------------------------------------------------------
constexpr double FAR_NEAR = 4.0, FAR_COARSE = 10.0;
constexpr int GP16 = 0, GP7 = 16 * 3, GP_STRIDE = 23 * 3;

// Gauss point i of the triangle with vertices P (integrator.h:59-61).
__device__ __forceinline__ d3 gauss_point(int order, int i, const d3 P[3]) {
    const double b0 = c_rules[order][i][0], b1 = c_rules[order][i][1], b2 = c_rules[order][i][2];
    return mk(b0 * P[0].x + b1 * P[1].x + b2 * P[2].x,
              b0 * P[0].y + b1 * P[1].y + b2 * P[2].y,
              b0 * P[0].z + b1 * P[1].z + b2 * P[2].z);
}

// Triangle "analytic side" of OperatorS, in registers.
struct OpS {
    d3 p0, p1, p2, n, e0, e1, e2;
    double area, l0, l1, l2;
};

__device__ __forceinline__ OpS load_ops(const TriRec& t) {
    OpS o;
    o.p0 = ld3(t.p[0]); o.p1 = ld3(t.p[1]); o.p2 = ld3(t.p[2]);
    o.n = ld3(t.n);
    o.e0 = ld3(t.enrm[0]); o.e1 = ld3(t.enrm[1]); o.e2 = ld3(t.enrm[2]);
    o.area = t.area;
    o.l0 = t.len[0]; o.l1 = t.len[1]; o.l2 = t.len[2];
    return o;
}

// OperatorS::operator() (analytics.h:34-55) including the De Munck solid
// angle of triangle.h:123-136.
__device__ __forceinline__ double ops_eval(const OpS& o, d3 x) {
    const d3 r0 = sub(o.p0, x), r1 = sub(o.p1, x), r2 = sub(o.p2, x);
    const double d0 = norm(r0), d1 = norm(r1), d2 = norm(r2);
    // atanh(l/R) written as log1p(2l/(R-l))/2: one division instead of two
    double sum = dot(r0, o.e0) * (0.5 * log1p(2 * o.l0 / (d0 + d1 - o.l0)));
    sum += dot(r1, o.e1) * (0.5 * log1p(2 * o.l1 / (d1 + d2 - o.l1)));
    sum += dot(r2, o.e2) * (0.5 * log1p(2 * o.l2 / (d2 + d0 - o.l2)));
    const double hn = dot(r0, o.n);
    const double det = 2 * o.area * hn;
    double omega = 0.0;
    if (fabs(det) >= 1e-10) {
        const double prod = d0 * d1 * d2;
        const double den = prod + (d0 * dot(r1, r2) + d1 * dot(r2, r0) + d2 * dot(r0, r1));
        omega = 2 * atan2(det, den);
    }
    return 2 * sum - hn * omega;
}

// Triangle "analytic side" of P1SingleLayerPotential, in registers.
struct P1S {
    d3 p0, p1, p2, n, U0, U1, U2;
    double area, inv_area;
};

__device__ __forceinline__ P1S load_p1s(const TriRec& t) {
    P1S o;
    o.p0 = ld3(t.p[0]); o.p1 = ld3(t.p[1]); o.p2 = ld3(t.p[2]);
    o.n = ld3(t.n);
    o.U0 = ld3(t.U[0]); o.U1 = ld3(t.U[1]); o.U2 = ld3(t.U[2]);
    o.area = t.area;
    o.inv_area = t.inv_area;
    return o;
}

// P1SingleLayerPotential::operator() (analytics.h:87-119).
__device__ __forceinline__ d3 p1s_eval(const P1S& o, d3 x) {
    const d3 r0 = sub(o.p0, x), r1 = sub(o.p1, x), r2 = sub(o.p2, x);
    const double h = dot(r0, o.n);
    if (fabs(h) < 1e-10) return mk(0.0, 0.0, 0.0);
    const double d0 = norm(r0), d1 = norm(r1), d2 = norm(r2);
    // S = sum_i U_i log((d_{i+1} + r_{i+1}.U_i) / (d_i + r_i.U_i))
    const double g0 = log((d1 + dot(r1, o.U0)) / (d0 + dot(r0, o.U0)));
    const double g1 = log((d2 + dot(r2, o.U1)) / (d1 + dot(r1, o.U1)));
    const double g2 = log((d0 + dot(r0, o.U2)) / (d2 + dot(r2, o.U2)));
    const d3 S = add(add(scl(g0, o.U0), scl(g1, o.U1)), scl(g2, o.U2));
    const double dd = 2 * h * o.area;
    const double d1v = d0 * d1 * d2 + (d0 * dot(r1, r2) + d1 * dot(r2, r0) + d2 * dot(r0, r1));
    const double omega = atan2(dd, d1v);
    const d3 D0 = sub(o.p1, o.p0), D1 = sub(o.p2, o.p1), D2 = sub(o.p0, o.p2);
    const double E10 = dot(cross(r1, r2), o.n), E11 = dot(cross(r2, r0), o.n), E12 = dot(cross(r0, r1), o.n);
    const double E20 = dot(D1, S), E21 = dot(D2, S), E22 = dot(D0, S);
    const double hh = 0.5 * h;
    return mk((omega * E10 + hh * E20) * o.inv_area, (omega * E11 + hh * E21) * o.inv_area,
              (omega * E12 + hh * E22) * o.inv_area);
}

// Dipole::potential (dipole.h:29-35); the 1/(4 pi) factor is applied outside.
__device__ __forceinline__ double dip_potential(d3 r0, d3 q, d3 r) {
    const d3 x = sub(r, r0);
    const double n2 = dot(x, x);
    return dot(q, x) / (n2 * sqrt(n2));
}

// Dipole::gradient . n (dipole.h:38-46).
__device__ __forceinline__ double dip_gradient_n(d3 r0, d3 q, d3 r, d3 n) {
    const d3 x = sub(r, r0);
    const double inv = 1.0 / dot(x, x);
    const d3 g = sub(q, scl(3 * dot(q, x) * inv, x));
    return dot(g, n) * (inv * sqrt(inv));
}

// Far-field classification of a pair from the tabulated centroids (the
// first point of the 16-point rule): 0 near (analytic), 1 far (16-point
// inner rule), 2 farther (7-point inner rule).
__device__ __forceinline__ int far_class(const double* __restrict__ g1, const double* __restrict__ g2, double diam) {
    const double d = norm(sub(ld3(g1), ld3(g2)));
    return d >= FAR_COARSE * diam ? 2 : (d >= FAR_NEAR * diam ? 1 : 0);
}

// sum over the NO outer points y_k (table go at OOFF, rule OO) of
// w_k sum_a w_a / |y_k - x_a| over the NI inner points x_a (table gi at IOFF,
// rule IO), the inner points held in registers.  Times the inner area it is
// the outer Gauss sum of OperatorS (analytics.h:19-67) of the inner triangle.
template <int NO, int OO, int OOFF, int NI, int IO, int IOFF>
__device__ __forceinline__ double s_far_n(const double* __restrict__ go, const double* __restrict__ gi) {
    d3 x[NI];
#pragma unroll
    for (int a = 0; a < NI; ++a) x[a] = ld3(gi + IOFF + 3 * a);
    double res = 0.0;
    for (int k = 0; k < NO; ++k) {
        const d3 y = ld3(go + OOFF + 3 * k);
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < NI; ++a) {
            const d3 r = sub(y, x[a]);
            acc += c_rules[IO][a][3] * rsqrt(dot(r, r));
        }
        res += c_rules[OO][k][3] * acc;
    }
    return res;
}

// Far class 1: 16 x 16 points; 2: 7 x 7 points (16- and 7-point rules).
__device__ __forceinline__ double s_far(const double* __restrict__ go, const double* __restrict__ gi, int fc) {
    return fc == 2 ? s_far_n<7, 2, GP7, 7, 2, GP7>(go, gi) : s_far_n<16, 3, GP16, 16, 3, GP16>(go, gi);
}

// sum over the NO outer points x_k (table go) of w_k sum_b w_b lambda_b
// n.(y_b - x_k) / |y_b - x_k|^3 over the NI inner points y_b (table gi,
// registers) with their barycentric coordinates lambda_b: times the two
// areas, the outer Gauss sum of the P1 double layer (analytics.h:75-128).
template <int NO, int OO, int OOFF, int NI, int IO, int IOFF>
__device__ __forceinline__ d3 d_far_n(const double* __restrict__ go, const double* __restrict__ gi, d3 n) {
    d3 y[NI];
#pragma unroll
    for (int a = 0; a < NI; ++a) y[a] = ld3(gi + IOFF + 3 * a);
    double rx = 0.0, ry = 0.0, rz = 0.0;
    for (int k = 0; k < NO; ++k) {
        const d3 x = ld3(go + OOFF + 3 * k);
        double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
        for (int a = 0; a < NI; ++a) {
            const d3 r = sub(y[a], x);
            const double ir = rsqrt(dot(r, r));
            const double kw = c_rules[IO][a][3] * dot(n, r) * (ir * ir * ir);
            ax += kw * c_rules[IO][a][0];
            ay += kw * c_rules[IO][a][1];
            az += kw * c_rules[IO][a][2];
        }
        const double w = c_rules[OO][k][3];
        rx += w * ax;
        ry += w * ay;
        rz += w * az;
    }
    return mk(rx, ry, rz);
}

__device__ __forceinline__ d3 d_far(const double* __restrict__ go, const double* __restrict__ gi, d3 n, int fc) {
    return fc == 2 ? d_far_n<7, 2, GP7, 7, 2, GP7>(go, gi, n) : d_far_n<16, 3, GP16, 16, 3, GP16>(go, gi, n);
}

// 1/sqrt(x) for finite x > 0: the hardware estimate and one Newton step
// with its cubic term (y = y0 + y0 e (1/2 + 3/8 e), e = 1 - x y0^2).
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// One far-field triangle pair (t, u) for all three pair operators from a
// single set of Gauss-point distances.  With x_p the points of t and y_q
// those of u (NP-point rule R, table offset OFF), d = y_q - x_p:
//   s   = sum_p w_p sum_q w_q / |d|                          (single layer)
//   a_c = sum_q w_q l_qc sum_p w_p (n_u . d) / |d|^3         (P1 double layer of u, t outer)
//   b_c = -sum_p w_p l_pc sum_q w_q (n_t . d) / |d|^3        (P1 double layer of t, u outer)
// (l: barycentric coordinates of the rule's points).  s times the two areas
// is s_far's pair value, a and b those of d_far for (t, u) and (u, t); a
// (b) is left zero unless DA (DB).  n . d is taken as n . y_q - n . x_p
// (far pairs: |d| is at least 4 triangle diameters, the difference loses
// no more than a few units of the last place relative to |n . d| + |d|).
// PRE: keep n . y_q of both normals in registers (the 7-point rule; the
// 16-point rule recomputes them, for its register budget).
template <int NP, int R, int OFF, bool DA, bool DB>
__device__ __forceinline__ void pair_far_n(const double* __restrict__ gt, const double* __restrict__ gu, d3 nt, d3 nu,
                                           double& s, d3& a, d3& b) {
    constexpr bool PRE = NP <= 8;
    constexpr int NPRE = PRE ? NP : 1;
    d3 y[NP];
    double uy[NPRE], ty[NPRE], G[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        y[q] = ld3(gu + OFF + 3 * q);
        if (PRE) {
            uy[q % NPRE] = dot(nu, y[q]);
            ty[q % NPRE] = dot(nt, y[q]);
        }
        G[q] = 0.0;
    }
    s = 0.0;
    b = mk(0.0, 0.0, 0.0);
#pragma unroll 1
    for (int p = 0; p < NP; ++p) {
        const d3 x = ld3(gt + OFF + 3 * p);
        const double ux = dot(nu, x), tx = dot(nt, x);
        const double wp = c_rules[R][p][3];
        double sp = 0.0, tp = 0.0;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const d3 d = sub(y[q], x);
            const double ir = rsqrt_pos(dot(d, d));
            const double wq = c_rules[R][q][3];
            sp = fma(wq, ir, sp);
            if (DA || DB) {
                const double ir3 = ir * ir * ir;
                const double nud = PRE ? uy[q % NPRE] - ux : dot(nu, d);
                const double ntd = PRE ? ty[q % NPRE] - tx : dot(nt, d);
                if (DA) G[q] = fma(wp, nud * ir3, G[q]);
                if (DB) tp = fma(ntd, wq * ir3, tp);
            }
        }
        s = fma(wp, sp, s);
        if (DB) {
            const double wt = wp * tp;
            b.x -= wt * c_rules[R][p][0];
            b.y -= wt * c_rules[R][p][1];
            b.z -= wt * c_rules[R][p][2];
        }
    }
    a = mk(0.0, 0.0, 0.0);
    if (DA) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const double g = c_rules[R][q][3] * G[q];
            a.x = fma(g, c_rules[R][q][0], a.x);
            a.y = fma(g, c_rules[R][q][1], a.y);
            a.z = fma(g, c_rules[R][q][2], a.z);
        }
    }
}

// Far classes of far_class: 1 (16-point rules), 2 (7-point rules); DA / DB
// select the double-layer values computed.
template <bool DA, bool DB>
__device__ __forceinline__ void pair_far(const double* __restrict__ gt, const double* __restrict__ gu, d3 nt, d3 nu,
                                         int fc, double& s, d3& a, d3& b) {
    if (fc == 2) pair_far_n<7, 2, GP7, DA, DB>(gt, gu, nt, nu, s, a, b);
    else pair_far_n<16, 3, GP16, DA, DB>(gt, gu, nt, nu, s, a, b);
}

}  // namespace omh
