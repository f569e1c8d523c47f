// Adaptive triangle quadrature on the device: Integrator::integrate
// (integrator.h:44-48) with the recursion of adaptive_integration
// (integrator.h:69-93) unrolled onto an explicit stack of at most
// ADAPT_MAX_FRAMES frames.  The integrand is a functor f(x, out[NC]); the
// tolerance test uses |.| for NC = 1 and the Euclidean norm for NC = 3, as
// Integrator::norm.
#pragma once

#include <hip/hip_runtime.h>

#include "bem_device.hpp"

namespace omh {

constexpr int ADAPT_MAX_FRAMES = 12;

template <int NC, class F>
__device__ __forceinline__ void tri_integration(const F& f, int order, d3 A, d3 B, d3 C, double area, double* res) {
    const int np = c_nb_pts[order];
    double acc[NC];
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (int i = 0; i < np; ++i) {
        const double b0 = c_rules[order][i][0], b1 = c_rules[order][i][1], b2 = c_rules[order][i][2];
        const d3 x = mk(b0 * A.x + b1 * B.x + b2 * C.x, b0 * A.y + b1 * B.y + b2 * C.y, b0 * A.z + b1 * B.z + b2 * C.z);
        double v[NC];
        f(x, v);
        const double w = c_rules[order][i][3];
        for (int c = 0; c < NC; ++c) acc[c] += w * v[c];
    }
    for (int c = 0; c < NC; ++c) res[c] = area * acc[c];
}

// 4-way split of integrator.h:71-75.
__device__ __forceinline__ void child_tri(const d3 P[3], int c, d3 out[3]) {
    const d3 m0 = scl(0.5, add(P[1], P[2])), m1 = scl(0.5, add(P[2], P[0])), m2 = scl(0.5, add(P[0], P[1]));
    switch (c) {
        case 0: out[0] = P[0]; out[1] = m1; out[2] = m2; break;
        case 1: out[0] = m0; out[1] = P[1]; out[2] = m2; break;
        case 2: out[0] = m0; out[1] = m1; out[2] = P[2]; break;
        default: out[0] = m0; out[1] = m1; out[2] = m2; break;
    }
}

template <int NC>
__device__ __forceinline__ double vnorm(const double* a) {
    if (NC == 1) return fabs(a[0]);
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

template <int NC>
__device__ __forceinline__ bool converged(const double* coarse, const double* refined, double tol) {
    double d[NC];
    for (int c = 0; c < NC; ++c) d[c] = coarse[c] - refined[c];
    return vnorm<NC>(d) <= tol * vnorm<NC>(coarse);
}

template <int NC>
struct Frame {
    d3 P[3];
    double area;
    double integ[4][NC];
    double sum[NC];
    int child;
    int level;
};

// Integrator::integrate (integrator.h:44-48) with the adaptive recursion of
// integrator.h:69-93 unrolled onto an explicit stack; the child results are
// summed in the reference's order.
template <int NC, class F>
__device__ void integrate(const F& f, int order, int levels, double tol, const d3 P[3], double area, double* res) {
    double coarse[NC];
    tri_integration<NC>(f, order, P[0], P[1], P[2], area, coarse);
    if (levels <= 0) {
        for (int c = 0; c < NC; ++c) res[c] = coarse[c];
        return;
    }
    Frame<NC> st[ADAPT_MAX_FRAMES];
    int top = 0;
    st[0].P[0] = P[0]; st[0].P[1] = P[1]; st[0].P[2] = P[2];
    st[0].area = area;
    double refined[NC];
    for (int c = 0; c < NC; ++c) refined[c] = 0.0;
    for (int i = 0; i < 4; ++i) {
        d3 C[3];
        child_tri(P, i, C);
        tri_integration<NC>(f, order, C[0], C[1], C[2], 0.25 * area, st[0].integ[i]);
        for (int c = 0; c < NC; ++c) refined[c] += st[0].integ[i][c];
    }
    if (converged<NC>(coarse, refined, tol)) {
        for (int c = 0; c < NC; ++c) res[c] = refined[c];
        return;
    }
    st[0].level = levels;
    st[0].child = 0;
    for (int c = 0; c < NC; ++c) st[0].sum[c] = 0.0;
    while (true) {
        Frame<NC>& fr = st[top];
        if (fr.child == 4) {
            if (top == 0) {
                for (int c = 0; c < NC; ++c) res[c] = fr.sum[c];
                return;
            }
            --top;
            for (int c = 0; c < NC; ++c) st[top].sum[c] += fr.sum[c];
            st[top].child++;
            continue;
        }
        // call adaptive(child, area/4, integ[child], level-1)
        d3 C[3];
        child_tri(fr.P, fr.child, C);
        const double ca = 0.25 * fr.area;
        double integ[4][NC], ref[NC];
        for (int c = 0; c < NC; ++c) ref[c] = 0.0;
        for (int i = 0; i < 4; ++i) {
            d3 G[3];
            child_tri(C, i, G);
            tri_integration<NC>(f, order, G[0], G[1], G[2], 0.25 * ca, integ[i]);
            for (int c = 0; c < NC; ++c) ref[c] += integ[i][c];
        }
        const int lvl = fr.level - 1;
        if (lvl == 0 || converged<NC>(fr.integ[fr.child], ref, tol) || top + 1 >= ADAPT_MAX_FRAMES) {
            for (int c = 0; c < NC; ++c) fr.sum[c] += ref[c];
            fr.child++;
        } else {
            Frame<NC>& nf = st[top + 1];
            nf.P[0] = C[0]; nf.P[1] = C[1]; nf.P[2] = C[2];
            nf.area = ca;
            for (int i = 0; i < 4; ++i)
                for (int c = 0; c < NC; ++c) nf.integ[i][c] = integ[i][c];
            for (int c = 0; c < NC; ++c) nf.sum[c] = 0.0;
            nf.child = 0;
            nf.level = lvl;
            ++top;
        }
    }
}

}  // namespace omh
