/*
 * CPU oracle: a plain-C restatement of OpenMEEG's symmetric-BEM EEG forward
 * path (integration kernels, HeadMat block assembly, DipSourceMat, Danielsson
 * closest-point search).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, the smoke() check and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Storage follows the reference: the HeadMat is a packed upper symmetric
 * matrix with index i + j*(j+1)/2, i <= j (OpenMEEGMaths/include/symmatrix.h:52-62);
 * dense matrices are column-major (OpenMEEGMaths/include/matrix.h:81-92).
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp, -ffp-contract=off so that
 * the arithmetic follows the C++ expression order of the reference).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ------------------------------------------------------------------ */
/* Vect3 helpers (OpenMEEG/include/vect3.h:72-114)                      */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } V3;

static inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 scl(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double sqr(double a) { return a * a; }
static inline double nrm2(V3 a) { return sqr(a.x) + sqr(a.y) + sqr(a.z); }
static inline double nrm(V3 a) { return sqrt(nrm2(a)); }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ------------------------------------------------------------------ */
/* Triangle (OpenMEEG/include/triangle.h:57-62, 123-136)               */
/* ------------------------------------------------------------------ */
typedef struct { V3 p[3]; V3 n; double area; } Tri;

static Tri make_tri(V3 a, V3 b, V3 c) {
    Tri t; t.p[0] = a; t.p[1] = b; t.p[2] = c;
    V3 on = cross(sub(b, a), sub(c, a));
    double nr = nrm(on);
    t.area = 0.5 * nr;
    t.n = v3(on.x / nr, on.y / nr, on.z / nr);
    return t;
}

/* De Munck solid angle, triangle.h:123-136. */
static double tri_solid_angle(const Tri* t, V3 v) {
    V3 r0 = sub(t->p[0], v), r1 = sub(t->p[1], v), r2 = sub(t->p[2], v);
    V3 d = v3(nrm(r0), nrm(r1), nrm(r2));
    V3 pr = v3(dot(r1, r2), dot(r2, r0), dot(r0, r1));
    double prod = d.x * d.y * d.z;
    double det = 2 * t->area * dot(r0, t->n);
    return (fabs(det) < 1e-10) ? 0.0 : 2 * atan2(det, prod + dot(d, pr));
}

/* ------------------------------------------------------------------ */
/* Quadrature rules, integrator.h:95,108-159 (Bonnet, appendix B.3)     */
/* ------------------------------------------------------------------ */
static const int nb_pts[4] = {3, 6, 7, 16};
static const double rules[4][16][4] = {
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

/* Gauss point i of a triangle given by its 3 points: the multadd sequence of
 * integrator.h:59-61. */
static inline V3 gauss_point(int order, int i, const V3 P[3]) {
    V3 v = v3(0.0, 0.0, 0.0);
    for (int j = 0; j < 3; ++j) {
        double b = rules[order][i][j];
        v = v3(v.x + b * P[j].x, v.y + b * P[j].y, v.z + b * P[j].z);
    }
    return v;
}

/* ------------------------------------------------------------------ */
/* OperatorS, analytics.h:19-67 (Ferguson, Zhang & Stroink 1994)        */
/* ------------------------------------------------------------------ */
typedef struct { const Tri* t; double len[3]; V3 enrm[3]; } OpS;

static void ops_init(OpS* o, const Tri* t) {
    static const int next[3] = {1, 2, 0};
    o->t = t;
    for (int i = 0; i < 3; ++i) {
        V3 vref = sub(t->p[i], t->p[next[i]]);
        double len = nrm(vref);
        V3 unit = v3(vref.x / len, vref.y / len, vref.z / len);
        o->len[i] = len;
        o->enrm[i] = cross(t->n, unit);
    }
}

static double ops_eval(const OpS* o, V3 x) {
    const Tri* t = o->t;
    V3 r[3] = {sub(t->p[0], x), sub(t->p[1], x), sub(t->p[2], x)};
    double d[3] = {nrm(r[0]), nrm(r[1]), nrm(r[2])};
    double R[3] = {d[0] + d[1], d[1] + d[2], d[2] + d[0]};
    double sum = 0.0;
    for (int i = 0; i < 3; ++i)
        sum += dot(r[i], o->enrm[i]) * atanh(o->len[i] / R[i]);
    return 2 * sum - dot(r[0], t->n) * tri_solid_angle(t, x);
}

/* Non-adaptive integral of OperatorS(t1) over t2, integrator.h:55-66. */
static double integrate_S(const OpS* o, const Tri* t2, int order) {
    double res = 0.0;
    for (int i = 0; i < nb_pts[order]; ++i)
        res += rules[order][i][3] * ops_eval(o, gauss_point(order, i, t2->p));
    return t2->area * res;
}

/* ------------------------------------------------------------------ */
/* P1SingleLayerPotential, analytics.h:75-128 (Graglia 1993)            */
/* ------------------------------------------------------------------ */
typedef struct { const Tri* t; V3 D[3]; V3 U[3]; } P1S;

static void p1s_init(P1S* o, const Tri* t) {
    o->t = t;
    o->D[0] = sub(t->p[1], t->p[0]);
    o->D[1] = sub(t->p[2], t->p[1]);
    o->D[2] = sub(t->p[0], t->p[2]);
    for (int i = 0; i < 3; ++i) {
        double l = nrm(o->D[i]);
        o->U[i] = v3(o->D[i].x / l, o->D[i].y / l, o->D[i].z / l);
    }
}

static V3 p1s_eval(const P1S* o, V3 x) {
    static const int next[3] = {1, 2, 0};
    const Tri* t = o->t;
    V3 r[3] = {sub(t->p[0], x), sub(t->p[1], x), sub(t->p[2], x)};
    double ds[3] = {nrm(r[0]), nrm(r[1]), nrm(r[2])};
    V3 Z[3] = {cross(r[1], r[2]), cross(r[2], r[0]), cross(r[0], r[1])};
    double pr[3] = {dot(r[1], r[2]), dot(r[2], r[0]), dot(r[0], r[1])};
    const V3 n = t->n;
    double h = dot(r[0], n);
    double dd = 2 * h * t->area;
    if (fabs(h) < 1e-10) return v3(0.0, 0.0, 0.0);
    V3 S = v3(0.0, 0.0, 0.0);
    for (int i = 0; i < 3; ++i) {
        int i1 = next[i];
        double lg = log((ds[i1] + dot(r[i1], o->U[i])) / (ds[i] + dot(r[i], o->U[i])));
        S = add(S, scl(lg, o->U[i]));
    }
    double d1 = ds[0] * ds[1] * ds[2] + (ds[0] * pr[0] + ds[1] * pr[1] + ds[2] * pr[2]);
    double omega = atan2(dd, d1);
    V3 E1 = v3(dot(Z[0], n), dot(Z[1], n), dot(Z[2], n));
    V3 E2 = v3(dot(o->D[1], S), dot(o->D[2], S), dot(o->D[0], S));
    V3 res = add(scl(omega, E1), scl(0.5 * h, E2));
    return v3(res.x / t->area, res.y / t->area, res.z / t->area);
}

/* Integral over t1 (Gauss) of P1SLP(t2): the inner call of BlocksBase::D,
 * operators.h:104-106. */
static V3 integrate_D(const P1S* o, const Tri* t1, int order) {
    V3 res = v3(0.0, 0.0, 0.0);
    for (int i = 0; i < nb_pts[order]; ++i)
        res = add(res, scl(rules[order][i][3], p1s_eval(o, gauss_point(order, i, t1->p))));
    return scl(t1->area, res);
}

/* ------------------------------------------------------------------ */
/* Adaptive integration, integrator.h:44-93                             */
/* ------------------------------------------------------------------ */
typedef struct { V3 r0, q; } Dip;     /* dipole.h:18-52 */

static double dip_potential(const Dip* d, V3 r) {     /* dipole.h:29-35 */
    V3 x = sub(r, d->r0);
    double n2 = nrm2(x);
    return dot(d->q, x) / (n2 * sqrt(n2));
}
static double dip_gradient(const Dip* d, V3 r, V3 n) { /* dipole.h:38-46 */
    V3 x = sub(r, d->r0);
    double inv = 1.0 / nrm2(x);
    V3 gr = sub(d->q, scl(3 * dot(d->q, x) * inv, x));
    return dot(gr, n) * (inv * sqrt(inv));
}

/* BarycentricCoordinates, BarycentricCoordinates.h:30-70 */
typedef struct { V3 p0, e10, e20; double d11, d12, d22; } Bary;
static void bary_init(Bary* b, const Tri* t) {
    b->p0 = t->p[0];
    b->e10 = sub(t->p[1], t->p[0]);
    b->e20 = sub(t->p[2], t->p[0]);
    double c11 = dot(b->e10, b->e10), c12 = dot(b->e10, b->e20), c22 = dot(b->e20, b->e20);
    double inv = 1.0 / (c11 * c22 - c12 * c12);
    b->d11 = c22 * inv; b->d12 = -c12 * inv; b->d22 = c11 * inv;
}
static V3 bary_eval(const Bary* b, V3 r) {
    V3 v = sub(r, b->p0);
    double d1 = dot(v, b->e10), d2 = dot(v, b->e20);
    double l1 = b->d11 * d1 + b->d12 * d2;
    double l2 = b->d12 * d1 + b->d22 * d2;
    return v3(1.0 - l1 - l2, l1, l2);
}

/* The integrand: kind 0 = Dipole::potential (scalar in .x),
 * kind 1 = PotentialDerivative (analytics.h:132-158), kind 2 = OperatorS of
 * a triangle (scalar, the S pair integral), kind 3 = P1SingleLayerPotential
 * of a triangle (the D pair integral). */
typedef struct { int kind; Dip dip; Bary bary; V3 n; const OpS* ops; const P1S* p1s; } Fn;
static V3 fn_eval(const Fn* f, V3 r) {
    if (f->kind == 0) return v3(dip_potential(&f->dip, r), 0, 0);
    if (f->kind == 2) return v3(ops_eval(f->ops, r), 0, 0);
    if (f->kind == 3) return p1s_eval(f->p1s, r);
    V3 P1 = bary_eval(&f->bary, r);
    double g = dip_gradient(&f->dip, r, f->n);
    return scl(g, P1);
}
static double fn_norm(int kind, V3 a) { return (kind == 0 || kind == 2) ? fabs(a.x) : nrm(a); }

static V3 tri_integration(const Fn* f, int order, const V3 P[3], double area) {
    V3 res = v3(0, 0, 0);
    for (int i = 0; i < nb_pts[order]; ++i)
        res = add(res, scl(rules[order][i][3], fn_eval(f, gauss_point(order, i, P))));
    return scl(area, res);
}

static V3 adaptive(const Fn* f, int order, double tol, const V3 P[3], double area,
                   V3 coarse, unsigned level) {
    V3 mid[3] = {scl(0.5, add(P[1], P[2])), scl(0.5, add(P[2], P[0])), scl(0.5, add(P[0], P[1]))};
    V3 nt[4][3] = {{P[0], mid[1], mid[2]}, {mid[0], P[1], mid[2]},
                   {mid[0], mid[1], P[2]}, {mid[0], mid[1], mid[2]}};
    double sa = 0.25 * area;
    V3 refined = v3(0, 0, 0), integ[4];
    for (int i = 0; i < 4; ++i) {
        integ[i] = tri_integration(f, order, nt[i], sa);
        refined = add(refined, integ[i]);
    }
    if (fn_norm(f->kind, sub(coarse, refined)) <= tol * fn_norm(f->kind, coarse) || level == 0)
        return refined;
    V3 sum = v3(0, 0, 0);
    for (int i = 0; i < 4; ++i)
        sum = add(sum, adaptive(f, order, tol, nt[i], sa, integ[i], level - 1));
    return sum;
}

static V3 integrate_fn(const Fn* f, int order, unsigned levels, double tol, const Tri* t) {
    V3 coarse = tri_integration(f, order, t->p, t->area);
    return levels == 0 ? coarse : adaptive(f, order, tol, t->p, t->area, coarse, levels);
}

/* ------------------------------------------------------------------ */
/* Geometry tables handed over from oracle/geometry.py                  */
/* ------------------------------------------------------------------ */
typedef struct {
    i64 nv;                 /* number of geometry vertices              */
    const double* verts;    /* nv x 3                                   */
    i64 nt;                 /* number of triangles over all meshes      */
    const i64* tris;        /* nt x 3 global vertex ids                 */
    const i64* tri_index;   /* nt: unknown index (-1 if none)           */
    const i64* vert_index;  /* nv: unknown index (-1 if invalid)        */
    i64 nm;                 /* number of meshes                         */
    const i64* tri_off;     /* nm+1: triangles of mesh m                */
    const i64* mv_off;      /* nm+1: vertex slots of mesh m             */
    const i64* mv_ids;      /* vertex slots -> global vertex id         */
    const i64* adj_off;     /* (slots+1): slot -> triangles (CSR)       */
    const i64* adj_tri;     /* global triangle ids (mesh order)         */
    const i64* barrier;     /* nm: current_barrier flag                 */
} Geom;

static inline V3 vert(const Geom* g, i64 v) {
    return v3(g->verts[3 * v], g->verts[3 * v + 1], g->verts[3 * v + 2]);
}
static Tri geom_tri(const Geom* g, i64 t) {
    return make_tri(vert(g, g->tris[3 * t]), vert(g, g->tris[3 * t + 1]), vert(g, g->tris[3 * t + 2]));
}

static inline i64 pidx(i64 i, i64 j) { /* symmatrix.h:52-62 */
    if (i > j) { i64 k = i; i = j; j = k; }
    return i + j * (j + 1) / 2;
}

/* --------------------- S blocks, operators.h:209-237, 391-425 ------ */
/* Writes coeff * S(t1,t2) for t1 in mesh m1 (analytic side) and t2 in mesh
 * m2 (Gauss side).  When dst_packed is non-null the value is assigned to
 * the HeadMat entry (tri_index(t1), tri_index(t2)); otherwise into the dense
 * temporary block `tmp` (row t1-local, column t2-local, column-major with
 * leading dimension n1) -- the SymBloc/Bloc temporaries of operators.h:159-173,337-352. */
static void s_block(const Geom* g, i64 m1, i64 m2, int order, unsigned levels, double tol, double coeff,
                    double* dst_packed, double* tmp) {
    const i64 a0 = g->tri_off[m1], a1 = g->tri_off[m1 + 1];
    const i64 b0 = g->tri_off[m2], b1 = g->tri_off[m2 + 1];
    const i64 n1 = a1 - a0;
    const int diag = (m1 == m2);
    #pragma omp parallel for schedule(dynamic, 4)
    for (i64 i = a0; i < a1; ++i) {
        Tri t1 = geom_tri(g, i);
        OpS o; ops_init(&o, &t1);
        for (i64 j = diag ? i : b0; j < b1; ++j) {
            Tri t2 = geom_tri(g, j);
            double v;
            if (levels == 0) {
                v = integrate_S(&o, &t2, order) * coeff;
            } else {   /* Integrator::integrate with subdivision, integrator.h:44-93 */
                Fn f = {0};
                f.kind = 2;
                f.ops = &o;
                v = integrate_fn(&f, order, levels, tol, &t2).x * coeff;
            }
            if (dst_packed) dst_packed[pidx(g->tri_index[i], g->tri_index[j])] = v;
            else {
                tmp[(i - a0) + (j - b0) * n1] = v;
                if (diag) tmp[(j - a0) + (i - b0) * n1] = v;
            }
        }
    }
}

/* ------------------- D blocks, operators.h:84-115 ---------------- */
/* H(tri_index(t1), vert_index(vertex(t2,k))) += coeff * int_{t1} P1SLP(t2)_k */
static void d_block(const Geom* g, i64 mrow, i64 mcol, int order, unsigned levels, double tol, double coeff,
                    double* H) {
    const i64 a0 = g->tri_off[mrow], a1 = g->tri_off[mrow + 1];
    const i64 b0 = g->tri_off[mcol], b1 = g->tri_off[mcol + 1];
    /* Each row t1 is owned by one thread; within a row the t2 loop runs in
     * mesh order, so the accumulation order of every entry is the reference's. */
    #pragma omp parallel for schedule(dynamic, 4)
    for (i64 i = a0; i < a1; ++i) {
        Tri t1 = geom_tri(g, i);
        const i64 row = g->tri_index[i];
        for (i64 j = b0; j < b1; ++j) {
            Tri t2 = geom_tri(g, j);
            P1S o; p1s_init(&o, &t2);
            V3 tot;
            if (levels == 0) {
                tot = integrate_D(&o, &t1, order);
            } else {
                Fn f = {0};
                f.kind = 3;
                f.p1s = &o;
                tot = integrate_fn(&f, order, levels, tol, &t1);
            }
            double c[3] = {tot.x, tot.y, tot.z};
            for (int k = 0; k < 3; ++k)
                H[pidx(row, g->vert_index[g->tris[3 * j + k]])] += c[k] * coeff;
        }
    }
}

/* edge(V) of a triangle: the two other vertices in the order of
 * triangle.h:87-90,154 ({1,2},{2,0},{0,1}); returns CB = A - B. */
static V3 opposite_edge(const Geom* g, i64 t, i64 v) {
    static const int ind[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    int k = 0;
    while (g->tris[3 * t + k] != v) ++k;
    V3 A = vert(g, g->tris[3 * t + ind[k][0]]);
    V3 B = vert(g, g->tris[3 * t + ind[k][1]]);
    return sub(A, B);
}

/* N kernel, operators.h:134-148.  S values are read either from the packed
 * HeadMat (by triangle unknown index, scaled by `sscale`) or from the dense
 * temporary block (t1-local rows, t2-local columns). */
static double n_entry(const Geom* g, double factor, i64 s1, i64 s2,
                      const double* Hs, double sscale, const double* tmp, i64 n1,
                      i64 t0a, i64 t0b) {
    double res = 0.0;
    const i64 v1 = g->mv_ids[s1], v2 = g->mv_ids[s2];
    for (i64 p = g->adj_off[s1]; p < g->adj_off[s1 + 1]; ++p) {
        const i64 tp1 = g->adj_tri[p];
        V3 CB1 = opposite_edge(g, tp1, v1);
        Tri T1 = geom_tri(g, tp1);
        for (i64 q = g->adj_off[s2]; q < g->adj_off[s2 + 1]; ++q) {
            const i64 tp2 = g->adj_tri[q];
            V3 CB2 = opposite_edge(g, tp2, v2);
            Tri T2 = geom_tri(g, tp2);
            double s;
            if (Hs) s = Hs[pidx(g->tri_index[tp1], g->tri_index[tp2])] * sscale;
            else s = tmp[(tp1 - t0a) + (tp2 - t0b) * n1];
            res -= factor * dot(CB1, CB2) * s / (T1.area * T2.area);
        }
    }
    return res;
}

/* N blocks: DiagonalBlock::N operators.h:240-248,272-296 and
 * NonDiagonalBlock::N :428-436,455-480. */
static void n_block(const Geom* g, i64 m1, i64 m2, double coeff, const double* Hs,
                    double sscale, const double* tmp, double* H) {
    const i64 s0 = g->mv_off[m1], s1e = g->mv_off[m1 + 1];
    const i64 r0 = g->mv_off[m2], r1e = g->mv_off[m2 + 1];
    const i64 n1 = g->tri_off[m1 + 1] - g->tri_off[m1];
    const i64 t0a = g->tri_off[m1], t0b = g->tri_off[m2];
    const int same = (m1 == m2);
    for (i64 a = s0; a < s1e; ++a) {
        const i64 v1 = g->mv_ids[a];
        const i64 hi = same ? a + 1 : r1e;
        #pragma omp parallel for schedule(static)
        for (i64 b = same ? s0 : r0; b < hi; ++b) {
            const i64 v2 = g->mv_ids[b];
            double factor = same ? 0.25 : ((v1 == v2) ? 0.5 : 0.25);
            double val = n_entry(g, factor, a, b, Hs, sscale, tmp, n1, t0a, t0b);
            H[pidx(g->vert_index[v1], g->vert_index[v2])] += val * coeff;
        }
    }
}

/* One communicating mesh pair: HeadMatrixBlocks::set_blocks,
 * operators.h:517-525, with the block rules of DiagonalBlock /
 * NonDiagonalBlock (operators.h:180-197, 363-384). */
int orc_mesh_pair(const Geom* g, i64 m1, i64 m2, int order, unsigned levels, double tol,
                  double cS, double cN, double cD, double* H) {
    const int b1 = (int)g->barrier[m1], b2 = (int)g->barrier[m2];
    const int same = (m1 == m2);
    double Scoeff = 0.0;
    /* set_S_block */
    if (same ? !b1 : (!b1 && !b2)) {
        s_block(g, m1, m2, order, levels, tol, cS, H, NULL);
        Scoeff = cS;
    }
    /* set_N_block */
    if (Scoeff != 0.0) {
        n_block(g, m1, m2, cN / Scoeff, H, 1.0, NULL, H);
    } else {
        const i64 n1 = g->tri_off[m1 + 1] - g->tri_off[m1];
        const i64 n2 = g->tri_off[m2 + 1] - g->tri_off[m2];
        double* tmp = (double*)calloc((size_t)(n1 * n2), sizeof(double));
        if (!tmp) return -1;
        s_block(g, m1, m2, order, levels, tol, 1.0, NULL, tmp);
        n_block(g, m1, m2, cN, NULL, 1.0, tmp, H);
        free(tmp);
    }
    /* set_D_block */
    if (!b1) d_block(g, m1, m2, order, levels, tol, cD, H);
    /* set_Dstar_block */
    if (!same && !b2) d_block(g, m2, m1, order, levels, tol, cD, H);
    return 0;
}

/* Deflation, assembleHeadMat.cpp:28-60: for one outermost mesh add coef to
 * every (vertex,vertex) pair of that mesh (upper half). */
void orc_deflate_mesh(const Geom* g, i64 m, double coef, double* H) {
    const i64 s0 = g->mv_off[m], s1 = g->mv_off[m + 1];
    for (i64 a = s0; a < s1; ++a)
        for (i64 b = a; b < s1; ++b)
            H[pidx(g->vert_index[g->mv_ids[a]], g->vert_index[g->mv_ids[b]])] += coef;
}

double orc_packed_get(const double* H, i64 i, i64 j) { return H[pidx(i, j)]; }

/* One DSM column contribution of one mesh (assembleSourceMat.cpp:84-99):
 * operatorPotentialDerivative (operators.cpp:63-88) into the vertex rows and,
 * when `with_potential`, operatorPotential (operators.cpp:41-60) into the
 * triangle rows.  `col` is the rhs column (length n). */
void orc_dsm_mesh(const Geom* g, i64 m, const double* dip6, int order, unsigned levels,
                  double tol, double coeffD, double coeffP, int with_potential, double* col) {
    Dip d; d.r0 = v3(dip6[0], dip6[1], dip6[2]); d.q = v3(dip6[3], dip6[4], dip6[5]);
    const i64 a0 = g->tri_off[m], a1 = g->tri_off[m + 1];
    for (i64 t = a0; t < a1; ++t) {
        Tri T = geom_tri(g, t);
        Fn f; f.kind = 1; f.dip = d; f.n = T.n; bary_init(&f.bary, &T);
        V3 v = integrate_fn(&f, order, levels, tol, &T);
        double c[3] = {v.x, v.y, v.z};
        for (int j = 0; j < 3; ++j)
            col[g->vert_index[g->tris[3 * t + j]]] += c[j] * coeffD;
    }
    if (with_potential) {
        for (i64 t = a0; t < a1; ++t) {
            Tri T = geom_tri(g, t);
            Fn f; f.kind = 0; f.dip = d;
            V3 v = integrate_fn(&f, order, levels, tol, &T);
            col[g->tri_index[t]] += v.x * coeffP;
        }
    }
}

/* Mesh::solid_angle, mesh.cpp:239-244. */
double orc_mesh_solid_angle(const Geom* g, i64 m, const double* p) {
    V3 P = v3(p[0], p[1], p[2]);
    double s = 0.0;
    for (i64 t = g->tri_off[m]; t < g->tri_off[m + 1]; ++t) {
        Tri T = geom_tri(g, t);
        s += tri_solid_angle(&T, P);
    }
    return s;
}

/* ------------------------------------------------------------------ */
/* Danielsson closest point, danielsson.cpp:26-102                      */
/* ------------------------------------------------------------------ */
static double dpc(V3 p, const V3 P[3], double al[3], int nb, int* idx, int* inside) {
    if (nb == 1) {
        al[idx[0]] = 1.0;
        return nrm(sub(p, P[idx[0]]));
    }
    V3 A0Ai[3];
    for (int i = 1; i < nb; ++i) A0Ai[i] = sub(P[idx[i]], P[idx[0]]);
    V3 A0M = sub(p, P[idx[0]]);
    if (nb == 2) {
        al[idx[1]] = dot(A0M, A0Ai[1]) / dot(A0Ai[1], A0Ai[1]);
        al[idx[0]] = 1.0 - al[idx[1]];
    } else {
        double a00 = dot(A0Ai[1], A0Ai[1]), a10 = dot(A0Ai[1], A0Ai[2]), a11 = dot(A0Ai[2], A0Ai[2]);
        double b0 = dot(A0M, A0Ai[1]), b1 = dot(A0M, A0Ai[2]);
        double d = a00 * a11 - a10 * a10;
        al[idx[1]] = (b0 * a11 - b1 * a10) / d;
        al[idx[2]] = (a00 * b1 - a10 * b0) / d;
        al[idx[0]] = 1.0 - al[idx[1]] - al[idx[2]];
    }
    for (int i = 0; i < nb; ++i)
        if (al[idx[i]] < 0) {
            *inside = 0;
            al[idx[i]] = 0;
            int tmp = idx[i]; idx[i] = idx[nb - 1]; idx[nb - 1] = tmp;
            return dpc(p, P, al, nb - 1, idx, inside);
        }
    V3 MH = scl(-1.0, A0M);
    for (int i = 1; i < nb; ++i) MH = add(MH, scl(al[idx[i]], A0Ai[i]));
    return nrm(MH);
}

/* dist_point_interface restricted to one mesh: first strict minimum over the
 * mesh's triangles in order.  Returns the distance, the global triangle id
 * and the barycentric weights. */
double orc_dist_point_mesh(const Geom* g, i64 m, const double* p, i64* tri, double* alphas) {
    V3 P = v3(p[0], p[1], p[2]);
    double best = 1.79769313486231570815e+308;
    for (i64 t = g->tri_off[m]; t < g->tri_off[m + 1]; ++t) {
        V3 Q[3] = {vert(g, g->tris[3 * t]), vert(g, g->tris[3 * t + 1]), vert(g, g->tris[3 * t + 2])};
        double al[3] = {0, 0, 0};
        int idx[3] = {0, 1, 2}, inside = 1;
        double dist = dpc(P, Q, al, 3, idx, &inside);
        if (dist < best) {
            best = dist; *tri = t;
            alphas[0] = al[0]; alphas[1] = al[1]; alphas[2] = al[2];
        }
    }
    return best;
}

/* Raw kernel probes used by the invariant tests. */
double orc_operator_s(const double* tri9, const double* x) {
    Tri t = make_tri(v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]));
    OpS o; ops_init(&o, &t);
    return ops_eval(&o, v3(x[0], x[1], x[2]));
}
void orc_p1slp(const double* tri9, const double* x, double* out3) {
    Tri t = make_tri(v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]));
    P1S o; p1s_init(&o, &t);
    V3 r = p1s_eval(&o, v3(x[0], x[1], x[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
double orc_solid_angle(const double* tri9, const double* x) {
    Tri t = make_tri(v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]));
    return tri_solid_angle(&t, v3(x[0], x[1], x[2]));
}
/* Surf2VolMat's PartialBlock (operators.h:302-331) of mesh m at point x:
 * row[vertex of T] += cD P1SingleLayerPotential_T(x) for the triangles T of
 * m in mesh order, and row[T] = cS OperatorS_T(x) when with_s. */
void orc_surf2vol_mesh(const Geom* g, i64 m, const double* x, double cD, double cS, int with_s, double* row) {
    const V3 X = v3(x[0], x[1], x[2]);
    for (i64 t = g->tri_off[m]; t < g->tri_off[m + 1]; ++t) {
        Tri T = geom_tri(g, t);
        P1S o; p1s_init(&o, &T);
        const V3 r = p1s_eval(&o, X);
        row[g->vert_index[g->tris[3 * t]]] += cD * r.x;
        row[g->vert_index[g->tris[3 * t + 1]]] += cD * r.y;
        row[g->vert_index[g->tris[3 * t + 2]]] += cD * r.z;
        if (with_s) {
            OpS s; ops_init(&s, &T);
            row[g->tri_index[t]] = cS * ops_eval(&s, X);
        }
    }
}

/* D pair of the D blocks (operators.h:103-106): the P1 double layer of
 * tri_col integrated over tri_row (Gauss), one value per vertex of tri_col. */
void orc_integrate_d_pair(const double* tri_row, const double* tri_col, int order, double* out3) {
    Tri r = make_tri(v3(tri_row[0], tri_row[1], tri_row[2]), v3(tri_row[3], tri_row[4], tri_row[5]),
                     v3(tri_row[6], tri_row[7], tri_row[8]));
    Tri c = make_tri(v3(tri_col[0], tri_col[1], tri_col[2]), v3(tri_col[3], tri_col[4], tri_col[5]),
                     v3(tri_col[6], tri_col[7], tri_col[8]));
    P1S o; p1s_init(&o, &c);
    V3 tot = integrate_D(&o, &r, order);
    out3[0] = tot.x; out3[1] = tot.y; out3[2] = tot.z;
}

double orc_integrate_s_pair(const double* tri_a, const double* tri_b, int order) {
    Tri a = make_tri(v3(tri_a[0], tri_a[1], tri_a[2]), v3(tri_a[3], tri_a[4], tri_a[5]), v3(tri_a[6], tri_a[7], tri_a[8]));
    Tri b = make_tri(v3(tri_b[0], tri_b[1], tri_b[2]), v3(tri_b[3], tri_b[4], tri_b[5]), v3(tri_b[6], tri_b[7], tri_b[8]));
    OpS o; ops_init(&o, &a);
    return integrate_S(&o, &b, order);
}

/* The adaptive pair integrals of a HeadMat with Integrator(order, levels,
 * tol): kind 2 = the S integral of OperatorS(a) over b (out3[0]); kind 3 =
 * the D integral over a of P1SLP(b) (out3[0..2]). */
void orc_integrate_pair_adaptive(const double* tri_a, const double* tri_b, int kind, int order, unsigned levels,
                                 double tol, double* out3) {
    Tri a = make_tri(v3(tri_a[0], tri_a[1], tri_a[2]), v3(tri_a[3], tri_a[4], tri_a[5]), v3(tri_a[6], tri_a[7], tri_a[8]));
    Tri b = make_tri(v3(tri_b[0], tri_b[1], tri_b[2]), v3(tri_b[3], tri_b[4], tri_b[5]), v3(tri_b[6], tri_b[7], tri_b[8]));
    OpS o; P1S p;
    Fn f = {0};
    f.kind = kind;
    V3 r;
    if (kind == 2) {
        ops_init(&o, &a);
        f.ops = &o;
        r = integrate_fn(&f, order, levels, tol, &b);
    } else {
        p1s_init(&p, &b);
        f.p1s = &p;
        r = integrate_fn(&f, order, levels, tol, &a);
    }
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

/* Sample of the HeadMat work for the CPU-baseline timing: rows [r0, r1) of
 * mesh m1 against all triangles of m2, S pairs (mode 1, analytic side m1 as
 * s_block) or D pairs (mode 2, as d_block), accumulated into a scratch sum
 * so that nothing is optimised away.  Returns the number of pairs done. */
double orc_sample_rows(const Geom* g, i64 m1, i64 m2, int order, i64 r0, i64 r1, int mode, double* sink) {
    const i64 a0 = g->tri_off[m1];
    const i64 b0 = g->tri_off[m2], b1 = g->tri_off[m2 + 1];
    double acc = 0.0;
    #pragma omp parallel for reduction(+:acc) schedule(dynamic, 1)
    for (i64 i = a0 + r0; i < a0 + r1; ++i) {
        Tri t1 = geom_tri(g, i);
        OpS o; ops_init(&o, &t1);
        for (i64 j = b0; j < b1; ++j) {
            Tri t2 = geom_tri(g, j);
            if (mode == 1) {
                acc += integrate_S(&o, &t2, order);
            } else {
                P1S p; p1s_init(&p, &t2);
                V3 v = integrate_D(&p, &t1, order);
                acc += v.x + v.y + v.z;
            }
        }
    }
    *sink = acc;
    return (double)(r1 - r0) * (double)(b1 - b0);
}
