// om_gain: the gain matrices of the reference's command-line tool
// (apps/gain.cpp:60-170) on the GPU, FP64 throughout.
//
//   -EEG           HeadMatInv SourceMat Head2EEGMat output            (gain.h:69-76)
//   -EEGadjoint    geometry conductivity dipoles HeadMat Head2EEGMat output   (gain.h:78-89)
//   -MEG           HeadMatInv SourceMat Head2MEGMat Source2MEGMat output      (gain.h:60-67)
//   -MEGadjoint    geometry conductivity dipoles HeadMat Head2MEGMat Source2MEGMat output
//   -EEGMEGadjoint geometry conductivity dipoles HeadMat Head2EEGMat Head2MEGMat Source2MEGMat
//                  outputEEG outputMEG                                       (gain.h:105-128)
//   -InternalPotential, -IP     HeadMatInv SourceMat Head2IPMat Source2IPMat output (gain.h:138-144)
//   -EITInternalPotential, -EITIP HeadMatInv SourceMat Head2IPMat output         (gain.h:146-152)
// The adjoint forms factor the HeadMat once (pivot-free signed Cholesky in
// the triangles-first unknown order of the geometry) and solve with the
// sensor rows; the dipole source matrix is assembled on the device.
#include <cstring>
#include <iostream>
#include <memory>

#include "cli.hpp"

using namespace cli;

namespace {

// Dense sensors x n matrix of a Head2EEGMat file.
Full dense_of(const Sparse& s) {
    Full m;
    m.nlin = s.nlin;
    m.ncol = s.ncol;
    m.a.assign((size_t)(s.nlin * s.ncol), 0.0);
    for (size_t k = 0; k < s.vals.size(); ++k) {
        if (s.rows[k] < 0 || s.rows[k] >= s.nlin || s.cols[k] < 0 || s.cols[k] >= s.ncol)
            throw std::runtime_error("sparse matrix entry out of range");
        m(s.rows[k], s.cols[k]) = s.vals[k];
    }
    return m;
}

Full product(const Dev<double>& A, int64_t m, int64_t k, const Dev<double>& B, int64_t n, int ta = 0) {
    Dev<double> C((size_t)(m * n));
    check(omh_gemm_f64(ta, 0, m, n, k, 1.0, A.p, ta ? k : m, B.p, k, 0.0, C.p, m, nullptr));
    Full out;
    out.nlin = m;
    out.ncol = n;
    out.a = C.down();
    return out;
}

// (sensors x n) x H^-1 x (n x ns) with the inverse read from a file.
Full gain_with_inverse(const std::string& hinv, const Full& S2N, const Full& src) {
    int64_t n = 0;
    const std::vector<double> packed = read_symmetric(hinv, n);
    if (S2N.ncol != n || src.nlin != n) throw std::runtime_error("matrix sizes do not match the HeadMatInv");
    Dev<double> P(packed.data(), packed.size());
    Dev<double> Hi((size_t)(n * n));
    check(omh_sym_unpack(P.p, n, n, Hi.p, n, OMH_ORDER_REFERENCE, nullptr));
    Dev<double> M(S2N.a.data(), S2N.a.size());
    Dev<double> MH((size_t)(S2N.nlin * n));
    check(omh_gemm_f64(0, 0, S2N.nlin, n, n, 1.0, M.p, S2N.nlin, Hi.p, n, 0.0, MH.p, S2N.nlin, nullptr));
    Dev<double> D(src.a.data(), src.a.size());
    return product(MH, S2N.nlin, n, D, src.ncol);
}

// The adjoint solve: Y = H^-1 R^T for the sensor rows R (r x n, reference
// column order), returned on the device in the triangles-first row order,
// and the dipole source matrix in the same order.
struct Adjoint {
    int64_t n = 0, nd = 0;
    std::unique_ptr<Dev<double>> Y, D;
};

Adjoint adjoint(const std::string& geom, const std::string& cond, const std::string& dipf, const std::string& hm,
                const std::vector<const Full*>& rows) {
    Geom g(geom, cond);
    Adjoint a;
    const int64_t n = g.n(), nv = g.nv();
    a.n = n;
    int64_t nh = 0;
    const std::vector<double> packed = read_symmetric(hm, nh);
    if (nh != n) throw std::runtime_error("the HeadMat does not match the geometry");
    Dev<double> P(packed.data(), packed.size());
    Dev<double> H((size_t)(n * n));
    check(omh_sym_unpack(P.p, n, nv, H.p, n, OMH_ORDER_TRIANGLES_FIRST, nullptr));
    Dev<unsigned char> W((size_t)omh_factor_workspace_bytes(n, 8));
    int32_t info = 0;
    check(omh_sym_factor_f64(H.p, n, n, W.p, &info, nullptr));
    if (info != 0) throw std::runtime_error("the pivot-free factorisation broke down at pivot " + std::to_string(info - 1));
    int64_t r = 0;
    for (const Full* f : rows) {
        if (f->ncol != n) throw std::runtime_error("a sensor matrix does not match the geometry");
        r += f->nlin;
    }
    std::vector<double> rhs((size_t)(n * r), 0.0);
    int64_t c0 = 0;
    for (const Full* f : rows) {
        for (int64_t i = 0; i < f->nlin; ++i)
            for (int64_t u = 0; u < n; ++u) {
                const int64_t t = u >= nv ? u - nv : u + (n - nv);   // triangles-first row
                rhs[t + (c0 + i) * n] = (*f)(i, u);
            }
        c0 += f->nlin;
    }
    a.Y = std::make_unique<Dev<double>>(rhs.data(), rhs.size());
    check(omh_sym_solve_f64(H.p, n, n, W.p, a.Y->p, n, r, nullptr));
    const std::vector<double> dip = read_dipoles(dipf, a.nd);
    Dev<double> dd(dip.data(), dip.size());
    a.D = std::make_unique<Dev<double>>((size_t)(n * a.nd));
    check(omh_dip_source_mat(g.g, dd.p, a.nd, nullptr, 3, 10, 0.001, a.D->p, n, OMH_ORDER_TRIANGLES_FIRST, nullptr));
    return a;
}

// G = Y(:, c0:c0+m)^T D (+ S2M).
Full adjoint_gain(const Adjoint& a, int64_t c0, int64_t m, const Full* s2m) {
    Dev<double> G((size_t)(m * a.nd));
    double beta = 0.0;
    if (s2m) {
        if (s2m->nlin != m || s2m->ncol != a.nd) throw std::runtime_error("Source2MEGMat size mismatch");
        G.up(s2m->a.data());
        beta = 1.0;
    }
    check(omh_gemm_f64(1, 0, m, a.nd, a.n, 1.0, a.Y->p + c0 * a.n, a.n, a.D->p, a.n, beta, G.p, m, nullptr));
    Full out;
    out.nlin = m;
    out.ncol = a.nd;
    out.a = G.down();
    return out;
}

bool is(const char* a, const char* b) { return !std::strcmp(a, b); }

void need_args(int argc, int k, int count, const char* opt) {
    if (argc < k + 1 + count) throw std::runtime_error(std::string(opt) + " needs " + std::to_string(count) + " file arguments");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << argv[0] << " [-EEG | -EEGadjoint | -MEG | -MEGadjoint | -EEGMEGadjoint] [filepaths...]\n";
        return 0;
    }
    try {
        const char* o = argv[1];
        const int k = 1;
        if (is(o, "-EEG")) {
            need_args(argc, k, 4, o);
            const Full E = dense_of(read_sparse(argv[4]));
            write_full(argv[5], gain_with_inverse(argv[2], E, read_full(argv[3])));
        } else if (is(o, "-MEG")) {
            need_args(argc, k, 5, o);
            const Full M = read_full(argv[4]);
            Full G = gain_with_inverse(argv[2], M, read_full(argv[3]));
            const Full S = read_full(argv[5]);
            if (S.nlin != G.nlin || S.ncol != G.ncol) throw std::runtime_error("Source2MEGMat size mismatch");
            for (size_t q = 0; q < G.a.size(); ++q) G.a[q] = S.a[q] + G.a[q];
            write_full(argv[6], G);
        } else if (is(o, "-InternalPotential") || is(o, "-IP") || is(o, "-ip")) {
            need_args(argc, k, 5, o);   // HeadMatInv SourceMat Head2IPMat Source2IPMat output (gain.h:138-144)
            Full G = gain_with_inverse(argv[2], read_full(argv[4]), read_full(argv[3]));
            const Full S = read_full(argv[5]);
            if (S.nlin != G.nlin || S.ncol != G.ncol) throw std::runtime_error("Source2IPMat size mismatch");
            for (size_t q = 0; q < G.a.size(); ++q) G.a[q] = S.a[q] + G.a[q];
            write_full(argv[6], G);
        } else if (is(o, "-EITInternalPotential") || is(o, "-EITIP") || is(o, "-eitip")) {
            need_args(argc, k, 4, o);   // HeadMatInv SourceMat Head2IPMat output (gain.h:146-152)
            write_full(argv[5], gain_with_inverse(argv[2], read_full(argv[4]), read_full(argv[3])));
        } else if (is(o, "-EEGadjoint")) {
            need_args(argc, k, 6, o);
            const Full E = dense_of(read_sparse(argv[6]));
            const Adjoint a = adjoint(argv[2], argv[3], argv[4], argv[5], {&E});
            write_full(argv[7], adjoint_gain(a, 0, E.nlin, nullptr));
        } else if (is(o, "-MEGadjoint")) {
            need_args(argc, k, 7, o);
            const Full M = read_full(argv[6]);
            const Full S = read_full(argv[7]);
            const Adjoint a = adjoint(argv[2], argv[3], argv[4], argv[5], {&M});
            write_full(argv[8], adjoint_gain(a, 0, M.nlin, &S));
        } else if (is(o, "-EEGMEGadjoint")) {
            need_args(argc, k, 9, o);
            const Full E = dense_of(read_sparse(argv[6]));
            const Full M = read_full(argv[7]);
            const Full S = read_full(argv[8]);
            const Adjoint a = adjoint(argv[2], argv[3], argv[4], argv[5], {&E, &M});
            write_full(argv[9], adjoint_gain(a, 0, E.nlin, nullptr));
            write_full(argv[10], adjoint_gain(a, E.nlin, M.nlin, &S));
        } else {
            throw std::runtime_error(std::string("option ") + o + " is not supported by this build");
        }
    } catch (const std::exception& e) {
        std::cerr << argv[0] << ": " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
