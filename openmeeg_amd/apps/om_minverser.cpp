// om_minverser: inverse of a HeadMat file (apps/minverser.cpp:30-60,
// SymMatrix::invert, symmatrix.h:280-297) on the GPU.
//
// The pivot-free factorisation needs the triangles-first unknown order.
// Without a geometry the split between vertex and triangle unknowns is read
// from the matrix: in the reference order a HeadMat has its vertex unknowns
// first, with negative diagonal entries (N blocks, deflated), then the
// triangle unknowns with positive ones (S blocks).  A matrix without that
// sign pattern, or whose factorisation breaks down, is refused.
#include <iostream>

#include "cli.hpp"

using namespace cli;

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cout << argv[0] << " HeadMat HeadMatInv\n  (formats from the suffixes: .mat, .bin, .txt)\n";
        return argc == 1 ? 0 : 1;
    }
    try {
        int64_t n = 0;
        const std::vector<double> packed = read_symmetric(argv[1], n);
        int64_t nv = 0;
        while (nv < n && packed[nv + nv * (nv + 1) / 2] < 0.0) ++nv;
        for (int64_t i = nv; i < n; ++i)
            if (!(packed[i + i * (i + 1) / 2] > 0.0))
                throw std::runtime_error("not a HeadMat: its diagonal is not negative on the vertex unknowns then "
                                         "positive on the triangle unknowns");
        Dev<double> P(packed.data(), packed.size());
        Dev<double> H((size_t)(n * n));
        check(omh_sym_unpack(P.p, n, nv, H.p, n, OMH_ORDER_TRIANGLES_FIRST, nullptr));
        const int64_t wb = omh_factor_workspace_bytes(n, 8);
        Dev<unsigned char> W((size_t)wb);
        int32_t info = 0;
        check(omh_sym_factor_f64(H.p, n, n, W.p, &info, nullptr));
        if (info != 0) throw std::runtime_error("the pivot-free factorisation broke down at pivot " + std::to_string(info - 1));
        std::vector<double> eye((size_t)(n * n), 0.0);
        for (int64_t i = 0; i < n; ++i) eye[i + i * n] = 1.0;
        Dev<double> X(eye.data(), eye.size());
        check(omh_sym_solve_f64(H.p, n, n, W.p, X.p, n, n, nullptr));
        check(omh_sym_pack(X.p, n, nv, n, OMH_ORDER_TRIANGLES_FIRST, P.p, nullptr));
        write_symmetric(argv[2], P.down(), n);
    } catch (const std::exception& e) {
        std::cerr << argv[0] << ": " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
