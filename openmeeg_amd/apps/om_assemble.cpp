// om_assemble: the head matrices of the reference's command-line tool
// (apps/assemble.cpp:30-330) for the EEG / MEG forward path, computed by
// libopenmeeg_amd on the GPU and written in the format of the output
// file's suffix (.mat, .bin, .txt).
//
//   -HeadMat, -HM, -hm                 geometry conductivity output
//   -DipSourceMat, -DSM, -dsm          geometry conductivity dipoles output [domain]
//   -DipSourceMatNoAdapt, -DSMNA, -dsmna   (the same without adaptive integration)
//   -SurfSourceMat, -SSM, -ssm         geometry conductivity sourcemesh output
//   -Head2EEGMat, -H2EM, -h2em         geometry conductivity electrodes output
//   -Head2MEGMat, -H2MM, -h2mm         geometry conductivity squids output
//   -DipSource2MEGMat, -DS2MM, -ds2mm  dipoles squids output
//   -SurfSource2MEGMat, -SS2MM, -ss2mm sourcemesh squids output
//   -EITSourceMat, -EITSM, -EITsm               geometry conductivity electrodes output
//   -Head2InternalPotMat, -H2IPM, -h2ipm        geometry conductivity points output
//   -DipSource2InternalPotMat, -DS2IPM, -ds2ipm geometry conductivity dipoles points output [domain]
//   -old-ordering (with any of the above): unknowns numbered mesh by mesh,
//   (V1, p1, V2, p2, V3) instead of (V1, V2, V3, p1, p2) (apps/assemble.cpp:41)
#include <algorithm>
#include <cstring>
#include <iostream>

#include "cli.hpp"

using namespace cli;

namespace {

bool g_old_ordering = false;   // -old-ordering

bool is(const char* a, std::initializer_list<const char*> names) {
    for (const char* n : names)
        if (!std::strcmp(a, n)) return true;
    return false;
}

void need_args(int argc, int k, int count, const char* opt) {
    if (argc < k + 1 + count) throw std::runtime_error(std::string(opt) + " needs " + std::to_string(count) + " file arguments");
}

void head_mat(const std::string& geom, const std::string& cond, const std::string& out) {
    Geom g(geom, cond);
    const int64_t n = g.n();
    Dev<double> H((size_t)(n * n));
    check(omh_head_mat(g.g, 3, H.p, n, OMH_ORDER_REFERENCE, nullptr));
    if (g_old_ordering) {   // packed upper storage (symmatrix.h:52-62) in the old numbering
        const std::vector<int64_t> perm = g.old_order();
        std::vector<int64_t> inv((size_t)n);
        for (int64_t u = 0; u < n; ++u) inv[(size_t)perm[(size_t)u]] = u;
        const std::vector<double> h = H.down();
        std::vector<double> P((size_t)(n * (n + 1) / 2));
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i <= j; ++i) P[(size_t)(i + j * (j + 1) / 2)] = h[(size_t)(inv[i] + inv[j] * n)];
        write_symmetric(out, P, n);
        return;
    }
    Dev<double> P((size_t)(n * (n + 1) / 2));
    check(omh_sym_pack(H.p, n, n, n, OMH_ORDER_REFERENCE, P.p, nullptr));
    write_symmetric(out, P.down(), n);
}

// The unknown rows / columns of a matrix in the numbering the options ask for.
Full rows_out(const Geom& g, const Full& m) { return g_old_ordering ? permute_rows(m, g.old_order()) : m; }
Full cols_out(const Geom& g, const Full& m) { return g_old_ordering ? permute_cols(m, g.old_order()) : m; }

void dip_source_mat(const std::string& geom, const std::string& cond, const std::string& dipf, const std::string& out,
                    const char* domain, int levels) {
    Geom g(geom, cond);
    int64_t nd = 0;
    const std::vector<double> dip = read_dipoles(dipf, nd);
    const int64_t n = g.n();
    std::vector<int32_t> dom;
    if (domain) {
        int32_t k = -1;
        check(omh_geometry_domain_index(g.g, domain, &k));
        dom.assign((size_t)nd, k);
    }
    Dev<double> dd(dip.data(), dip.size());
    Dev<double> D((size_t)(n * nd));
    check(omh_dip_source_mat(g.g, dd.p, nd, domain ? dom.data() : nullptr, 3, levels, 0.001, D.p, n,
                             OMH_ORDER_REFERENCE, nullptr));
    Full m;
    m.nlin = n;
    m.ncol = nd;
    m.a = D.down();
    write_full(out, rows_out(g, m));
}

void surf_source_mat(const std::string& geom, const std::string& cond, const std::string& meshf,
                     const std::string& out) {
    Geom g(geom, cond);
    omh_mesh* mesh = nullptr;
    check(omh_mesh_load(meshf.c_str(), &mesh));
    int64_t ms[2];
    check(omh_mesh_sizes(mesh, ms));
    const int64_t n = g.n();
    Dev<double> D((size_t)(n * ms[0]));
    const int rc = omh_surf_source_mat(g.g, mesh, 3, 0, D.p, n, OMH_ORDER_REFERENCE, nullptr);
    omh_mesh_free(mesh);
    check(rc);
    Full m;
    m.nlin = n;
    m.ncol = ms[0];
    m.a = D.down();
    write_full(out, rows_out(g, m));
}

void head2eeg(const std::string& geom, const std::string& cond, const std::string& elf, const std::string& out) {
    Geom g(geom, cond);
    const Sensors s = read_sensors(elf);
    const int64_t ne = s.npts();
    Dev<double> de(s.pos.data(), s.pos.size());
    std::vector<int64_t> cols((size_t)(3 * ne));
    std::vector<double> vals((size_t)(3 * ne));
    check(omh_head2eeg_mat(g.g, de.p, ne, cols.data(), vals.data(), nullptr));
    // SparseMatrix(i, col) = val, a map: a repeated (i, col) keeps the last value
    if (g_old_ordering) {
        const std::vector<int64_t> perm = g.old_order();
        for (auto& c : cols) c = perm[(size_t)c];
    }
    std::map<std::pair<int64_t, int64_t>, double> e;
    for (int64_t i = 0; i < ne; ++i)
        for (int k = 0; k < 3; ++k) e[{i, cols[3 * i + k]}] = vals[3 * i + k];
    Sparse sp;
    sp.nlin = ne;
    sp.ncol = g.n();
    for (const auto& x : e) {
        sp.rows.push_back(x.first.first);
        sp.cols.push_back(x.first.second);
        sp.vals.push_back(x.second);
    }
    write_sparse(out, sp);
}

std::vector<double> dirs_of(const Sensors& s, const std::string& path) {
    if (!s.oriented) throw std::runtime_error(path + ": MEG sensors need orientations (6 or 7 columns)");
    return s.dir;
}

void head2meg(const std::string& geom, const std::string& cond, const std::string& sqf, const std::string& out) {
    Geom g(geom, cond);
    const Sensors s = read_sensors(sqf);
    const int64_t np = s.npts(), n = g.n();
    Dev<double> dp(s.pos.data(), s.pos.size());
    const std::vector<double> dirs = dirs_of(s, sqf);
    Dev<double> dd(dirs.data(), dirs.size());
    Dev<double> P((size_t)(np * n));
    check(omh_head2meg_mat(g.g, dp.p, dd.p, np, P.p, np, OMH_ORDER_REFERENCE, nullptr));
    write_full(out, cols_out(g, sensor_sum(s, P.down(), n)));
}

void dip2meg(const std::string& dipf, const std::string& sqf, const std::string& out) {
    int64_t nd = 0;
    const std::vector<double> dip = read_dipoles(dipf, nd);
    const Sensors s = read_sensors(sqf);
    const int64_t np = s.npts();
    Dev<double> dp(s.pos.data(), s.pos.size());
    const std::vector<double> dirs = dirs_of(s, sqf);
    Dev<double> dd(dirs.data(), dirs.size());
    Dev<double> ddip(dip.data(), dip.size());
    Dev<double> P((size_t)(np * nd));
    check(omh_dip_source2meg_mat(ddip.p, nd, dp.p, dd.p, np, P.p, np, nullptr));
    write_full(out, sensor_sum(s, P.down(), nd));
}

void surf2meg(const std::string& meshf, const std::string& sqf, const std::string& out) {
    omh_mesh* mesh = nullptr;
    check(omh_mesh_load(meshf.c_str(), &mesh));
    int64_t ms[2];
    check(omh_mesh_sizes(mesh, ms));
    const Sensors s = read_sensors(sqf);
    const int64_t np = s.npts();
    Dev<double> dp(s.pos.data(), s.pos.size());
    const std::vector<double> dirs = dirs_of(s, sqf);
    Dev<double> dd(dirs.data(), dirs.size());
    Dev<double> P((size_t)(np * ms[0]));
    const int rc = omh_surf_source2meg_mat(mesh, dp.p, dd.p, np, P.p, np, nullptr);
    omh_mesh_free(mesh);
    check(rc);
    write_full(out, sensor_sum(s, P.down(), ms[0]));
}

// Points: a full n x 3 (or more) matrix file (Matrix(filename)), row-major.
std::vector<double> read_points(const std::string& path, int64_t& np) {
    const Full m = read_full(path);
    if (m.ncol < 3) throw std::runtime_error(path + ": points need 3 columns");
    np = m.nlin;
    std::vector<double> p((size_t)(np * 3));
    for (int64_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) p[i * 3 + c] = m(i, c);
    return p;
}

void head2ip(const std::string& geom, const std::string& cond, const std::string& ptf, const std::string& out) {
    Geom g(geom, cond);
    int64_t np = 0;
    const std::vector<double> pts = read_points(ptf, np);
    const int64_t n = g.n(), ld = std::max<int64_t>(1, np);
    Dev<double> dp(pts.data(), pts.size());
    Dev<double> M((size_t)(ld * n));
    int64_t nk = 0;
    check(omh_surf2vol_mat(g.g, dp.p, np, M.p, ld, OMH_ORDER_REFERENCE, nullptr, &nk, nullptr));
    const std::vector<double> h = M.down();
    Full m;
    m.nlin = nk;
    m.ncol = n;
    m.a.resize((size_t)(nk * n));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < nk; ++i) m(i, j) = h[i + j * ld];
    write_full(out, cols_out(g, m));
}

void dip2ip(const std::string& geom, const std::string& cond, const std::string& dipf, const std::string& ptf,
            const std::string& out, const char* domain) {
    Geom g(geom, cond);
    int64_t nd = 0, np = 0;
    const std::vector<double> dip = read_dipoles(dipf, nd);
    const std::vector<double> pts = read_points(ptf, np);
    const int64_t ld = std::max<int64_t>(1, np);
    std::vector<int32_t> dom;
    if (domain) {
        int32_t k = -1;
        check(omh_geometry_domain_index(g.g, domain, &k));
        dom.assign((size_t)nd, k);
    }
    Dev<double> dd(dip.data(), dip.size()), dp(pts.data(), pts.size());
    Dev<double> M((size_t)(ld * std::max<int64_t>(1, nd)));
    int64_t nk = 0;
    check(omh_dip_source2internal_pot_mat(g.g, dd.p, nd, domain ? dom.data() : nullptr, dp.p, np, M.p, ld, nullptr,
                                          &nk, nullptr));
    const std::vector<double> h = M.down();
    Full m;
    m.nlin = nk;
    m.ncol = nd;
    m.a.resize((size_t)(nk * nd));
    for (int64_t j = 0; j < nd; ++j)
        for (int64_t i = 0; i < nk; ++i) m(i, j) = h[i + j * ld];
    write_full(out, m);
}

void eit_source_mat(const std::string& geom, const std::string& cond, const std::string& elf, const std::string& out) {
    Geom g(geom, cond);
    const Sensors s = read_sensors(elf);
    const int64_t ne = s.npts(), n = g.n();
    Dev<double> dp(s.pos.data(), s.pos.size());
    Dev<double> D((size_t)(n * ne));
    check(omh_eit_source_mat(g.g, dp.p, s.radius.data(), ne, 3, D.p, n, OMH_ORDER_REFERENCE, nullptr));
    Full m;
    m.nlin = n;
    m.ncol = ne;
    m.a = D.down();
    write_full(out, rows_out(g, m));
}

void help(const char* cmd) {
    std::cout << cmd << " [-option] [filepaths...]\n\n"
              << "   -HeadMat, -HM, -hm :                 geometry conductivity output\n"
              << "   -DipSourceMat, -DSM, -dsm :          geometry conductivity dipoles output [domain]\n"
              << "   -DipSourceMatNoAdapt, -DSMNA, -dsmna: geometry conductivity dipoles output [domain]\n"
              << "   -SurfSourceMat, -SSM, -ssm :         geometry conductivity sourcemesh output\n"
              << "   -Head2EEGMat, -H2EM, -h2em :         geometry conductivity electrodes output\n"
              << "   -Head2MEGMat, -H2MM, -h2mm :         geometry conductivity squids output\n"
              << "   -DipSource2MEGMat, -DS2MM, -ds2mm :  dipoles squids output\n"
              << "   -SurfSource2MEGMat, -SS2MM, -ss2mm : sourcemesh squids output\n"
              << "   -EITSourceMat, -EITSM, -EITsm :      geometry conductivity electrodes output\n"
              << "   -Head2InternalPotMat, -H2IPM, -h2ipm : geometry conductivity points output\n"
              << "   -DipSource2InternalPotMat, -DS2IPM, -ds2ipm : geometry conductivity dipoles points output [domain]\n"
              << "   -old-ordering : with any of the above, unknowns numbered (V1, p1, V2, p2, V3)\n"
              << "Output format from the suffix: .mat (MAT 7.3), .bin, .txt\n";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || is(argv[1], {"-h", "--help", "-help"})) {
        help(argv[0]);
        return 0;
    }
    try {
        for (int k = 1; k < argc; ++k)
            if (is(argv[k], {"-old-ordering"})) g_old_ordering = true;
        int done = 0;
        for (int k = 1; k < argc; ++k) {
            const char* o = argv[k];
            if (o[0] != '-') continue;
            if (is(o, {"-old-ordering"})) continue;
            if (is(o, {"-HeadMat", "-HM", "-hm"})) {
                need_args(argc, k, 3, o);
                head_mat(argv[k + 1], argv[k + 2], argv[k + 3]);
            } else if (is(o, {"-DipSourceMat", "-DSM", "-dsm", "-DipSourceMatNoAdapt", "-DSMNA", "-dsmna"})) {
                need_args(argc, k, 4, o);
                const bool noadapt = is(o, {"-DipSourceMatNoAdapt", "-DSMNA", "-dsmna"});
                const char* domain = (k + 5 < argc && argv[k + 5][0] != '-') ? argv[k + 5] : nullptr;
                dip_source_mat(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4], domain, noadapt ? 0 : 10);
            } else if (is(o, {"-SurfSourceMat", "-SSM", "-ssm"})) {
                need_args(argc, k, 4, o);
                surf_source_mat(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4]);
            } else if (is(o, {"-Head2EEGMat", "-H2EM", "-h2em"})) {
                need_args(argc, k, 4, o);
                head2eeg(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4]);
            } else if (is(o, {"-Head2MEGMat", "-H2MM", "-h2mm"})) {
                need_args(argc, k, 4, o);
                head2meg(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4]);
            } else if (is(o, {"-DipSource2MEGMat", "-DS2MM", "-ds2mm"})) {
                need_args(argc, k, 3, o);
                dip2meg(argv[k + 1], argv[k + 2], argv[k + 3]);
            } else if (is(o, {"-SurfSource2MEGMat", "-SS2MM", "-ss2mm"})) {
                need_args(argc, k, 3, o);
                surf2meg(argv[k + 1], argv[k + 2], argv[k + 3]);
            } else if (is(o, {"-EITSourceMat", "-EITSM", "-EITsm"})) {
                need_args(argc, k, 4, o);
                eit_source_mat(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4]);
            } else if (is(o, {"-Head2InternalPotMat", "-H2IPM", "-h2ipm"})) {
                need_args(argc, k, 4, o);
                head2ip(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4]);
            } else if (is(o, {"-DipSource2InternalPotMat", "-DS2IPM", "-ds2ipm"})) {
                need_args(argc, k, 5, o);
                const char* domain = (k + 6 < argc && argv[k + 6][0] != '-') ? argv[k + 6] : nullptr;
                dip2ip(argv[k + 1], argv[k + 2], argv[k + 3], argv[k + 4], argv[k + 5], domain);
            } else {
                throw std::runtime_error(std::string("option ") + o + " is not supported by this build");
            }
            ++done;
            break;   // one matrix per call (assert_non_conflicting_options)
        }
        if (!done) {
            help(argv[0]);
            return 1;
        }
    } catch (const std::exception& e) {
        std::cerr << argv[0] << ": " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
