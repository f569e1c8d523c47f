// Shared host code of the command-line tools (om_assemble, om_minverser,
// om_gain): matrices in the reference's file formats (libopenmeeg_amd_io),
// sensor files, device buffers and the C ABI of libopenmeeg_amd.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/openmeeg_amd.h"
#include "../../include/openmeeg_amd_io.h"

namespace cli {

inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(omh_last_error());
}
inline void check_io(int rc) {
    if (rc != 0) throw std::runtime_error(omh_io_last_error());
}
inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer of doubles (or bytes), freed on scope exit.
template <class T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    explicit Dev(size_t count) : n(count) {
        check_hip(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (count ? count : 1)), "hipMalloc");
    }
    Dev(const T* host, size_t count) : Dev(count) { up(host); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    void up(const T* host) {
        if (n) check_hip(hipMemcpy(p, host, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    std::vector<T> down() const {
        std::vector<T> h(n);
        if (n) check_hip(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
};

// Dense column-major host matrix.
struct Full {
    int64_t nlin = 0, ncol = 0;
    std::vector<double> a;
    double& operator()(int64_t i, int64_t j) { return a[i + j * nlin]; }
    double operator()(int64_t i, int64_t j) const { return a[i + j * nlin]; }
};

struct Sparse {
    int64_t nlin = 0, ncol = 0;
    std::vector<int64_t> rows, cols;
    std::vector<double> vals;
};

inline omh_linop_info info_of(const std::string& path) {
    omh_linop_info li{};
    check_io(omh_io_info(path.c_str(), &li));
    return li;
}

inline Full read_full(const std::string& path) {
    const omh_linop_info li = info_of(path);
    if (li.storage != OMH_IO_FULL) throw std::runtime_error(path + ": not a full matrix");
    Full m;
    m.nlin = li.nlin;
    m.ncol = li.ncol;
    m.a.resize((size_t)(li.nlin * li.ncol));
    check_io(omh_io_read(path.c_str(), &li, m.a.data(), nullptr, nullptr));
    return m;
}

inline std::vector<double> read_symmetric(const std::string& path, int64_t& n) {
    const omh_linop_info li = info_of(path);
    if (li.storage != OMH_IO_SYMMETRIC) throw std::runtime_error(path + ": not a symmetric matrix");
    n = li.nlin;
    std::vector<double> p((size_t)(n * (n + 1) / 2));
    check_io(omh_io_read(path.c_str(), &li, p.data(), nullptr, nullptr));
    return p;
}

inline Sparse read_sparse(const std::string& path) {
    const omh_linop_info li = info_of(path);
    if (li.storage != OMH_IO_SPARSE) throw std::runtime_error(path + ": not a sparse matrix");
    Sparse s;
    s.nlin = li.nlin;
    s.ncol = li.ncol;
    s.rows.resize(li.nnz);
    s.cols.resize(li.nnz);
    s.vals.resize(li.nnz);
    check_io(omh_io_read(path.c_str(), &li, s.vals.data(), s.rows.data(), s.cols.data()));
    return s;
}

inline void write_full(const std::string& path, const Full& m) {
    omh_linop_info li{};
    li.storage = OMH_IO_FULL;
    li.dimension = 2;
    li.nlin = m.nlin;
    li.ncol = m.ncol;
    check_io(omh_io_write(path.c_str(), &li, m.a.data(), nullptr, nullptr));
}

inline void write_symmetric(const std::string& path, const std::vector<double>& packed, int64_t n) {
    omh_linop_info li{};
    li.storage = OMH_IO_SYMMETRIC;
    li.dimension = 2;
    li.nlin = li.ncol = n;
    check_io(omh_io_write(path.c_str(), &li, packed.data(), nullptr, nullptr));
}

inline void write_sparse(const std::string& path, const Sparse& s) {
    omh_linop_info li{};
    li.storage = OMH_IO_SPARSE;
    li.dimension = 2;
    li.nlin = s.nlin;
    li.ncol = s.ncol;
    li.nnz = (int64_t)s.vals.size();
    check_io(omh_io_write(path.c_str(), &li, s.vals.data(), s.rows.data(), s.cols.data()));
}

// Sensors::load (sensors.cpp:36-143) of an electrode or MEG file: optional
// label column (a first token that is not a number with one '.'),
// positions, orientations (6 columns), weights (7 columns); points with the
// same label form one sensor, labels in order of first appearance.
struct Sensors {
    std::vector<double> pos, dir, weight;   // npts x 3 row-major, npts x 3, npts
    std::vector<double> radius;             // EIT electrodes (4 columns), else 0
    std::vector<int64_t> sensor;            // sensor index of each point
    int64_t nsensors = 0;
    bool oriented = false;
    int64_t npts() const { return (int64_t)weight.size(); }
};

inline Sensors read_sensors(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open sensors file " + path);
    std::vector<std::vector<std::string>> rows;
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::vector<std::string> tok;
        std::string t;
        while (ss >> t) tok.push_back(t);
        if (!tok.empty() && tok[0][0] != '#') rows.push_back(tok);
    }
    if (rows.empty()) throw std::runtime_error("empty sensors file " + path);
    for (const auto& r : rows)
        if (r.size() != rows[0].size())
            throw std::runtime_error("Problem while reading Sensors file: each line should have the same number of elements");
    // labelled unless some row's first token is a number with one '.'
    // (sensors.cpp:67-69 scans every row)
    bool labeled = true;
    for (const auto& r : rows)
        if (std::count(r[0].begin(), r[0].end(), '.') == 1) labeled = false;
    const size_t off = labeled ? 1 : 0;
    const size_t ncol = rows[0].size() - off;
    if (ncol != 3 && ncol != 4 && ncol != 6 && ncol != 7)
        throw std::runtime_error(path + ": sensors need 3, 4, 6 or 7 numeric columns");
    Sensors s;
    s.oriented = ncol >= 6;
    std::map<std::string, int64_t> index;
    for (size_t i = 0; i < rows.size(); ++i) {
        double v[7] = {0, 0, 0, 0, 0, 0, 1};
        for (size_t c = 0; c < ncol; ++c) v[c] = std::stod(rows[i][off + c]);
        s.radius.push_back(ncol == 4 ? v[3] : 0.0);
        if (ncol == 4) v[3] = 0.0;
        for (int c = 0; c < 3; ++c) {
            s.pos.push_back(v[c]);
            s.dir.push_back(v[3 + c]);
        }
        s.weight.push_back(ncol == 7 ? v[6] : 1.0);
        if (labeled) {
            auto it = index.find(rows[i][0]);
            if (it == index.end()) it = index.emplace(rows[i][0], (int64_t)index.size()).first;
            s.sensor.push_back(it->second);
        } else {
            s.sensor.push_back((int64_t)i);
        }
    }
    s.nsensors = labeled ? (int64_t)index.size() : (int64_t)rows.size();
    return s;
}

// Dipoles: a full n x 6 matrix in any of the formats (Matrix(filename)),
// returned row-major (x y z qx qy qz per dipole).
inline std::vector<double> read_dipoles(const std::string& path, int64_t& nd) {
    const Full m = read_full(path);
    if (m.ncol != 6) throw std::runtime_error("Dipoles File Format Error");
    nd = m.nlin;
    std::vector<double> d((size_t)(nd * 6));
    for (int64_t i = 0; i < nd; ++i)
        for (int c = 0; c < 6; ++c) d[i * 6 + c] = m(i, c);
    return d;
}

struct Geom {
    omh_geometry* g = nullptr;
    int64_t sizes[8] = {};
    Geom(const std::string& geom, const std::string& cond) {
        check(omh_geometry_load(geom.c_str(), cond.c_str(), &g));
        check(omh_geometry_sizes(g, sizes));
    }
    ~Geom() { omh_geometry_free(g); }
    int64_t n() const { return sizes[0]; }
    int64_t nv() const { return sizes[3]; }
    // The reference's old unknown numbering (Geometry::generate_indices with
    // OLD_ORDERING, geometry.cpp:256-295): mesh by mesh in geometry order,
    // its vertices then its triangles (none for a current barrier or an
    // isolated mesh); nested geometries only (geometry.cpp:270).  Returns
    // perm with perm[u] = the old index of reference-order unknown u < n().
    std::vector<int64_t> old_order() const {
        if (!sizes[7]) throw std::runtime_error("-old-ordering needs a nested geometry");
        const int64_t n = this->n();
        std::vector<int64_t> perm((size_t)n, -1);
        int64_t idx = 0;
        for (int32_t m = 0; m < (int32_t)sizes[5]; ++m) {
            int64_t info[8];
            check(omh_geometry_mesh_info(g, m, info));
            if (info[2] < 0) throw std::runtime_error("-old-ordering needs meshes that share no vertices");
            for (int64_t k = 0; k < info[4]; ++k) perm[(size_t)(info[2] + k)] = idx++;
            if (!(info[5] & 3))
                for (int64_t t = 0; t < info[1]; ++t)
                    if (info[0] + t < n) perm[(size_t)(info[0] + t)] = idx++;
        }
        for (int64_t u = 0; u < n; ++u)
            if (perm[(size_t)u] < 0 || perm[(size_t)u] >= n) throw std::runtime_error("old ordering: inconsistent numbering");
        return perm;
    }
};

// Rows (or columns) of a full matrix moved to the old numbering: row u goes
// to row perm[u].
inline Full permute_rows(const Full& a, const std::vector<int64_t>& perm) {
    Full b = a;
    for (int64_t j = 0; j < a.ncol; ++j)
        for (int64_t i = 0; i < a.nlin; ++i) b(perm[(size_t)i], j) = a(i, j);
    return b;
}
inline Full permute_cols(const Full& a, const std::vector<int64_t>& perm) {
    Full b = a;
    for (int64_t j = 0; j < a.ncol; ++j)
        for (int64_t i = 0; i < a.nlin; ++i) b(i, perm[(size_t)j]) = a(i, j);
    return b;
}

// Sums integration points into sensors: out (nsensors x ncols) = W P with W
// the weights matrix (Sensors::getWeightsMatrix, sensors.h:191-196).
inline Full sensor_sum(const Sensors& s, const std::vector<double>& P, int64_t ncols) {
    Full out;
    out.nlin = s.nsensors;
    out.ncol = ncols;
    out.a.assign((size_t)(s.nsensors * ncols), 0.0);
    const int64_t np = s.npts();
    for (int64_t j = 0; j < ncols; ++j)
        for (int64_t i = 0; i < np; ++i) out(s.sensor[i], j) += s.weight[i] * P[i + j * np];
    return out;
}


}  // namespace cli
