// On-disk matrix formats of the reference's maths library (host only):
// MAT-file 7.3 through the HDF5 C library, the raw binary format and the
// ASCII format.  C ABI in include/openmeeg_amd_io.h.
//
// Reference behaviour (file:line in OpenMEEG 2.5.16):
//   format identification     OpenMEEGMaths/src/MathsIO.C:71-155
//   MAT 7.3 variables         OpenMEEGMaths/include/MatlabIO.H:28-120, 268-394
//   raw binary                OpenMEEGMaths/include/TrivialBinIO.H:34-195
//   ASCII                     OpenMEEGMaths/include/AsciiIO.H:34-265
// The MAT files store their variables as the reference's matio writes them:
// dense datasets (those with a MATLAB class) in DEFLATE-9 chunks whose sides
// are powers of two of at most 4096 elements in all (the chunk shapes of the
// reference's own files), sparse components contiguous.  Byte identity is
// not possible: the header records the creating library and time.
#include <hdf5.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/openmeeg_amd_io.h"

namespace {

thread_local std::string g_err = "no error";

struct FileError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = std::string("invalid argument: ") + e.what();
        return 2;
    } catch (const FileError& e) {
        g_err = e.what();
        return 6;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

std::string suffix_of(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    return path.substr(dot + 1);
}

// ---------------------------------------------------------------------------
// format identification (MathsIO.C:82-108, ReadTag + identify of each IO in
// priority order: MATLAB 5, ASCII 9, raw binary 20)
// ---------------------------------------------------------------------------
const char kMagic[] = "MATLAB";

int identify(const std::string& path) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw FileError("cannot open " + path + " for reading");
    char buf[64] = {};
    is.read(buf, sizeof(buf) - 1);
    const std::string tag(buf, (size_t)is.gcount());
    if (tag.compare(0, 6, kMagic) == 0) return OMH_IO_FORMAT_MAT;
    std::stringstream ss(tag);
    double d;
    if (ss >> d) return OMH_IO_FORMAT_ASCII;
    return OMH_IO_FORMAT_BIN;
}

// ---------------------------------------------------------------------------
// raw binary
// ---------------------------------------------------------------------------
omh_linop_info bin_info(const std::string& path) {
    std::ifstream is(path, std::ios::binary | std::ios::ate);
    if (!is) throw FileError("cannot open " + path);
    const int64_t size = (int64_t)is.tellg() - 4;
    is.seekg(0);
    uint32_t ui = 0;
    is.read(reinterpret_cast<char*>(&ui), 4);
    if (!is) throw FileError(path + ": truncated binary matrix");
    omh_linop_info li{};
    li.format = OMH_IO_FORMAT_BIN;
    li.nlin = ui;
    const int64_t n = ui;
    if (size == n * 8) {
        li.storage = OMH_IO_FULL;
        li.dimension = 1;
        li.ncol = 1;
    } else if (size == n * (n + 1) / 2 * 8) {
        li.storage = OMH_IO_SYMMETRIC;
        li.dimension = 2;
        li.ncol = n;
    } else {
        is.read(reinterpret_cast<char*>(&ui), 4);
        if (!is) throw FileError(path + ": truncated binary matrix");
        li.dimension = 2;
        li.ncol = ui;
        if (size - 4 == li.nlin * li.ncol * 8) {
            li.storage = OMH_IO_FULL;
        } else {
            li.storage = OMH_IO_SPARSE;
            li.nnz = (size - 4) / 16;
        }
    }
    return li;
}

void bin_read(const std::string& path, const omh_linop_info& li, double* data, int64_t* rows, int64_t* cols) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw FileError("cannot open " + path);
    const bool two = li.dimension == 2 && li.storage != OMH_IO_SYMMETRIC;
    is.seekg(two ? 8 : 4);
    if (li.storage == OMH_IO_SPARSE) {
        std::vector<std::pair<std::pair<int64_t, int64_t>, double>> e;
        for (int64_t k = 0; k < li.nnz; ++k) {
            uint32_t ij[2];
            double v;
            is.read(reinterpret_cast<char*>(ij), 8);
            is.read(reinterpret_cast<char*>(&v), 8);
            if (!is) throw FileError(path + ": truncated sparse matrix");
            e.push_back({{ij[0], ij[1]}, v});
        }
        std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < e.size(); ++k) {
            rows[k] = e[k].first.first;
            cols[k] = e[k].first.second;
            data[k] = e[k].second;
        }
        return;
    }
    const int64_t count = li.storage == OMH_IO_SYMMETRIC ? li.nlin * (li.nlin + 1) / 2 : li.nlin * li.ncol;
    is.read(reinterpret_cast<char*>(data), 8 * count);
    if (!is) throw FileError(path + ": truncated binary matrix");
}

void bin_write(const std::string& path, const omh_linop_info& li, const double* data, const int64_t* rows,
               const int64_t* cols) {
    std::ofstream os(path, std::ios::binary);
    if (!os) throw FileError("cannot open " + path + " for writing");
    if (li.nlin > UINT32_MAX || li.ncol > UINT32_MAX) throw std::invalid_argument("dimension above 2^32 - 1");
    const uint32_t nlin = (uint32_t)li.nlin, ncol = (uint32_t)li.ncol;
    os.write(reinterpret_cast<const char*>(&nlin), 4);
    if (li.storage != OMH_IO_SYMMETRIC && li.dimension != 1) os.write(reinterpret_cast<const char*>(&ncol), 4);
    if (li.storage == OMH_IO_SPARSE) {
        for (int64_t k = 0; k < li.nnz; ++k) {
            const uint32_t ij[2] = {(uint32_t)rows[k], (uint32_t)cols[k]};
            os.write(reinterpret_cast<const char*>(ij), 8);
            os.write(reinterpret_cast<const char*>(&data[k]), 8);
        }
    } else {
        const int64_t count = li.storage == OMH_IO_SYMMETRIC ? li.nlin * (li.nlin + 1) / 2 : li.nlin * li.ncol;
        os.write(reinterpret_cast<const char*>(data), 8 * count);
    }
    if (!os) throw FileError("write error on " + path);
}

// ---------------------------------------------------------------------------
// ASCII
// ---------------------------------------------------------------------------
unsigned numbers_in(const std::string& line) {
    std::stringstream ss(line);
    unsigned k = 0;
    double d;
    while (ss >> d) ++k;
    return k;
}

omh_linop_info ascii_info(const std::string& path) {
    std::ifstream is(path);
    if (!is) throw FileError("cannot open " + path);
    std::string line;
    std::getline(is, line);
    omh_linop_info li{};
    li.format = OMH_IO_FORMAT_ASCII;
    const unsigned len = numbers_in(line);
    li.ncol = len;
    std::string line2;
    if (!std::getline(is, line2)) {   // AsciiIO::set_type: one line = 1 x len
        li.storage = OMH_IO_FULL;
        li.dimension = 2;
        li.nlin = 1;
        return li;
    }
    const unsigned len1 = numbers_in(line2);
    if (len1 == len) li.storage = OMH_IO_FULL;
    else if (len1 + 1 == len) li.storage = OMH_IO_SYMMETRIC;
    else if (len == 2 && len1 == 3) li.storage = OMH_IO_SPARSE;
    else throw FileError(path + ": rows of an ASCII matrix with inconsistent lengths");
    li.dimension = (li.storage == OMH_IO_FULL && len == 1) ? 1 : 2;
    if (li.storage == OMH_IO_SPARSE) {
        std::stringstream ss(line);
        ss >> li.nlin >> li.ncol;
        int64_t k = 0;
        while (std::getline(is, line2))
            if (numbers_in(line2) == 3) ++k;
        li.nnz = k + 1;
        return li;
    }
    std::ifstream is2(path);
    int64_t n = 0;
    while (std::getline(is2, line))
        if (!line.empty()) ++n;
    li.nlin = n;
    if (li.storage == OMH_IO_SYMMETRIC && li.nlin != li.ncol)
        throw FileError(path + ": symmetric ASCII matrix that is not square");
    return li;
}

void ascii_read(const std::string& path, const omh_linop_info& li, double* data, int64_t* rows, int64_t* cols) {
    std::ifstream is(path);
    if (!is) throw FileError("cannot open " + path);
    std::string line;
    if (li.storage == OMH_IO_SPARSE) {
        std::getline(is, line);
        std::vector<std::pair<std::pair<int64_t, int64_t>, double>> e;
        while (std::getline(is, line)) {
            std::stringstream ss(line);
            int64_t i, j;
            double v;
            if (ss >> i >> j >> v) e.push_back({{i, j}, v});
        }
        std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        if ((int64_t)e.size() != li.nnz) throw FileError(path + ": sparse ASCII matrix changed while read");
        for (size_t k = 0; k < e.size(); ++k) {
            rows[k] = e[k].first.first;
            cols[k] = e[k].first.second;
            data[k] = e[k].second;
        }
        return;
    }
    for (int64_t i = 0; i < li.nlin; ++i) {
        if (!std::getline(is, line)) throw FileError(path + ": truncated ASCII matrix");
        std::stringstream ss(line);
        if (li.storage == OMH_IO_SYMMETRIC) {
            for (int64_t j = i; j < li.ncol; ++j) ss >> data[i + j * (j + 1) / 2];
        } else {
            for (int64_t j = 0; j < li.ncol; ++j) ss >> data[i + j * li.nlin];
        }
        if (ss.fail()) throw FileError(path + ": bad number in an ASCII matrix");
    }
}

void ascii_write(const std::string& path, const omh_linop_info& li, const double* data, const int64_t* rows,
                 const int64_t* cols) {
    std::ofstream os(path);
    if (!os) throw FileError("cannot open " + path + " for writing");
    if (li.storage == OMH_IO_SPARSE) {
        os << li.nlin << " " << li.ncol << std::endl;
        for (int64_t k = 0; k < li.nnz; ++k) os << rows[k] << " " << cols[k] << " " << data[k] << std::endl;
    } else if (li.storage == OMH_IO_SYMMETRIC) {
        for (int64_t i = 0; i < li.nlin; ++i)
            for (int64_t j = i; j < li.ncol; ++j)
                os << data[i + j * (j + 1) / 2] << ((j != li.ncol - 1) ? "\t" : "\n");
    } else if (li.dimension == 1) {
        for (int64_t i = 0; i < li.nlin; ++i) os << data[i] << std::endl;
    } else {
        for (int64_t i = 0; i < li.nlin; ++i)
            for (int64_t j = 0; j < li.ncol; ++j)
                os << data[i + j * li.nlin] << ((j != li.ncol - 1) ? "\t" : "\n");
    }
    if (!os) throw FileError("write error on " + path);
}

// ---------------------------------------------------------------------------
// MAT-file 7.3
// ---------------------------------------------------------------------------
struct H5Handle {   // closes an HDF5 identifier with its own close routine
    hid_t id;
    herr_t (*close)(hid_t);
    H5Handle(hid_t i, herr_t (*c)(hid_t), const char* what) : id(i), close(c) {
        if (id < 0) throw FileError(std::string("HDF5: ") + what);
    }
    ~H5Handle() {
        if (id >= 0) close(id);
    }
    H5Handle(const H5Handle&) = delete;
    H5Handle& operator=(const H5Handle&) = delete;
    operator hid_t() const { return id; }
};

void h5check(herr_t rc, const char* what) {
    if (rc < 0) throw FileError(std::string("HDF5: ") + what);
}

std::string read_class(hid_t obj) {
    if (H5Aexists(obj, "MATLAB_class") <= 0) return "";
    H5Handle a(H5Aopen(obj, "MATLAB_class", H5P_DEFAULT), H5Aclose, "open MATLAB_class");
    H5Handle t(H5Aget_type(a), H5Tclose, "attribute type");
    const size_t sz = H5Tget_size(t);
    std::vector<char> buf(sz + 1, 0);
    H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
    h5check(H5Tset_size(mt, sz), "string size");
    h5check(H5Aread(a, mt, buf.data()), "read MATLAB_class");
    return std::string(buf.data());
}

void write_class(hid_t obj, const char* cls) {
    H5Handle t(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
    h5check(H5Tset_size(t, std::strlen(cls)), "string size");
    h5check(H5Tset_strpad(t, H5T_STR_NULLTERM), "string pad");
    H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
    H5Handle a(H5Acreate2(obj, "MATLAB_class", t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "create MATLAB_class");
    h5check(H5Awrite(a, t, cls), "write MATLAB_class");
}

std::vector<hsize_t> dims_of(hid_t dset) {
    H5Handle sp(H5Dget_space(dset), H5Sclose, "dataspace");
    const int rank = H5Sget_simple_extent_ndims(sp);
    if (rank < 0) throw FileError("HDF5: rank");
    std::vector<hsize_t> d(rank);
    H5Sget_simple_extent_dims(sp, d.data(), nullptr);
    return d;
}

// The first variable at the root of the file (matio reads the first one).
std::string first_variable(hid_t f) {
    H5G_info_t gi;
    h5check(H5Gget_info(f, &gi), "root group");
    for (hsize_t k = 0; k < gi.nlinks; ++k) {
        char name[256] = {};
        if (H5Lget_name_by_idx(f, ".", H5_INDEX_NAME, H5_ITER_INC, k, name, sizeof(name), H5P_DEFAULT) < 0) continue;
        if (name[0] != '#') return name;   // "#refs#" and the like are not variables
    }
    throw FileError("MAT file without a variable");
}

omh_linop_info mat_info(const std::string& path) {
    H5Handle f(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, ("open " + path).c_str());
    const std::string var = first_variable(f);
    H5O_info_t oi;
    h5check(H5Oget_info_by_name(f, var.c_str(), &oi, H5P_DEFAULT), "object info");
    omh_linop_info li{};
    li.format = OMH_IO_FORMAT_MAT;
    if (oi.type == H5O_TYPE_DATASET) {   // full matrix or vector
        H5Handle d(H5Dopen2(f, var.c_str(), H5P_DEFAULT), H5Dclose, "open dataset");
        if (read_class(d) != "double") throw FileError(path + ": variable is not of class double");
        const auto dims = dims_of(d);
        li.storage = OMH_IO_FULL;
        if (dims.size() == 1) {
            li.dimension = 1;
            li.nlin = (int64_t)dims[0];
            li.ncol = 1;
        } else if (dims.size() == 2) {
            li.dimension = 2;
            li.nlin = (int64_t)dims[1];
            li.ncol = (int64_t)dims[0];
        } else {
            throw FileError(path + ": variable of rank above 2");
        }
        return li;
    }
    H5Handle g(H5Gopen2(f, var.c_str(), H5P_DEFAULT), H5Gclose, "open group");
    const std::string cls = read_class(g);
    if (cls == "struct") {   // symmetric matrix: fields size and data
        H5Handle s(H5Dopen2(g, "size", H5P_DEFAULT), H5Dclose, "symmatrix.size");
        uint64_t n = 0;
        h5check(H5Dread(s, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &n), "read symmatrix.size");
        H5Handle d(H5Dopen2(g, "data", H5P_DEFAULT), H5Dclose, "symmatrix.data");
        hsize_t count = 1;
        for (hsize_t x : dims_of(d)) count *= x;
        if (count != n * (n + 1) / 2) throw FileError(path + ": symmatrix data size does not match its size field");
        li.storage = OMH_IO_SYMMETRIC;
        li.dimension = 2;
        li.nlin = li.ncol = (int64_t)n;
        return li;
    }
    if (H5Aexists(g, "MATLAB_sparse") > 0) {
        H5Handle a(H5Aopen(g, "MATLAB_sparse", H5P_DEFAULT), H5Aclose, "MATLAB_sparse");
        uint64_t nlin = 0;
        h5check(H5Aread(a, H5T_NATIVE_UINT64, &nlin), "read MATLAB_sparse");
        H5Handle jc(H5Dopen2(g, "jc", H5P_DEFAULT), H5Dclose, "sparse jc");
        const auto dj = dims_of(jc);
        li.storage = OMH_IO_SPARSE;
        li.dimension = 2;
        li.nlin = (int64_t)nlin;
        if (dj.empty() || dj[0] < 1) throw FileError(path + ": sparse matrix without column starts");
        li.ncol = (int64_t)dj[0] - 1;
        // stored entries = jc[ncol] (MATLAB may pad ir/data to nzmax)
        std::vector<uint64_t> jcv(dj[0]);
        h5check(H5Dread(jc, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, jcv.data()), "read jc");
        li.nnz = (int64_t)jcv.back();
        if (li.nnz < 0) throw FileError(path + ": bad sparse column starts");
        return li;
    }
    throw FileError(path + ": variable is not a matrix, a symmatrix struct or a sparse matrix");
}

void mat_read(const std::string& path, const omh_linop_info& li, double* data, int64_t* rows, int64_t* cols) {
    H5Handle f(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, ("open " + path).c_str());
    const std::string var = first_variable(f);
    if (li.storage == OMH_IO_FULL) {
        H5Handle d(H5Dopen2(f, var.c_str(), H5P_DEFAULT), H5Dclose, "open dataset");
        h5check(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "read matrix");
        return;
    }
    H5Handle g(H5Gopen2(f, var.c_str(), H5P_DEFAULT), H5Gclose, "open group");
    if (li.storage == OMH_IO_SYMMETRIC) {
        H5Handle d(H5Dopen2(g, "data", H5P_DEFAULT), H5Dclose, "symmatrix.data");
        h5check(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "read symmatrix.data");
        return;
    }
    // compressed columns -> (row, col) ordered triplets; every dataset is
    // read into a buffer of its own size, then checked against jc
    std::vector<uint64_t> jc, ir;
    std::vector<double> v;
    {
        H5Handle d(H5Dopen2(g, "jc", H5P_DEFAULT), H5Dclose, "sparse jc");
        const auto dims = dims_of(d);
        if (dims.empty() || (int64_t)dims[0] != li.ncol + 1) throw FileError(path + ": sparse jc size changed");
        jc.resize(dims[0]);
        h5check(H5Dread(d, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, jc.data()), "read jc");
    }
    const uint64_t nnz = (uint64_t)li.nnz;
    if (jc[0] != 0 || jc[li.ncol] != nnz) throw FileError(path + ": sparse jc does not match the entry count");
    for (int64_t j = 0; j < li.ncol; ++j)
        if (jc[j] > jc[j + 1]) throw FileError(path + ": sparse jc is not non-decreasing");
    if (nnz > 0) {
        H5Handle d(H5Dopen2(g, "data", H5P_DEFAULT), H5Dclose, "sparse data");
        const auto dd = dims_of(d);
        if (dd.empty() || dd[0] < nnz) throw FileError(path + ": sparse data shorter than jc[ncol]");
        v.resize(dd[0]);
        h5check(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read data");
        H5Handle r(H5Dopen2(g, "ir", H5P_DEFAULT), H5Dclose, "sparse ir");
        const auto dr = dims_of(r);
        if (dr.empty() || dr[0] < nnz) throw FileError(path + ": sparse ir shorter than jc[ncol]");
        ir.resize(dr[0]);
        h5check(H5Dread(r, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, ir.data()), "read ir");
        for (uint64_t k = 0; k < nnz; ++k)
            if (ir[k] >= (uint64_t)li.nlin) throw FileError(path + ": sparse row index out of range");
    }
    std::vector<std::pair<std::pair<int64_t, int64_t>, double>> e;
    for (int64_t j = 0; j < li.ncol; ++j)
        for (uint64_t k = jc[j]; k < jc[j + 1]; ++k) e.push_back({{(int64_t)ir[k], j}, v[k]});
    std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t k = 0; k < e.size(); ++k) {
        rows[k] = e[k].first.first;
        cols[k] = e[k].first.second;
        data[k] = e[k].second;
    }
}

// Chunk of a compressed dataset: per dimension the largest power of two not
// above it, then the largest side halved while the chunk holds more than 4096
// elements.
std::vector<hsize_t> chunk_of(const std::vector<hsize_t>& dims) {
    std::vector<hsize_t> c(dims.size());
    hsize_t total = 1;
    for (size_t i = 0; i < dims.size(); ++i) {
        hsize_t p = 1;
        while (2 * p <= dims[i]) p *= 2;
        c[i] = p;
        total *= p;
    }
    while (total > 4096) {
        size_t big = c.size() - 1;
        for (size_t i = c.size(); i-- > 0;)
            if (c[i] > c[big]) big = i;
        c[big] /= 2;
        total /= 2;
    }
    return c;
}

// Dense datasets (with a MATLAB class) are written in DEFLATE-9 chunks, the
// components of sparse matrices (no class) contiguous.
void write_dataset(hid_t loc, const char* name, hid_t file_type, hid_t mem_type, const std::vector<hsize_t>& dims,
                   const void* buf, const char* cls) {
    H5Handle sp(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose, "dataspace");
    hsize_t count = 1;
    for (hsize_t x : dims) count *= x;
    H5Handle dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "dataset creation properties");
    if (cls && count > 0) {
        const std::vector<hsize_t> chunk = chunk_of(dims);
        h5check(H5Pset_chunk(dcpl, (int)chunk.size(), chunk.data()), "chunk");
        h5check(H5Pset_deflate(dcpl, 9), "deflate");
    }
    H5Handle d(H5Dcreate2(loc, name, file_type, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose, name);
    if (count > 0) h5check(H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), name);
    if (cls) write_class(d, cls);
}

// The 128-byte MAT-file header in the 512-byte user block: 116 bytes of
// text, 8 bytes of subsystem offset, version 0x0200, endian tag "IM".
void write_mat_header(const std::string& path) {
    char hdr[128];
    std::memset(hdr, ' ', sizeof(hdr));
    char text[117];
    const std::time_t now = std::time(nullptr);
    char date[64] = {};
    std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y", std::localtime(&now));
    std::snprintf(text, sizeof(text), "MATLAB 7.3 MAT-file, Platform: x86_64-linux-gnu, Created by: openmeeg_amd on %s",
                  date);
    std::memcpy(hdr, text, std::strlen(text));
    hdr[124] = 0x00;
    hdr[125] = 0x02;
    hdr[126] = 'I';
    hdr[127] = 'M';
    std::FILE* fp = std::fopen(path.c_str(), "r+b");
    if (!fp) throw FileError("cannot reopen " + path);
    const size_t w = std::fwrite(hdr, 1, sizeof(hdr), fp);
    std::fclose(fp);
    if (w != sizeof(hdr)) throw FileError("cannot write the MAT header of " + path);
}

void mat_write(const std::string& path, const omh_linop_info& li, const double* data, const int64_t* rows,
               const int64_t* cols) {
    {
        H5Handle fcpl(H5Pcreate(H5P_FILE_CREATE), H5Pclose, "file creation properties");
        h5check(H5Pset_userblock(fcpl, 512), "user block");
        H5Handle f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), H5Fclose, ("create " + path).c_str());
        if (li.storage == OMH_IO_FULL) {
            std::vector<hsize_t> dims;
            if (li.dimension == 1) dims = {(hsize_t)li.nlin};
            else dims = {(hsize_t)li.ncol, (hsize_t)li.nlin};
            write_dataset(f, "linop", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, dims, data, "double");
        } else if (li.storage == OMH_IO_SYMMETRIC) {
            H5Handle g(H5Gcreate2(f, "symmatrix", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "symmatrix");
            write_class(g, "struct");
            // MATLAB_fields: variable-length sequences of 1-character strings
            H5Handle ct(H5Tcopy(H5T_C_S1), H5Tclose, "char type");
            h5check(H5Tset_size(ct, 1), "char size");
            H5Handle vt(H5Tvlen_create(ct), H5Tclose, "vlen type");
            const char* names[2] = {"size", "data"};
            hvl_t fields[2];
            for (int k = 0; k < 2; ++k) {
                fields[k].len = std::strlen(names[k]);
                fields[k].p = const_cast<char*>(names[k]);
            }
            const hsize_t two = 2;
            H5Handle sp(H5Screate_simple(1, &two, nullptr), H5Sclose, "fields space");
            H5Handle a(H5Acreate2(g, "MATLAB_fields", vt, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "MATLAB_fields");
            h5check(H5Awrite(a, vt, fields), "write MATLAB_fields");
            const uint32_t n32 = (uint32_t)li.nlin;
            write_dataset(g, "size", H5T_STD_U32LE, H5T_NATIVE_UINT32, {1, 1}, &n32, "uint32");
            write_dataset(g, "data", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, {1, (hsize_t)(li.nlin * (li.nlin + 1) / 2)},
                          data, "double");
        } else {   // sparse: compressed columns, rows ascending in each column
            std::vector<std::pair<std::pair<int64_t, int64_t>, double>> e(li.nnz);
            for (int64_t k = 0; k < li.nnz; ++k) e[k] = {{cols[k], rows[k]}, data[k]};
            std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            std::vector<uint64_t> ir(li.nnz), jc(li.ncol + 1, 0);
            std::vector<double> v(li.nnz);
            for (int64_t k = 0; k < li.nnz; ++k) {
                ir[k] = (uint64_t)e[k].first.second;
                v[k] = e[k].second;
                jc[e[k].first.first + 1]++;
            }
            for (int64_t j = 0; j < li.ncol; ++j) jc[j + 1] += jc[j];
            H5Handle g(H5Gcreate2(f, "matrix", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "matrix");
            write_class(g, "double");
            const uint64_t nlin = (uint64_t)li.nlin;
            H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
            H5Handle a(H5Acreate2(g, "MATLAB_sparse", H5T_STD_U64LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
                       "MATLAB_sparse");
            h5check(H5Awrite(a, H5T_NATIVE_UINT64, &nlin), "write MATLAB_sparse");
            write_dataset(g, "data", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, {(hsize_t)li.nnz}, v.data(), nullptr);
            write_dataset(g, "ir", H5T_STD_U64LE, H5T_NATIVE_UINT64, {(hsize_t)li.nnz}, ir.data(), nullptr);
            write_dataset(g, "jc", H5T_STD_U64LE, H5T_NATIVE_UINT64, {(hsize_t)(li.ncol + 1)}, jc.data(), nullptr);
        }
    }
    write_mat_header(path);
}

void check_info(const omh_linop_info* li) {
    if (!li) throw std::invalid_argument("info is NULL");
    if (li->storage != OMH_IO_FULL && li->storage != OMH_IO_SYMMETRIC && li->storage != OMH_IO_SPARSE)
        throw std::invalid_argument("unknown storage type");
    if (li->dimension != 1 && li->dimension != 2) throw std::invalid_argument("dimension must be 1 or 2");
    if (li->nlin < 0 || li->ncol < 0 || li->nnz < 0) throw std::invalid_argument("negative size");
    if (li->storage == OMH_IO_SYMMETRIC && li->nlin != li->ncol) throw std::invalid_argument("symmetric matrix not square");
    if (li->dimension == 1 && (li->storage != OMH_IO_FULL || li->ncol != 1))
        throw std::invalid_argument("a vector is full with one column");
}

}  // namespace

extern "C" {

const char* omh_io_last_error(void) { return g_err.c_str(); }

int omh_io_info(const char* path, omh_linop_info* info) {
    return guard([&] {
        if (!path || !info) throw std::invalid_argument("NULL argument");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        switch (identify(path)) {
            case OMH_IO_FORMAT_MAT: *info = mat_info(path); break;
            case OMH_IO_FORMAT_ASCII: *info = ascii_info(path); break;
            default: *info = bin_info(path); break;
        }
    });
}

int omh_io_read(const char* path, const omh_linop_info* info, double* data, int64_t* rows, int64_t* cols) {
    return guard([&] {
        if (!path) throw std::invalid_argument("path is NULL");
        check_info(info);
        const int64_t count = info->storage == OMH_IO_SPARSE     ? info->nnz
                              : info->storage == OMH_IO_SYMMETRIC ? info->nlin * (info->nlin + 1) / 2
                                                                  : info->nlin * info->ncol;
        if (count > 0 && !data) throw std::invalid_argument("data is NULL");
        if (info->storage == OMH_IO_SPARSE && info->nnz > 0 && (!rows || !cols))
            throw std::invalid_argument("rows / cols are NULL");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        switch (info->format) {
            case OMH_IO_FORMAT_MAT: mat_read(path, *info, data, rows, cols); break;
            case OMH_IO_FORMAT_ASCII: ascii_read(path, *info, data, rows, cols); break;
            case OMH_IO_FORMAT_BIN: bin_read(path, *info, data, rows, cols); break;
            default: throw std::invalid_argument("unknown format");
        }
    });
}

int omh_io_write(const char* path, const omh_linop_info* info, const double* data, const int64_t* rows,
                 const int64_t* cols) {
    return guard([&] {
        if (!path) throw std::invalid_argument("path is NULL");
        check_info(info);
        // sparse entries in (row, col) order, the iteration order of the
        // reference's SparseMatrix (a map keyed by (i, j))
        std::vector<int64_t> sr, sc;
        std::vector<double> sv;
        if (info->storage == OMH_IO_SPARSE && info->nnz > 0) {
            if (!rows || !cols || !data) throw std::invalid_argument("sparse entries are NULL");
            std::vector<std::pair<std::pair<int64_t, int64_t>, double>> e(info->nnz);
            for (int64_t k = 0; k < info->nnz; ++k) {
                if (rows[k] < 0 || rows[k] >= info->nlin || cols[k] < 0 || cols[k] >= info->ncol)
                    throw std::invalid_argument("sparse entry out of range");
                e[k] = {{rows[k], cols[k]}, data[k]};
            }
            std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t k = 1; k < e.size(); ++k)
                if (e[k].first == e[k - 1].first) throw std::invalid_argument("duplicate sparse entry");
            for (const auto& x : e) {
                sr.push_back(x.first.first);
                sc.push_back(x.first.second);
                sv.push_back(x.second);
            }
            rows = sr.data();
            cols = sc.data();
            data = sv.data();
        } else if (info->storage != OMH_IO_SPARSE && info->nlin * info->ncol > 0 && !data) {
            throw std::invalid_argument("data is NULL");
        }
        const std::string sfx = suffix_of(path);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        // Matrix::save / SymMatrix::save (matrix.cpp:219-226, symmatrix.cpp:
        // 208-216): the format of the suffix, else (no or unknown suffix)
        // the first format in priority order that stores the matrix, MAT
        if (sfx == "bin") bin_write(path, *info, data, rows, cols);
        else if (sfx == "txt") ascii_write(path, *info, data, rows, cols);
        else mat_write(path, *info, data, rows, cols);
    });
}

}  // extern "C"
