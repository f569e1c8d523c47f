/*
 * openmeeg_amd_io: the on-disk matrix formats of the reference's maths
 * library, host only (no GPU), C ABI.
 *
 *   .mat  MAT-file version 7.3 (HDF5 with a 512-byte MATLAB user block):
 *         a full matrix is the double dataset "linop" (ncol x nlin, i.e. the
 *         column-major nlin x ncol data; a vector is a rank-1 dataset); a
 *         symmetric matrix is the struct "symmatrix" with fields "size"
 *         (uint32 1x1) and "data" (the packed upper triangle, 1 x n(n+1)/2);
 *         a sparse matrix is the sparse variable "matrix" (compressed
 *         columns: "data", "ir", "jc", attribute MATLAB_sparse = nlin).
 *         OpenMEEGMaths/include/MatlabIO.H:28-120,268-394.
 *   .bin  raw little-endian: u32 nlin [u32 ncol] then doubles (column-major;
 *         the packed upper triangle for a symmetric matrix; no ncol for a
 *         vector or a symmetric matrix), or u32 nlin, u32 ncol then
 *         (u32 i, u32 j, f64 v) records for a sparse matrix.
 *         OpenMEEGMaths/include/TrivialBinIO.H:34-195.
 *   .txt  ASCII rows (tab separated, default stream precision); a symmetric
 *         matrix writes row i from column i on; a sparse matrix writes
 *         "nlin ncol" then "i j v" lines.  OpenMEEGMaths/include/AsciiIO.H:
 *         34-265.
 *
 * Reading identifies the format from the content as the reference does
 * (MathsIO.C:82-108: MATLAB magic tag, else a leading number = ASCII, else
 * raw binary); writing picks it from the file suffix (MathsIO.C:71-81),
 * and a file without a known suffix (Head1.hm, Head1.dsm, ...) is written
 * in the first format that stores the matrix, MAT (matrix.cpp:219-226).
 * Every entry point returns 0 on success and a non-zero status otherwise
 * (1 runtime error, 2 invalid argument, 6 file error); omh_io_last_error()
 * gives the message (per thread).
 */
#ifndef OPENMEEG_AMD_IO_H
#define OPENMEEG_AMD_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LinOpInfo::StorageType values (OpenMEEGMaths/include/linop.h:40). */
#define OMH_IO_FULL 0
#define OMH_IO_SYMMETRIC 1
#define OMH_IO_SPARSE 4

#define OMH_IO_FORMAT_MAT 0
#define OMH_IO_FORMAT_BIN 1
#define OMH_IO_FORMAT_ASCII 2

/* LinOpInfo (linop.h): storage, dimension (1 = vector, 2 = matrix), sizes,
 * and for a sparse matrix its number of stored entries. */
typedef struct {
    int32_t storage;
    int32_t dimension;
    int64_t nlin;
    int64_t ncol;
    int64_t nnz;
    int32_t format; /* on read: the format found; ignored on write */
} omh_linop_info;

const char* omh_io_last_error(void);

/* maths::info(name), MathsIO.C:132-155. */
int omh_io_info(const char* path, omh_linop_info* info);

/* Reads the matrix of `path` (shape from omh_io_info):
 *   full:      data[nlin x ncol], column-major;
 *   symmetric: data[n(n+1)/2], packed upper (index i + j(j+1)/2, i <= j);
 *   sparse:    data[nnz] with rows[nnz], cols[nnz], ordered by (row, col).
 * `info` must be what omh_io_info returned for the file. */
int omh_io_read(const char* path, const omh_linop_info* info, double* data, int64_t* rows, int64_t* cols);

/* Writes a matrix in the format of the suffix of `path` (.bin, .txt; .mat
 * and any other suffix: MAT 7.3).  Layouts as omh_io_read; sparse entries
 * in any order (duplicates are an error). */
int omh_io_write(const char* path, const omh_linop_info* info, const double* data, const int64_t* rows,
                 const int64_t* cols);

#ifdef __cplusplus
}
#endif

#endif
