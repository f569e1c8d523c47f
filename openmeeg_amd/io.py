"""ctypes binding of libopenmeeg_amd_io.so (include/openmeeg_amd_io.h): the
reference's matrix file formats -- MAT-file 7.3 (.mat), raw binary (.bin)
and ASCII (.txt) -- for full matrices, vectors, packed symmetric matrices
and sparse matrices.  Host only; reading identifies the format from the
content, writing from the suffix (OpenMEEGMaths/src/MathsIO.C:71-155)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.environ.get("OMH_IO_LIBRARY", os.path.join(_HERE, "libopenmeeg_amd_io.so"))

FULL, SYMMETRIC, SPARSE = 0, 1, 4          # LinOpInfo::StorageType (linop.h:40)
FORMATS = {0: "mat", 1: "bin", 2: "txt"}


class LinOpInfo(C.Structure):
    _fields_ = [("storage", C.c_int32), ("dimension", C.c_int32), ("nlin", C.c_int64), ("ncol", C.c_int64),
                ("nnz", C.c_int64), ("format", C.c_int32)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError("%s is missing: build it with python -m openmeeg_amd.build" % IO_LIB_PATH)
        L = C.CDLL(IO_LIB_PATH)
        pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.omh_io_last_error.restype = C.c_char_p
        L.omh_io_info.argtypes = [C.c_char_p, C.POINTER(LinOpInfo)]
        L.omh_io_read.argtypes = [C.c_char_p, C.POINTER(LinOpInfo), pd, pi, pi]
        L.omh_io_write.argtypes = [C.c_char_p, C.POINTER(LinOpInfo), pd, pi, pi]
        _L = L
    return _L


def _check(rc):
    if rc == 0:
        return
    msg = lib().omh_io_last_error().decode()
    if rc == 2:
        raise ValueError(msg)
    if rc == 6:
        raise OSError(msg)
    raise RuntimeError(msg)


def info(path):
    li = LinOpInfo()
    _check(lib().omh_io_info(os.fsencode(path), C.byref(li)))
    return li


def read(path):
    """Returns (kind, value): ("full", F-ordered array; a vector is 1-D),
    ("symmetric", packed upper array), or ("sparse", (rows, cols, vals,
    (nlin, ncol))) with entries in (row, col) order."""
    li = info(path)
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int64)
    if li.storage == SPARSE:
        r = np.zeros(li.nnz, dtype=np.int64)
        c = np.zeros(li.nnz, dtype=np.int64)
        v = np.zeros(li.nnz)
        _check(lib().omh_io_read(os.fsencode(path), C.byref(li), v.ctypes.data_as(pd), r.ctypes.data_as(pi),
                                 c.ctypes.data_as(pi)))
        return "sparse", (r, c, v, (li.nlin, li.ncol))
    if li.storage == SYMMETRIC:
        p = np.zeros(li.nlin * (li.nlin + 1) // 2)
        _check(lib().omh_io_read(os.fsencode(path), C.byref(li), p.ctypes.data_as(pd), None, None))
        return "symmetric", p
    a = np.zeros(li.nlin * li.ncol)
    _check(lib().omh_io_read(os.fsencode(path), C.byref(li), a.ctypes.data_as(pd), None, None))
    if li.dimension == 1:
        return "full", a
    return "full", a.reshape(li.ncol, li.nlin).T.copy(order="F")


def _write(path, li, data, rows=None, cols=None):
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int64)
    _check(lib().omh_io_write(os.fsencode(path), C.byref(li), data.ctypes.data_as(pd),
                              None if rows is None else rows.ctypes.data_as(pi),
                              None if cols is None else cols.ctypes.data_as(pi)))


def write_full(path, a):
    """A matrix (2-D) or a vector (1-D)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        li = LinOpInfo(FULL, 1, a.size, 1, 0, 0)
        _write(path, li, np.ascontiguousarray(a))
    else:
        li = LinOpInfo(FULL, 2, a.shape[0], a.shape[1], 0, 0)
        _write(path, li, np.ascontiguousarray(a.ravel(order="F")))


def write_symmetric(path, packed):
    packed = np.ascontiguousarray(packed, dtype=np.float64)
    n = int((np.sqrt(8 * packed.size + 1) - 1) / 2)
    if n * (n + 1) // 2 != packed.size:
        raise ValueError("not a packed symmetric matrix")
    _write(path, LinOpInfo(SYMMETRIC, 2, n, n, 0, 0), packed)


def write_sparse(path, rows, cols, vals, shape):
    rows = np.ascontiguousarray(rows, dtype=np.int64).ravel()
    cols = np.ascontiguousarray(cols, dtype=np.int64).ravel()
    vals = np.ascontiguousarray(vals, dtype=np.float64).ravel()
    li = LinOpInfo(SPARSE, 2, int(shape[0]), int(shape[1]), len(vals), 0)
    _write(path, li, vals, rows, cols)
