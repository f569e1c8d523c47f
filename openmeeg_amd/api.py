"""Python surface of openmeeg_amd, mirroring the reference's ``openmeeg``
module for the symmetric-BEM EEG forward path
(wrapping/python/openmeeg/__init__.py:15-39, _openmeeg.i, _make_geometry.py).

Host code only: every computation runs in the HIP library through the C ABI
(_lib.py); PyTorch provides device memory and the current stream.  There is
no CPU fallback -- without the library or a GPU the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib, ORDER_REFERENCE, ORDER_TRIANGLES_FIRST

_torch = None


def _t():
    global _torch
    if _torch is None:
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("openmeeg_amd needs a HIP GPU (torch.cuda.is_available() is False)")
        _torch = torch
    return _torch


def _stream():
    return C.c_void_p(_t().cuda.current_stream().cuda_stream)


_SIDE_STREAMS = {}


def _side_stream():
    """A second torch stream of the current device (default priority) for
    device work that runs beside the caller's stream."""
    torch = _t()
    dev = torch.cuda.current_device()
    if dev not in _SIDE_STREAMS:
        _SIDE_STREAMS[dev] = torch.cuda.Stream(device=dev)
    return _SIDE_STREAMS[dev]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev_f64(n):
    return _t().empty(int(n), dtype=_t().float64, device="cuda")


def _to_dev(a, dtype=None):
    torch = _t()
    a = np.ascontiguousarray(a, dtype=np.float64 if dtype is None else dtype)
    return torch.from_numpy(a).to("cuda")


# ---------------------------------------------------------------------------
# Integrator (integrator.h:19-37)
# ---------------------------------------------------------------------------
class Integrator:
    """Integrator(order), Integrator(order, tol) (10 levels) or
    Integrator(order, levels, tol=0.0001); order is clamped to 1..3 as
    safe_order (integrator.h:24-29)."""

    def __init__(self, order, levels=None, tol=None):
        if levels is not None and tol is None and isinstance(levels, float):
            levels, tol = 10, levels
        if levels is None:
            levels, tol = 0, 0.0
        if tol is None:
            tol = 0.0001
        order = int(order)
        self.order = order if 0 < order < 4 else (1 if order < 1 else 3)
        self.levels = int(levels)
        self.tolerance = float(tol)

    def __repr__(self):
        return "Integrator(%d, %d, %g)" % (self.order, self.levels, self.tolerance)


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------
class Geometry:
    """Geometry(geom_file, cond_file) (geometry.h:67-69), or the result of
    make_nested_geometry / make_geometry."""

    def __init__(self, geom_file=None, cond_file=None):
        self._h = C.c_void_p()
        if geom_file is not None:
            if cond_file is None:
                raise ValueError("a conductivity file is needed")
            check(lib().omh_geometry_load(os.fsencode(geom_file), os.fsencode(cond_file), C.byref(self._h)))
            self._load_sizes()

    def _load_sizes(self):
        s = (C.c_int64 * 8)()
        check(lib().omh_geometry_sizes(self._h, s))
        (self._size, self._nbp, self._nbct, self._nv, self._nt, self._nm, self._nd, nested) = list(s)
        self._nested = bool(nested)

    def __del__(self):
        h = getattr(self, "_h", None)
        L = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and h.value and L is not None:
            L.omh_geometry_free(h)
            self._h = None

    # reference accessors (geometry.h:91,160,253)
    def nb_parameters(self):
        return self._nbp

    def nb_current_barrier_triangles(self):
        return self._nbct

    def is_nested(self):
        return self._nested

    @property
    def size(self):
        """HeadMat dimension nb_parameters - nb_current_barrier_triangles."""
        return self._size

    @property
    def n_vertex_unknowns(self):
        return self._nv

    def domain_index(self, name):
        idx = C.c_int32()
        check(lib().omh_geometry_domain_index(self._h, name.encode(), C.byref(idx)))
        return idx.value

    def domain_names(self):
        out = []
        buf = C.create_string_buffer(256)
        for i in range(self._nd):
            check(lib().omh_geometry_domain_name(self._h, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def conductivities(self):
        """Conductivity of every domain, in domain order (domain.h:77-79)."""
        v = C.c_double()
        out = []
        for d in range(self._nd):
            check(lib().omh_geometry_domain_conductivity(self._h, d, C.byref(v)))
            out.append(v.value)
        return np.array(out)

    def set_conductivity(self, domain, value):
        """Change the conductivity of a conductive domain (index or name);
        the HeadMat / source coefficients follow."""
        d = self.domain_index(domain) if isinstance(domain, str) else int(domain)
        check(lib().omh_geometry_set_conductivity(self._h, d, float(value)))

    def mesh_pairs(self):
        """Communicating mesh pairs (geometry.cpp:405-420) as dicts with the
        meshes, orientation, (S, N, D) coefficients and common domains."""
        cnt = C.c_int32()
        check(lib().omh_geometry_pair_count(self._h, C.byref(cnt)))
        out = []
        for i in range(cnt.value):
            m = (C.c_int32 * 2)()
            o, mask = C.c_int32(), C.c_int32()
            c = (C.c_double * 3)()
            check(lib().omh_geometry_pair_info(self._h, i, m, C.byref(o), c, C.byref(mask)))
            out.append(dict(meshes=(m[0], m[1]), orientation=o.value, coeffs=tuple(c),
                            domains=[d for d in range(self._nd) if mask.value >> d & 1]))
        return out

    def mesh_info(self, m):
        """Unknown ranges and flags of mesh m (reference numbering)."""
        a = (C.c_int64 * 8)()
        check(lib().omh_geometry_mesh_info(self._h, m, a))
        return dict(tri_first=a[0], ntris=a[1], vert_first=a[2], vert_end=a[3], nverts=a[4],
                    current_barrier=bool(a[5] & 1), isolated=bool(a[5] & 2), outermost=bool(a[5] & 4),
                    part=a[6], deflation_pivot=a[7])

    @property
    def n_meshes(self):
        return self._nm

    @property
    def n_domains(self):
        return self._nd

    def factor_segments(self):
        """Block structure of the HeadMat in triangles-first order for the
        factorisation and the solves: segments (each mesh's triangle unknowns
        in mesh order, then each mesh's vertex unknowns) and the
        lower-triangle pattern of the matrix over them, from the mesh-pair
        coupling (operators of assembleHeadMat.cpp:73-135: S between
        triangle sets, N between vertex sets, D/D* between the triangles of
        one mesh and the vertices of the other; the deflation couples the
        outermost meshes of a part).  Returns (starts, hfill) with hfill an
        nseg x nseg uint8 array, or None when a mesh's vertex unknowns are
        not one contiguous range."""
        nt, nv = self._nt, self._nv
        tseg, vseg = {}, {}
        for m in range(self._nm):
            inf = self.mesh_info(m)
            if inf["nverts"] > 0:
                if inf["vert_first"] < 0:
                    return None
                vseg[m] = (inf["vert_first"] + nt, inf["vert_end"] + nt)
            if inf["ntris"] > 0 and not inf["current_barrier"]:   # barrier meshes have no triangle unknowns
                tseg[m] = (inf["tri_first"] - nv, inf["tri_first"] - nv + inf["ntris"])
        segs = sorted([(a, b, "T", m) for m, (a, b) in tseg.items()] + [(a, b, "V", m) for m, (a, b) in vseg.items()])
        if not segs or segs[0][0] != 0 or segs[-1][1] != self.size:
            return None
        for (a0, b0, _, _), (a1, _, _, _) in zip(segs, segs[1:]):
            if b0 != a1:
                return None
        idx = {(k, m): i for i, (_, _, k, m) in enumerate(segs)}
        ns = len(segs)
        adj = np.eye(ns, dtype=bool)
        def link(a, b):
            if a in idx and b in idx:
                adj[idx[a], idx[b]] = adj[idx[b], idx[a]] = True
        for p in self.mesh_pairs():
            m1, m2 = p["meshes"]
            link(("T", m1), ("T", m2))
            link(("V", m1), ("V", m2))
            link(("T", m1), ("V", m2))
            link(("T", m2), ("V", m1))
        outer = {}
        for m in range(self._nm):
            inf = self.mesh_info(m)
            if inf["outermost"]:
                outer.setdefault(inf["part"], []).append(m)
        for ms in outer.values():
            for a in ms:
                for b in ms:
                    link(("V", a), ("V", b))
        hfill = np.tril(adj, -1).astype(np.uint8)
        return np.array([a for a, _, _, _ in segs], dtype=np.int64), hfill

    def factor_order(self):
        """Segment order of the factorisation: each mesh's triangle segment,
        then its vertex segment, mesh by mesh in the geometry's order, as a
        permutation of factor_segments()' segments; or None when the meshes
        do not form a chain (every communicating pair neighbours in mesh
        order, as in a nested head) or no reordering is possible.  Over a
        chain of meshes this order makes the HeadMat block tridiagonal, and
        its pivot-free factor fills no block outside the band (the
        triangles-first order fills the blocks between every vertex segment
        and the later triangle segments): at 4 layers a quarter fewer
        trailing-update flops."""
        seg = self.factor_segments()
        if seg is None:
            return None
        for p in self.mesh_pairs():
            m1, m2 = p["meshes"]
            if abs(m1 - m2) > 1:
                return None
        nt, nv = self._nt, self._nv
        where = {}
        for m in range(self._nm):
            inf = self.mesh_info(m)
            if inf["ntris"] > 0 and not inf["current_barrier"]:
                where[(m, 0)] = inf["tri_first"] - nv
            if inf["nverts"] > 0:
                where[(m, 1)] = inf["vert_first"] + nt
        starts = seg[0].tolist()
        index = {a: i for i, a in enumerate(starts)}
        order = [index[where[key]] for key in sorted(where)]
        if sorted(order) != list(range(len(starts))) or order == list(range(len(starts))):
            return None
        return np.array(order, dtype=np.int32)

    def tri_first_index(self, ref_rows):
        """Map reference-order unknown indices to triangles-first order."""
        r = np.asarray(ref_rows, dtype=np.int64)
        return np.where(r >= self._nv, r - self._nv, r + self._nt)


def _mesh_arrays(meshes):
    pts, tris = [], []
    for m in meshes:
        if isinstance(m, dict):
            v, t = m["vertices"], m["triangles"]
        else:
            v, t = m
        pts.append(np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3))
        tris.append(np.ascontiguousarray(t, dtype=np.int64).reshape(-1, 3))
    return pts, tris


def make_nested_geometry(meshes, conductivity):
    """_make_geometry.py:131-211: meshes (vertices, triangles) inner to outer."""
    if not isinstance(meshes, (list, tuple)) or len(meshes) < 1:
        raise ValueError("Wrong argument (should be a non-empty list of meshes)")
    if len(conductivity) != len(meshes):
        raise ValueError("one conductivity per layer is needed")
    pts, tris = _mesh_arrays(meshes)
    L = len(pts)
    nv = (C.c_int64 * L)(*[len(p) for p in pts])
    nt = (C.c_int64 * L)(*[len(t) for t in tris])
    pp = (_lib.p_f64 * L)(*[p.ctypes.data_as(_lib.p_f64) for p in pts])
    tp = (_lib.p_i64 * L)(*[t.ctypes.data_as(_lib.p_i64) for t in tris])
    cond = (C.c_double * L)(*[float(c) for c in conductivity])
    g = Geometry()
    check(lib().omh_geometry_nested(L, nv, pp, nt, tp, cond, C.byref(g._h)))
    g._load_sizes()
    return g


def make_geometry(meshes, interfaces, domains):
    """_make_geometry.py:15-128.  meshes: {name: (vertices, triangles)};
    interfaces: {name: [(mesh name, +1/-1)]}; domains: {name:
    ([(interface name, 'Inside'|'Outside')], conductivity)}."""
    if not isinstance(meshes, dict) or not meshes:
        raise ValueError("Wrong argument (should be a non empty dictionary of named meshes)")
    if not isinstance(interfaces, dict) or not interfaces:
        raise ValueError("Wrong argument (should be a non empty dictionary of named interfaces)")
    if not isinstance(domains, dict) or not domains:
        raise ValueError("Wrong argument (should be a non empty dictionary of named domains)")
    mnames = list(meshes)
    pts, tris = _mesh_arrays([meshes[k] for k in mnames])
    inames = list(interfaces)
    ioff, imesh, iori = [0], [], []
    for k in inames:
        for mname, o in interfaces[k]:
            if mname not in mnames:
                raise ValueError("interface %s uses unknown mesh %s" % (k, mname))
            imesh.append(mnames.index(mname))
            iori.append(int(o))
        ioff.append(len(imesh))
    dnames = list(domains)
    doff, dint, dins, dcond = [0], [], [], []
    for k in dnames:
        bnds, c = domains[k]
        for iname, side in bnds:
            if iname not in inames:
                raise ValueError("Domain %s contains an unknown interface %s." % (k, iname))
            if side not in ("Inside", "Outside", 0, 1):
                raise ValueError("Domain %s: interface %s has a wrong side direction (In/Out)" % (k, iname))
            dint.append(inames.index(iname))
            dins.append(1 if side in ("Inside", 0) else 0)
        doff.append(len(dint))
        dcond.append(float(c))
    M, I, D = len(mnames), len(inames), len(dnames)
    enc = lambda names: (C.c_char_p * len(names))(*[n.encode() for n in names])
    arr = lambda ctype, v: (ctype * max(1, len(v)))(*v)
    g = Geometry()
    check(lib().omh_geometry_create(
        M, enc(mnames), arr(C.c_int64, [len(p) for p in pts]),
        (_lib.p_f64 * M)(*[p.ctypes.data_as(_lib.p_f64) for p in pts]),
        arr(C.c_int64, [len(t) for t in tris]),
        (_lib.p_i64 * M)(*[t.ctypes.data_as(_lib.p_i64) for t in tris]),
        I, enc(inames), arr(C.c_int64, ioff), arr(C.c_int32, imesh), arr(C.c_int32, iori),
        D, enc(dnames), arr(C.c_double, dcond), arr(C.c_int64, doff), arr(C.c_int32, dint),
        arr(C.c_int32, dins), C.byref(g._h)))
    g._load_sizes()
    return g


def read_geometry(geom_file, cond_file):
    return Geometry(geom_file, cond_file)


# ---------------------------------------------------------------------------
# Matrices
# ---------------------------------------------------------------------------
class Matrix:
    """Dense column-major FP64 matrix (matrix.h:81-92), host numpy storage."""

    def __init__(self, a=None, ncol=None):
        if a is None:
            self._a = np.zeros((0, 0), order="F")
        elif isinstance(a, (int, np.integer)):
            self._a = np.zeros((int(a), int(ncol)), order="F")
        elif isinstance(a, Matrix):
            self._a = a._a
        elif isinstance(a, (SymMatrix, SparseMatrix)):   # Matrix(const SymMatrix&), matrix.h:90-92
            self._a = np.asfortranarray(a.array())
        elif isinstance(a, (str, os.PathLike)):           # Matrix(filename)
            self._a = load_matrix(str(a))
        else:
            arr = np.asarray(a)
            if arr.dtype != np.float64 or arr.ndim != 2:
                raise TypeError("Matrix needs a 2-D float64 array")
            self._a = np.asfortranarray(arr)

    def nlin(self):
        return self._a.shape[0]

    def ncol(self):
        return self._a.shape[1]

    def array(self):
        return self._a

    def load(self, fname):
        self._a = load_matrix(fname)

    def save(self, fname):
        save_matrix(fname, self._a)


class SparseMatrix:
    """Head2EEGMat storage: three (column, value) pairs per row
    (sparse_matrix.h:29 holds a map; the EEG interpolation has 3 per row)."""

    def __init__(self, nlin, ncol, cols, vals):
        self._n = (int(nlin), int(ncol))
        self.cols = np.asarray(cols, dtype=np.int64).reshape(-1, 3)
        self.vals = np.asarray(vals, dtype=np.float64).reshape(-1, 3)

    def nlin(self):
        return self._n[0]

    def ncol(self):
        return self._n[1]

    def array(self):
        E = np.zeros(self._n)
        for k in range(3):
            np.add.at(E, (np.arange(self._n[0]), self.cols[:, k]), self.vals[:, k])
        return E

    def save(self, fname):
        """SparseMatrix::save: (row, col, value) entries, .mat / .bin / .txt."""
        from . import io as omio
        rows = np.repeat(np.arange(self._n[0], dtype=np.int64), 3)
        omio.write_sparse(fname, rows, self.cols.ravel(), self.vals.ravel(), self._n)

    @classmethod
    def load(cls, fname):
        """A Head2EEGMat file (three entries per row)."""
        from . import io as omio
        kind, (r, c, v, shape) = omio.read(fname)
        if kind != "sparse":
            raise ValueError("%s holds a %s matrix, not a sparse one" % (fname, kind))
        if len(r) != 3 * shape[0] or np.any(r != np.repeat(np.arange(shape[0]), 3)):
            raise ValueError("%s: a Head2EEGMat has three entries per row" % fname)
        return cls(shape[0], shape[1], c.reshape(-1, 3), v.reshape(-1, 3))


class SymMatrix:
    """Symmetric FP64 matrix.  The reference stores it packed upper
    (symmatrix.h:52-62); here it lives on the device as a full matrix in the
    triangles-first unknown order (nv = number of vertex unknowns, known
    from the geometry), and is packed on the host on demand."""

    def __init__(self, n=0, nv=None):
        self._n = int(n)
        self.nv = nv
        self._dev = None       # torch float64 (n*n), column-major, triangles-first
        self._packed = None    # numpy packed upper, reference order

    @classmethod
    def _from_device(cls, dev, n, nv):
        m = cls(n, nv)
        m._dev = dev
        return m

    @classmethod
    def from_packed(cls, packed, nv=None):
        packed = np.ascontiguousarray(packed, dtype=np.float64)
        n = int((np.sqrt(8 * packed.size + 1) - 1) / 2)
        if n * (n + 1) // 2 != packed.size:
            raise ValueError("not a packed symmetric matrix")
        m = cls(n, nv)
        m._packed = packed
        return m

    def nlin(self):
        return self._n

    def ncol(self):
        return self._n

    def _ordering(self):
        return ORDER_TRIANGLES_FIRST if self.nv is not None else ORDER_REFERENCE

    def device(self):
        """The full device matrix (float64, n*n column-major)."""
        if self._dev is None:
            if self._packed is None:
                raise ValueError("empty SymMatrix")
            P = _to_dev(self._packed)
            D = _dev_f64(self._n * self._n)
            nv = self.nv if self.nv is not None else self._n
            check(lib().omh_sym_unpack(_ptr(P), self._n, nv, _ptr(D), self._n, self._ordering(), _stream()))
            self._dev = D
        return self._dev

    def data(self):
        """Packed upper storage in reference order (numpy)."""
        if self._packed is None:
            n = self._n
            P = _dev_f64(n * (n + 1) // 2)
            nv = self.nv if self.nv is not None else n
            check(lib().omh_sym_pack(_ptr(self._dev), n, nv, n, self._ordering(), _ptr(P), _stream()))
            self._packed = P.cpu().numpy()
        return self._packed

    def array(self):
        n = self._n
        P = self.data()
        j = np.repeat(np.arange(n), np.arange(1, n + 1))
        i = np.concatenate([np.arange(k + 1) for k in range(n)]) if n else np.zeros(0, np.int64)
        M = np.zeros((n, n), order="F")
        M[i, j] = P
        M[j, i] = P
        return M

    def _factor(self, precision="fp64"):
        if self.nv is None:
            raise ValueError("SymMatrix.invert/solveLin need the unknown layout of a HeadMat (nv)")
        n = self._n
        return factor(self.device().clone(), n, precision, n_positive=n - self.nv,
                      rebuild=lambda: self.device().clone())

    def inverse(self):
        """SymMatrix::inverse (symmatrix.h:259-278)."""
        n = self._n
        F = self._factor()
        X = _t().eye(n, dtype=_t().float64, device="cuda").reshape(-1).clone()
        F.solve(X, n)
        return SymMatrix._from_device(X, n, self.nv)

    def invert(self):
        """SymMatrix::invert (symmatrix.h:280-297), in place."""
        inv = self.inverse()
        self._dev = inv._dev
        self._packed = None

    def solveLin(self, rhs):
        """SymMatrix::solveLin(Matrix&) (symmatrix.cpp:133-152)."""
        B = rhs.array() if isinstance(rhs, Matrix) else np.asarray(rhs, dtype=np.float64)
        B = B.reshape(self._n, -1)
        F = self._factor()
        Bd = to_device_rows(B, self._n, self.nv)
        F.solve(Bd, B.shape[1])
        return Matrix(from_device_rows(Bd, self._n, self.nv, B.shape[1]))

    def save(self, fname):
        save_symmatrix(fname, self.data())

    def load(self, fname):
        P = load_symmatrix(fname)
        m = SymMatrix.from_packed(P, self.nv)
        self.__dict__.update(m.__dict__)


def to_device_rows(B, n, nv):
    """Host matrix (n x c, reference rows) -> device column-major buffer in
    triangles-first row order."""
    B = np.asarray(B, dtype=np.float64).reshape(n, -1)
    c = B.shape[1]
    src = _to_dev(np.asfortranarray(B).ravel(order="F"))
    dst = _dev_f64(n * c)
    check(lib().omh_permute_rows(_ptr(src), n, _ptr(dst), n, n, nv, c, ORDER_TRIANGLES_FIRST, 0, _stream()))
    return dst


def from_device_rows(D, n, nv, c):
    dst = _dev_f64(n * c)
    check(lib().omh_permute_rows(_ptr(D), n, _ptr(dst), n, n, nv, c, ORDER_TRIANGLES_FIRST, 1, _stream()))
    return dst.cpu().numpy().reshape(c, n).T.copy(order="F")


class Factor:
    """A stored factor of a device HeadMat: the pivot-free signed Cholesky
    (precision 'fp64', or 'mixed' for an FP32 factor refined against the FP64
    matrix) or, after a breakdown of it, the FP64 LU with partial pivoting
    ('lu'; `pivoted` is True).

    A mixed-precision factor whose refinement does not converge (the FP32
    factor too coarse for the matrix) is replaced by the FP64 factor of the
    FP64 matrix when fallback_fp64 is set (then `fell_back` is True);
    otherwise solve() raises NotConverged."""

    def __init__(self, L, n, work, precision, H64=None, structure=None, fallback_fp64=True, ipiv=None,
                 n_positive=None, order=None):
        self.L, self.n, self.work, self.precision, self.H64 = L, n, work, precision, H64
        self.ipiv = ipiv
        self.iterations = 0
        self.rel_update = 0.0
        self.fell_back = False
        self.pivoted = precision == "lu"
        self.fallback_fp64 = fallback_fp64
        self.n_positive = n_positive
        self._raw_structure = structure
        self._structure = _block_args(structure)
        # segment order of the factor (None: the matrix's own order); the
        # factor's block structure in that order
        self.order = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        self._fstructure = _block_args(_reorder_structure(structure, self.order, n)) if self.order is not None else None

    def _refactor_fp64(self):
        """Factor the FP64 matrix in place (it is no longer needed for
        residuals) and drop the FP32 factor; the LU with partial pivoting
        when the pivot-free factorisation breaks down on it."""
        H = self.H64
        backup = H.clone() if self.n_positive is not None else None
        F = factor(H, self.n, "fp64", structure=self._raw_structure, n_positive=self.n_positive,
                   rebuild=(lambda: backup) if backup is not None else None)
        self.L, self.work, self.precision, self.H64, self.ipiv = F.L, F.work, F.precision, None, F.ipiv
        self.order, self._fstructure = F.order, F._fstructure
        self.pivoted = F.pivoted
        self.fell_back = True

    def solve(self, B, nrhs, max_iter=10, tol=1e-12):
        """B <- H^-1 B (float64 device buffer, n x nrhs column-major).  With
        the block structure the factor was made with, the solves and the
        residuals skip its zero blocks."""
        if self.precision == "mixed":
            try:
                return self._solve(B, nrhs, max_iter, tol)
            except _lib.NotConverged:
                if not self.fallback_fp64:
                    raise
                self._refactor_fp64()
        return self._solve(B, nrhs, max_iter, tol)

    def _solve(self, B, nrhs, max_iter, tol):
        blocks = self._structure[2:] if self._structure else ()
        if self.precision == "lu":
            check(lib().omh_lu_solve_f64(_ptr(self.L), self.n, self.n, _ptr(self.ipiv), _ptr(self.work), _ptr(B),
                                         self.n, nrhs, _stream()))
            return B
        if self.precision == "fp64":
            if self.order is not None:
                # rows into the factor's order, solve there, rows back
                starts, optr = self._structure[3], self.order.ctypes.data_as(C.c_void_p)
                nseg = self._structure[2]
                T = _dev_f64(self.n * nrhs)
                check(lib().omh_reorder_rows_f64(_ptr(B), self.n, _ptr(T), self.n, self.n, nrhs, nseg, starts, optr,
                                                 0, _stream()))
                check(lib().omh_sym_solve_f64_blocks(_ptr(self.L), self.n, self.n, _ptr(self.work), _ptr(T), self.n,
                                                     nrhs, *self._fstructure[2:], _stream()))
                check(lib().omh_reorder_rows_f64(_ptr(T), self.n, _ptr(B), self.n, self.n, nrhs, nseg, starts, optr,
                                                 1, _stream()))
            elif blocks:
                check(lib().omh_sym_solve_f64_blocks(_ptr(self.L), self.n, self.n, _ptr(self.work), _ptr(B), self.n,
                                                     nrhs, *blocks, _stream()))
            else:
                check(lib().omh_sym_solve_f64(_ptr(self.L), self.n, self.n, _ptr(self.work), _ptr(B), self.n, nrhs,
                                              _stream()))
            return B
        X = _dev_f64(self.n * nrhs)
        it, rel = C.c_int32(), C.c_double()
        if self.order is not None:
            rc = lib().omh_mixed_solve_reordered(_ptr(self.H64), self.n, self.n, _ptr(self.L), self.n,
                                                 _ptr(self.work), _ptr(B), self.n, _ptr(X), self.n, nrhs, max_iter,
                                                 tol, *blocks, self.order.ctypes.data_as(C.c_void_p), C.byref(it),
                                                 C.byref(rel), _stream())
        elif blocks:
            rc = lib().omh_mixed_solve_blocks(_ptr(self.H64), self.n, self.n, _ptr(self.L), self.n, _ptr(self.work),
                                              _ptr(B), self.n, _ptr(X), self.n, nrhs, max_iter, tol, *blocks,
                                              C.byref(it), C.byref(rel), _stream())
        else:
            rc = lib().omh_mixed_solve(_ptr(self.H64), self.n, self.n, _ptr(self.L), self.n, _ptr(self.work),
                                       _ptr(B), self.n, _ptr(X), self.n, nrhs, max_iter, tol, C.byref(it),
                                       C.byref(rel), _stream())
        self.iterations, self.rel_update = it.value, rel.value
        check(rc)
        B.copy_(X)
        return B


def _block_args(structure):
    """(starts, hfill, nseg, starts pointer, hfill pointer) for the *_blocks
    entry points (the arrays are kept with the pointers), or None."""
    if structure is None:
        return None
    starts = np.ascontiguousarray(structure[0], dtype=np.int64)
    hfill = np.ascontiguousarray(structure[1], dtype=np.uint8)
    if hfill.shape != (len(starts), len(starts)):
        raise ValueError("hfill must be nseg x nseg")
    return (starts, hfill, len(starts), starts.ctypes.data_as(C.c_void_p), hfill.ctypes.data_as(C.c_void_p))


def _reorder_structure(structure, order, n):
    """The block structure (starts, hfill) of the matrix with its segments
    taken in `order`: segment q of the reordered matrix is segment order[q];
    the pattern is read from hfill's lower triangle, as the *_blocks entry
    points do (omh_mixed_solve_reordered derives the same one)."""
    starts = np.asarray(structure[0], dtype=np.int64)
    hfill = np.tril(np.asarray(structure[1], dtype=np.uint8) != 0, -1)
    o = np.asarray(order, dtype=np.int64)
    lens = np.append(starts[1:], n) - starts
    fstarts = np.concatenate([[0], np.cumsum(lens[o])[:-1]]).astype(np.int64)
    adj = hfill | hfill.T
    return fstarts, np.tril(adj[np.ix_(o, o)], -1).astype(np.uint8)


def _segment_signs(structure, order, n_positive):
    """Expected pivot sign per segment of the reordered matrix: + for the
    unknowns below n_positive in the matrix's order (a HeadMat's triangle
    unknowns), - for the rest."""
    starts = np.asarray(structure[0], dtype=np.int64)
    sg = np.where(starts < n_positive, 1, -1).astype(np.int8)
    return np.ascontiguousarray(sg[np.asarray(order, dtype=np.int64)])


def _room_for(nbytes):
    """True when the device can hold nbytes more (free memory plus torch's
    cached blocks) with 4 GiB to spare."""
    torch = _t()
    free, _ = torch.cuda.mem_get_info()
    cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
    return free + cached >= nbytes + (4 << 30)


def _lu_factor(Hdev, n):
    """FP64 LU with partial pivoting of Hdev in place (the pivoted fallback)."""
    torch = _t()
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    work = torch.empty(int(lib().omh_lu_workspace_bytes(n)), dtype=torch.uint8, device="cuda")
    info = C.c_int32()
    check(lib().omh_lu_factor_f64(_ptr(Hdev), n, n, _ptr(ipiv), _ptr(work), C.byref(info), _stream()))
    if info.value != 0:
        raise _lib.OpenMEEGError("the matrix is singular: zero pivot in column %d of the pivoted LU" % (info.value - 1))
    return Factor(Hdev, n, work, "lu", ipiv=ipiv)


def factor(Hdev, n, precision="fp64", structure=None, n_positive=None, rebuild=None, order=None):
    """Factor a device HeadMat (triangles-first order).  precision 'fp64'
    factors in place; 'mixed' keeps Hdev as the FP64 residual operator and
    factors an FP32 copy; 'mixed16' does the same with split-FP16 MFMA
    trailing updates; 'fp64_lu' goes straight to the pivoted LU.  Given
    structure = Geometry.factor_segments(), the trailing updates, and later
    the solves, skip the parts that only multiply structural zeros (the
    factor is the same).

    order (with structure; Geometry.factor_order()): factor the matrix with
    its segments in that order instead (the solves take right-hand sides and
    return solutions in Hdev's order).  The mixed paths build their FP32 copy
    reordered; 'fp64' factors a reordered FP64 copy when the device has room
    for one (Hdev is then left intact), else Hdev in place in its own order.
    A reordered factorisation that breaks down is redone in Hdev's order.

    The pivot-free factorisation is used when it is stable: it reports a
    vanishing pivot, and given n_positive (the triangle unknowns, which come
    first) any pivot whose sign differs from + for the first n_positive
    unknowns and - for the rest also counts as a breakdown.  Then the FP64
    matrix is factored by LU with partial pivoting (the reference's DSPTRF
    pivots too, symmatrix.cpp:133-152): for the mixed paths Hdev itself,
    which is intact; for 'fp64', whose failed factorisation overwrote Hdev,
    the matrix rebuild() returns (else OpenMEEGError)."""
    torch = _t()
    if precision == "fp64_lu":
        return _lu_factor(Hdev, n)
    if order is not None and structure is None:
        raise ValueError("a factor order needs the block structure")
    if order is not None and precision == "fp64" and not _room_for(8 * n * n):
        order = None
    info = C.c_int32()
    sargs = _block_args(structure)
    blocks = sargs[2:] if sargs else ()
    if order is not None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        fargs = _block_args(_reorder_structure(structure, order, n))
        fblocks = fargs[2:]
        optr = order.ctypes.data_as(C.c_void_p)
    if precision == "fp64":
        work = torch.empty(int(lib().omh_factor_workspace_bytes(n, 8)), dtype=torch.uint8, device="cuda")
        if order is not None:
            L = torch.empty(n * n, dtype=torch.float64, device="cuda")
            check(lib().omh_sym_reorder(_ptr(Hdev), n, _ptr(L), n, 8, n, blocks[0], blocks[1], optr, _stream()))
            check(lib().omh_sym_factor_f64_blocks(_ptr(L), n, n, _ptr(work), *fblocks, C.byref(info), _stream()))
        elif blocks:
            L = Hdev
            check(lib().omh_sym_factor_f64_blocks(_ptr(Hdev), n, n, _ptr(work), *blocks, C.byref(info), _stream()))
        else:
            L = Hdev
            check(lib().omh_sym_factor_f64(_ptr(Hdev), n, n, _ptr(work), C.byref(info), _stream()))
        F = Factor(L, n, work, "fp64", structure=structure, n_positive=n_positive, order=order)
        elem = 8
    elif precision in ("mixed", "mixed16"):
        L32 = torch.empty(n * n, dtype=torch.float32, device="cuda")
        if order is not None:
            check(lib().omh_sym_reorder(_ptr(Hdev), n, _ptr(L32), n, 4, n, blocks[0], blocks[1], optr, _stream()))
        else:
            check(lib().omh_convert_f64_f32(_ptr(Hdev), n, _ptr(L32), n, n, n, 1, _stream()))
        work = torch.empty(int(lib().omh_factor_workspace_bytes(n, 4)), dtype=torch.uint8, device="cuda")
        if blocks:
            fn = lib().omh_sym_factor_f32_blocks if precision == "mixed" else lib().omh_sym_factor_f32_split_blocks
            check(fn(_ptr(L32), n, n, _ptr(work), *(fblocks if order is not None else blocks), C.byref(info),
                     _stream()))
        else:
            fn = lib().omh_sym_factor_f32 if precision == "mixed" else lib().omh_sym_factor_f32_split
            check(fn(_ptr(L32), n, n, _ptr(work), C.byref(info), _stream()))
        F = Factor(L32, n, work, "mixed", H64=Hdev, structure=structure, n_positive=n_positive, order=order)
        elem = 4
    else:
        raise ValueError("precision must be 'fp64', 'mixed', 'mixed16' or 'fp64_lu'")
    reason = None
    if info.value != 0:
        reason = "vanishing pivot %d" % (info.value - 1)
    elif n_positive is not None:
        first = C.c_int64()
        if order is not None:
            sg = _segment_signs(structure, order, n_positive)
            check(lib().omh_factor_sign_mismatch_segments(_ptr(F.work), n, elem, fblocks[0], fblocks[1],
                                                          sg.ctypes.data_as(C.c_void_p), C.byref(first), _stream()))
        else:
            check(lib().omh_factor_sign_mismatch(_ptr(F.work), n, elem, int(n_positive), C.byref(first), _stream()))
        if first.value >= 0:
            reason = "pivot sign at unknown %d" % first.value
    if reason is None:
        return F
    del F
    if order is not None:
        # Hdev is intact: the factorisation in its own order
        F = factor(Hdev, n, precision, structure, n_positive, rebuild)
        F.reorder_breakdown = reason
        return F
    if precision == "fp64":
        if rebuild is None:
            raise _lib.OpenMEEGError("pivot-free factorisation broke down (%s)" % reason)
        Hdev = rebuild()
    F = _lu_factor(Hdev, n)
    F.breakdown = reason
    return F


# ---------------------------------------------------------------------------
# Assembly (assemble.h:32-35, assembleSourceMat.h:24-29)
# ---------------------------------------------------------------------------
HEADMAT_FAR_FIELD = 1


def _head_factor(geom, H, precision, integrator=None, far_field=False):
    """factor() of a pipeline's HeadMat H with the geometry's block
    structure, its mesh-by-mesh factor order (Geometry.factor_order; the
    environment variable OMH_FACTOR_ORDER=0 keeps the triangles-first order)
    and pivot-sign pattern (triangle unknowns +); an FP64 breakdown
    re-assembles H for the pivoted LU."""
    n = geom.size
    structure = geom.factor_segments()
    order = None
    if structure is not None and os.environ.get("OMH_FACTOR_ORDER", "1") != "0":
        order = geom.factor_order()
    return factor(H, n, precision, structure=structure, n_positive=n - geom.n_vertex_unknowns,
                  rebuild=lambda: _head_mat_device(geom, integrator, far_field), order=order)


def _head_mat_device(geom, integrator=None, far_field=False):
    """Device HeadMat (triangles-first order).  far_field: Gauss quadrature on
    both triangles of well-separated pairs (the mixed-precision path)."""
    integ = integrator or Integrator(3, 0, 0.005)
    n = geom.size
    H = _dev_f64(n * n)
    if integ.levels > 0:   # adaptive integration (the analytic pair integrals, refined)
        check(lib().omh_head_mat_integrator(geom._h, integ.order, integ.levels, integ.tolerance, _ptr(H), n,
                                            ORDER_TRIANGLES_FIRST, _stream()))
        return H
    check(lib().omh_head_mat_ex(geom._h, integ.order, HEADMAT_FAR_FIELD if far_field else 0, _ptr(H), n,
                                ORDER_TRIANGLES_FIRST, _stream()))
    return H


def HeadMat(geom, integrator=None):
    """HeadMat(geo, Integrator(3,0,0.005)) -- assemble.h:32."""
    return SymMatrix._from_device(_head_mat_device(geom, integrator), geom.size, geom.n_vertex_unknowns)


def _dipoles_array(dipoles):
    a = dipoles.array() if isinstance(dipoles, Matrix) else np.asarray(dipoles, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError("Dipoles must have 6 columns (x y z qx qy qz)")
    return np.ascontiguousarray(a)


def _dsm_device(geom, dipoles, domain_name="", integrator=None):
    integ = integrator or Integrator(3, 10, 0.001)
    dip = _dipoles_array(dipoles)
    n, nd = geom.size, len(dip)
    D = _dev_f64(n * max(nd, 1))
    dd = _to_dev(dip)
    doms = None
    if domain_name:
        k = geom.domain_index(domain_name)
        doms = (C.c_int32 * nd)(*([k] * nd))
    check(lib().omh_dip_source_mat(geom._h, _ptr(dd), nd, doms, integ.order, integ.levels, integ.tolerance, _ptr(D),
                                   n, ORDER_TRIANGLES_FIRST, _stream()))
    return D


def DipSourceMat(geom, dipoles, domain_name="", integrator=None):
    """DipSourceMat(geom, dipoles, domain) -- _openmeeg.i:399-413,
    SourceMatrix<Dipole> with the adaptive Integrator(3,10,0.001)."""
    D = _dsm_device(geom, dipoles, domain_name, integrator)
    nd = len(_dipoles_array(dipoles))
    return Matrix(from_device_rows(D, geom.size, geom.n_vertex_unknowns, nd))


def _electrodes_array(sensors):
    if isinstance(sensors, Sensors):
        return sensors.positions
    a = sensors.array() if isinstance(sensors, Matrix) else np.asarray(sensors, dtype=np.float64)
    return np.ascontiguousarray(a[:, :3])


def Head2EEGMat(geom, sensors):
    """Head2EEGMat(geo, electrodes) -- assemble.h:35."""
    el = _electrodes_array(sensors)
    ne = len(el)
    cols = np.zeros((ne, 3), dtype=np.int64)
    vals = np.zeros((ne, 3))
    if ne:
        de = _to_dev(el)
        check(lib().omh_head2eeg_mat(geom._h, _ptr(de), ne, cols.ctypes.data_as(_lib.p_i64),
                                     vals.ctypes.data_as(_lib.p_f64), _stream()))
    return SparseMatrix(ne, geom.size, cols, vals)


def _eeg_rhs(geom, h2em):
    """E^T (n x n_e, triangles-first rows) on the device, built there from the
    3 n_e non-zeros of the Head2EEGMat (gain.h:28-49)."""
    n, ne = geom.size, h2em.nlin()
    rows = _to_dev(np.ascontiguousarray(geom.tri_first_index(h2em.cols), dtype=np.int64).ravel(), np.int64)
    vals = _to_dev(np.ascontiguousarray(h2em.vals, dtype=np.float64).ravel())
    B = _dev_f64(n * ne)
    check(lib().omh_sparse_transpose_dense(_ptr(rows), _ptr(vals), ne, n, _ptr(B), n, _stream()))
    return B


def _gemm_tn(Y, D, n, ne, nd):
    G = _dev_f64(ne * nd)
    check(lib().omh_gemm_f64(1, 0, ne, nd, n, 1.0, _ptr(Y), n, _ptr(D), n, 0.0, _ptr(G), ne, _stream()))
    return G


def GainEEG(hminv, dsm, h2em):
    """GainEEG(HeadMatInv, SourceMat, Head2EEGMat) = (E H^-1) D -- gain.h:69-76."""
    n = hminv.nlin()
    nv = hminv.nv
    if nv is None:
        raise ValueError("GainEEG needs the HeadMat unknown layout (a SymMatrix from HeadMat)")
    Hi = hminv.device()
    ne = h2em.nlin()
    rows = np.where(h2em.cols >= nv, h2em.cols - nv, h2em.cols + (n - nv))
    dc = _to_dev(rows.ravel(), np.int64)
    dv = _to_dev(h2em.vals.ravel())
    EH = _dev_f64(ne * n)
    check(lib().omh_sparse_rows(_ptr(dc), _ptr(dv), ne, _ptr(Hi), n, n, _ptr(EH), ne, _stream()))
    D = dsm.array()
    Dd = to_device_rows(D, n, nv)
    nd = D.shape[1]
    G = _dev_f64(ne * nd)
    check(lib().omh_gemm_f64(0, 0, ne, nd, n, 1.0, _ptr(EH), ne, _ptr(Dd), n, 0.0, _ptr(G), ne, _stream()))
    return Matrix(G.cpu().numpy().reshape(nd, ne).T.copy(order="F"))


def GainEEGadjoint(geom, dipoles, hm, h2em, precision="fp64"):
    """GainEEGadjoint(geo, dipoles, HeadMat, Head2EEGMat) -- gain.h:78-89:
    Y = H^-1 E^T with the stored factor, then G = Y^T DSM."""
    n, ne = geom.size, h2em.nlin()
    F = hm._factor(precision)
    Y = _eeg_rhs(geom, h2em)
    F.solve(Y, ne)
    D = _dsm_device(geom, dipoles, "")
    nd = len(_dipoles_array(dipoles))
    G = _gemm_tn(Y, D, n, ne, nd)
    return Matrix(G.cpu().numpy().reshape(nd, ne).T.copy(order="F"))


class Sensors:
    """Sensors (sensors.h:76-216): Sensors::load (sensors.cpp:36-143) for
    labelled or unlabelled files -- positions (EEG electrodes), positions and
    orientations (6 columns), or MEG integration points with a weight (7
    columns: label x y z dx dy dz w; points with the same label are one
    sensor) -- or Sensors(positions)."""

    def __init__(self, positions=None, geom=None, orientations=None, weights=None):
        if isinstance(positions, (str, os.PathLike)):      # Sensors(filename), sensors.h:82-88
            self.__init__()
            self.load(str(positions))
            return
        self.labels = []
        self.positions = np.zeros((0, 3))
        self.radii = np.zeros(0)
        self.orientations = None
        self.weights = np.zeros(0)
        self.point_sensor = np.zeros(0, dtype=np.int64)
        if positions is not None:
            a = positions.array() if isinstance(positions, Matrix) else np.asarray(positions, dtype=np.float64)
            self.positions = np.ascontiguousarray(a[:, :3], dtype=np.float64)
            self.radii = np.zeros(len(self.positions))
            self._plain(len(self.positions))
        if orientations is not None:
            o = orientations.array() if isinstance(orientations, Matrix) else np.asarray(orientations, dtype=np.float64)
            if o.shape != self.positions.shape:
                raise ValueError("orientations must have one row of 3 per position")
            self.orientations = np.ascontiguousarray(o, dtype=np.float64)
        if weights is not None:
            w = np.asarray(weights, dtype=np.float64).ravel()
            if len(w) != len(self.positions):
                raise ValueError("one weight per position")
            self.weights = w.copy()

    def _plain(self, npts):
        self.weights = np.ones(npts)
        self.point_sensor = np.arange(npts, dtype=np.int64)

    def load(self, fname):
        rows = []
        with open(fname) as f:
            for ln in f:
                tok = ln.split()
                if tok and not tok[0].startswith("#"):
                    rows.append(tok)
        if not rows:
            raise ValueError("empty sensor file")
        if any(len(r) != len(rows[0]) for r in rows):
            raise ValueError("Problem while reading Sensors file: each line should have the same number of elements")
        # labelled unless some row's first token is a number with one '.'
        # (sensors.cpp:67-69 scans every row)
        labeled = not any(r[0].count(".") == 1 for r in rows)
        names = [r[0] for r in rows] if labeled else []
        if labeled:
            rows = [r[1:] for r in rows]
        mat = np.array(rows, dtype=np.float64)
        ncol = mat.shape[1]
        # 4 columns: EIT electrodes with a radius (sensors.cpp:104-111)
        self.radii = np.ascontiguousarray(mat[:, 3]) if ncol == 4 else np.zeros(len(mat))
        self.positions = np.ascontiguousarray(mat[:, :3])
        self.orientations = np.ascontiguousarray(mat[:, 3:6]) if ncol >= 6 else None
        self.weights = mat[:, 6].copy() if ncol == 7 else np.ones(len(mat))
        # sensor index per point: labels in order of first appearance (sensors.cpp:127-142)
        self.labels = []
        idx = {}
        ps = np.zeros(len(mat), dtype=np.int64)
        for i in range(len(mat)):
            if labeled:
                if names[i] not in idx:
                    idx[names[i]] = len(self.labels)
                    self.labels.append(names[i])
                ps[i] = idx[names[i]]
            else:
                ps[i] = i
        self.point_sensor = ps

    def getNumberOfSensors(self):
        return int(self.point_sensor.max()) + 1 if len(self.point_sensor) else 0

    def getNumberOfPositions(self):
        return len(self.positions)

    def getPositions(self):
        return Matrix(np.asfortranarray(self.positions))

    def hasOrientations(self):
        return self.orientations is not None

    def getOrientations(self):
        if self.orientations is None:
            raise ValueError("these sensors have no orientations")
        return Matrix(np.asfortranarray(self.orientations))

    def getWeights(self):
        return self.weights.copy()

    def getWeightsMatrix(self):
        """Sensors::getWeightsMatrix (sensors.h:191-196), dense: sensors x points."""
        W = np.zeros((self.getNumberOfSensors(), self.getNumberOfPositions()))
        W[self.point_sensor, np.arange(self.getNumberOfPositions())] = self.weights
        return W


def _meg_points(sensors):
    if not isinstance(sensors, Sensors) or sensors.orientations is None:
        raise ValueError("MEG needs Sensors with orientations (a .squids file)")
    return _to_dev(np.ascontiguousarray(sensors.positions)), _to_dev(np.ascontiguousarray(sensors.orientations))


def _sensor_sum(sensors, P, ncols):
    """Weights matrix (sensors x points) times P (points x ncols, device,
    column-major) on the device."""
    npts, ns = sensors.getNumberOfPositions(), sensors.getNumberOfSensors()
    W = _to_dev(np.asfortranarray(sensors.getWeightsMatrix()).ravel(order="F"))
    out = _dev_f64(max(1, ns * ncols))
    check(lib().omh_gemm_f64(0, 0, ns, ncols, npts, 1.0, _ptr(W), ns, _ptr(P), npts, 0.0, _ptr(out), ns,
                             _stream()))
    return out


def _head2meg_device(geom, sensors, ordering=ORDER_TRIANGLES_FIRST):
    """Head2MEGMat on the device (sensors x n, column-major, rows summed over
    the sensors' integration points)."""
    d_pts, d_dirs = _meg_points(sensors)
    npts, n = sensors.getNumberOfPositions(), geom.size
    P = _dev_f64(max(1, npts * n))
    check(lib().omh_head2meg_mat(geom._h, _ptr(d_pts), _ptr(d_dirs), npts, _ptr(P), max(1, npts), ordering,
                                 _stream()))
    return _sensor_sum(sensors, P, n)


def Head2MEGMat(geom, sensors):
    """Head2MEGMat(geo, sensors) -- assemble.h:38, assembleSensors.cpp:69-99."""
    M = _head2meg_device(geom, sensors, ORDER_REFERENCE)
    ns = sensors.getNumberOfSensors()
    return Matrix(M.cpu().numpy()[:ns * geom.size].reshape(geom.size, ns).T.copy(order="F"))


def _ds2meg_device(dipoles, sensors):
    d_pts, d_dirs = _meg_points(sensors)
    dip = _dipoles_array(dipoles)
    npts, nd = sensors.getNumberOfPositions(), len(dip)
    P = _dev_f64(max(1, npts * nd))
    dd = _to_dev(dip)          # held until the call has been enqueued
    check(lib().omh_dip_source2meg_mat(_ptr(dd), nd, _ptr(d_pts), _ptr(d_dirs), npts, _ptr(P),
                                       max(1, npts), _stream()))
    return _sensor_sum(sensors, P, nd)


def DipSource2MEGMat(dipoles, sensors):
    """DipSource2MEGMat(dipoles, sensors) -- assemble.h:41, assembleSensors.cpp:133-156."""
    ns, nd = sensors.getNumberOfSensors(), len(_dipoles_array(dipoles))
    M = _ds2meg_device(dipoles, sensors)
    return Matrix(M.cpu().numpy()[:ns * nd].reshape(nd, ns).T.copy(order="F"))


class _VertexList(list):
    def size(self):
        return len(self)


def _read_tri(path):
    """tri.h:26-56: '- N' then N lines 'x y z nx ny nz', '- T T T' then T
    lines of 0-based vertex indices."""
    with open(path) as f:
        tok = f.read().split()
    if not tok or tok[0] != "-":
        raise ValueError("%s: not a .tri mesh" % path)
    nv = int(tok[1])
    vals = np.array(tok[2:2 + 6 * nv], dtype=np.float64).reshape(nv, 6)
    pos = 2 + 6 * nv
    if tok[pos] != "-":
        raise ValueError("%s: not a .tri mesh" % path)
    nt = int(tok[pos + 1])
    tris = np.array(tok[pos + 4:pos + 4 + 3 * nt], dtype=np.int64).reshape(nt, 3)
    return vals[:, :3].copy(), tris


class Mesh:
    """Mesh (mesh.h) of a distributed source surface: Mesh(filename) for a
    .tri file (Mesh::load, mesh.cpp:255-265), or Mesh(vertices, triangles)
    with 0-based triangles.  Vertex i has index i (Mesh::generate_indices,
    mesh.cpp:266-272)."""

    def __init__(self, vertices=None, triangles=None):
        if isinstance(vertices, (str, os.PathLike)):
            if not str(vertices).endswith(".tri"):
                raise ValueError("only .tri source meshes are read")
            vertices, triangles = _read_tri(vertices)
        self._v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
        self._t = np.ascontiguousarray(np.asarray(triangles, dtype=np.int64).reshape(-1, 3))
        h = C.c_void_p()
        check(lib().omh_mesh_create(self._v.ctypes.data_as(_lib.p_f64), len(self._v),
                                    self._t.ctypes.data_as(_lib.p_i64), len(self._t), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().omh_mesh_free(h)
            self._h = None

    def vertices(self):
        return _VertexList(map(tuple, self._v))

    def triangles(self):
        return self._t.copy()

    def nb_vertices(self):
        return len(self._v)

    def nb_triangles(self):
        return len(self._t)


def mesh_intersects(geom, mesh):
    """Geometry::check(mesh) (geometry.cpp:329-343): True when a triangle of
    the source mesh meets a triangle of the geometry."""
    out = C.c_int32()
    check(lib().omh_mesh_intersects(geom._h, mesh._h, C.byref(out), _stream()))
    return bool(out.value)


def _ssm_device(geom, mesh, integrator=None, ordering=ORDER_TRIANGLES_FIRST, far_field=False):
    integ = integrator or Integrator(3, 0, 0.005)
    if integ.levels != 0:
        raise ValueError("adaptive SurfSourceMat integration is not supported on this path (levels must be 0)")
    n, ns = geom.size, mesh.nb_vertices()
    D = _dev_f64(n * ns)
    check(lib().omh_surf_source_mat(geom._h, mesh._h, integ.order, HEADMAT_FAR_FIELD if far_field else 0, _ptr(D),
                                    n, ordering, _stream()))
    return D


def SurfSourceMat(geom, mesh, integrator=None):
    """SurfSourceMat(geo, sources, Integrator(3,0,0.005)) --
    assembleSourceMat.h:20, assembleSourceMat.cpp:22-56: n x (source
    vertices), rows in the reference unknown order.  Raises
    OverlappingSourceMesh when the mesh meets the geometry."""
    if isinstance(mesh, (str, os.PathLike)):
        mesh = Mesh(mesh)
    D = _ssm_device(geom, mesh, integrator, ORDER_REFERENCE)
    n, ns = geom.size, mesh.nb_vertices()
    return Matrix(D.cpu().numpy().reshape(ns, n).T.copy(order="F"))


def _ss2meg_device(mesh, sensors):
    d_pts, d_dirs = _meg_points(sensors)
    npts, ns = sensors.getNumberOfPositions(), mesh.nb_vertices()
    P = _dev_f64(max(1, npts * ns))
    check(lib().omh_surf_source2meg_mat(mesh._h, _ptr(d_pts), _ptr(d_dirs), npts, _ptr(P), max(1, npts), _stream()))
    return _sensor_sum(sensors, P, ns)


def SurfSource2MEGMat(mesh, sensors):
    """SurfSource2MEGMat(sources, sensors) -- assemble.h:39,
    assembleSensors.cpp:102-124."""
    if isinstance(mesh, (str, os.PathLike)):
        mesh = Mesh(mesh)
    nsens, ns = sensors.getNumberOfSensors(), mesh.nb_vertices()
    M = _ss2meg_device(mesh, sensors)
    return Matrix(M.cpu().numpy()[:nsens * ns].reshape(ns, nsens).T.copy(order="F"))


def _points_array(points):
    a = points.array() if isinstance(points, Matrix) else np.asarray(points, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] < 3:
        raise ValueError("points must have 3 columns (x y z)")
    return np.ascontiguousarray(a[:, :3])


def _surf2vol_device(geom, points, ordering=ORDER_TRIANGLES_FIRST):
    """Surf2VolMat on the device: (kept points x n, column-major, leading
    dimension the number of points) and the kept point indices."""
    pts = _points_array(points)
    npts, n = len(pts), geom.size
    M = _dev_f64(max(1, npts) * n)
    kept = np.zeros(max(1, npts), dtype=np.int64)
    nk = C.c_int64()
    dp = _to_dev(pts)          # held until the call has been enqueued
    check(lib().omh_surf2vol_mat(geom._h, _ptr(dp), npts, _ptr(M), max(1, npts), ordering,
                                 kept.ctypes.data_as(_lib.p_i64), C.byref(nk), _stream()))
    return M, kept[:nk.value]


def Surf2VolMat(geom, points):
    """Surf2VolMat(geo, points) -- assemble.h:33, assembleHeadMat.cpp:325-363:
    the potential at the points inside the head from the unknowns (points in
    a non-conductive domain are dropped)."""
    pts = _points_array(points)
    M, kept = _surf2vol_device(geom, pts, ORDER_REFERENCE)
    ld = max(1, len(pts))
    A = M.cpu().numpy().reshape(geom.size, ld).T[:len(kept)]
    return Matrix(np.asfortranarray(A))


def DipSource2InternalPotMat(geom, dipoles, points, domain_name=""):
    """DipSource2InternalPotMat(geo, dipoles, points, domain) --
    assembleSourceMat.h:32, assembleSourceMat.cpp:154-186."""
    dip = _dipoles_array(dipoles)
    pts = _points_array(points)
    nd, npts = len(dip), len(pts)
    ld = max(1, npts)
    D = _dev_f64(ld * max(1, nd))
    doms = None
    if domain_name:
        doms = (C.c_int32 * nd)(*([geom.domain_index(domain_name)] * nd))
    kept = np.zeros(ld, dtype=np.int64)
    nk = C.c_int64()
    # both device copies are held until the call has been enqueued: a
    # temporary freed inside the argument list can be handed to the next
    # allocation (the points) and overwritten before the kernels run
    dd, dp = _to_dev(dip), _to_dev(pts)
    check(lib().omh_dip_source2internal_pot_mat(geom._h, _ptr(dd), nd, doms, _ptr(dp), npts,
                                                _ptr(D), ld, kept.ctypes.data_as(_lib.p_i64), C.byref(nk),
                                                _stream()))
    A = D.cpu().numpy()[:ld * nd].reshape(nd, ld).T[:nk.value]
    return Matrix(np.asfortranarray(A))


def EITSourceMat(geom, electrodes, integrator=None):
    """EITSourceMat(geo, electrodes, Integrator(3,0,0.005)) --
    assembleSourceMat.h:31, assembleSourceMat.cpp:109-152: n x electrodes;
    electrodes with a radius (a 4-column sensor file) inject a current
    density over their injection triangles, the others an intensity."""
    integ = integrator or Integrator(3, 0, 0.005)
    if isinstance(electrodes, Sensors):
        pos, rad = electrodes.positions, electrodes.radii
    else:
        pos = _electrodes_array(electrodes)
        rad = np.zeros(len(pos))
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rad = np.ascontiguousarray(rad, dtype=np.float64)
    n, ne = geom.size, len(pos)
    D = _dev_f64(n * max(1, ne))
    dp = _to_dev(pos)          # held until the call has been enqueued
    check(lib().omh_eit_source_mat(geom._h, _ptr(dp), rad.ctypes.data_as(_lib.p_f64), ne, integ.order, _ptr(D), n,
                                   ORDER_REFERENCE, _stream()))
    return Matrix(D.cpu().numpy()[:n * ne].reshape(ne, n).T.copy(order="F"))


def GainInternalPot(hminv, dsm, h2ip, ds2ip):
    """GainInternalPot(HeadMatInv, SourceMat, Head2IPMat, Source2IPMat) =
    Source2IP + (Head2IP H^-1) SourceMat -- gain.h:138-144."""
    return _gain_with_inverse(hminv, dsm, h2ip, ds2ip)


def GainEITInternalPot(hminv, sm, h2ip):
    """GainEITInternalPot(HeadMatInv, SourceMat, Head2IPMat) = (Head2IP
    H^-1) SourceMat -- gain.h:146-152."""
    return _gain_with_inverse(hminv, sm, h2ip, None)


def _gain_with_inverse(hminv, src, rows, add):
    n, nv = hminv.nlin(), hminv.nv
    if nv is None:
        raise ValueError("the gain needs the HeadMat unknown layout (a SymMatrix from HeadMat)")
    Hi = hminv.device()
    R = rows.array()
    nr = R.shape[0]
    Rd = to_device_rows(np.asfortranarray(R.T), n, nv)             # (n x nr) triangles-first rows
    RH = _dev_f64(nr * n)
    check(lib().omh_gemm_f64(1, 0, nr, n, n, 1.0, _ptr(Rd), n, _ptr(Hi), n, 0.0, _ptr(RH), nr, _stream()))
    D = src.array()
    Dd = to_device_rows(D, n, nv)
    nd = D.shape[1]
    G = _to_dev(np.asfortranarray(add.array()).ravel(order="F")) if add is not None else _dev_f64(nr * nd)
    check(lib().omh_gemm_f64(0, 0, nr, nd, n, 1.0, _ptr(RH), nr, _ptr(Dd), n, 1.0 if add is not None else 0.0,
                             _ptr(G), nr, _stream()))
    return Matrix(G.cpu().numpy().reshape(nd, nr).T.copy(order="F"))


def Forward(gain, sources, noise_level=0.0, rng=None):
    """Forward(GainMatrix, RealSourcesData, NoiseLevel) -- forward.h:14-36:
    gain x sources plus Gaussian noise of standard deviation noise_level."""
    G = gain.array() if isinstance(gain, Matrix) else np.asarray(gain, dtype=np.float64)
    X = sources.array() if isinstance(sources, Matrix) else np.asarray(sources, dtype=np.float64)
    Y = G @ X
    if noise_level > 0:
        Y = Y + (rng or np.random.default_rng()).normal(0.0, noise_level, Y.shape)
    return Matrix(np.asfortranarray(Y))


def GainMEG(hminv, dsm, h2mm, ds2mm):
    """GainMEG(HeadMatInv, SourceMat, Head2MEGMat, Source2MEGMat) =
    Source2MEG + (Head2MEG H^-1) SourceMat -- gain.h:60-67."""
    n = hminv.nlin()
    nv = hminv.nv
    if nv is None:
        raise ValueError("GainMEG needs the HeadMat unknown layout (a SymMatrix from HeadMat)")
    Hi = hminv.device()
    M = h2mm.array()
    ns = M.shape[0]
    Md = to_device_rows(np.asfortranarray(M.T), n, nv)           # (n x ns) triangles-first rows
    MH = _dev_f64(ns * n)
    check(lib().omh_gemm_f64(1, 0, ns, n, n, 1.0, _ptr(Md), n, _ptr(Hi), n, 0.0, _ptr(MH), ns, _stream()))
    D = dsm.array()
    Dd = to_device_rows(D, n, nv)
    nd = D.shape[1]
    G = _to_dev(np.asfortranarray(ds2mm.array()).ravel(order="F"))
    check(lib().omh_gemm_f64(0, 0, ns, nd, n, 1.0, _ptr(MH), ns, _ptr(Dd), n, 1.0, _ptr(G), ns, _stream()))
    return Matrix(G.cpu().numpy().reshape(nd, ns).T.copy(order="F"))


def GainMEGadjoint(geom, dipoles, hm, h2mm, ds2mm, precision="fp64"):
    """GainMEGadjoint(geo, dipoles, HeadMat, Head2MEGMat, Source2MEGMat) --
    gain.h:90-103: Y = H^-1 M^T with the stored factor, G = Y^T DSM +
    Source2MEG."""
    n, nv = geom.size, geom.n_vertex_unknowns
    M = h2mm.array()
    ns = M.shape[0]
    F = hm._factor(precision)
    Y = to_device_rows(np.asfortranarray(M.T), n, nv)
    F.solve(Y, ns)
    D = _dsm_device(geom, dipoles, "")
    nd = len(_dipoles_array(dipoles))
    G = _to_dev(np.asfortranarray(ds2mm.array()).ravel(order="F"))
    check(lib().omh_gemm_f64(1, 0, ns, nd, n, 1.0, _ptr(Y), n, _ptr(D), n, 1.0, _ptr(G), ns, _stream()))
    return Matrix(G.cpu().numpy().reshape(nd, ns).T.copy(order="F"))


class GainEEGMEGadjoint:
    """GainEEGMEGadjoint(geo, dipoles, HeadMat, Head2EEGMat, Head2MEGMat,
    Source2MEGMat) -- gain.h:105-128: one factor of the HeadMat and one solve
    with the EEG and MEG rows; EEGleadfield / MEGleadfield, saveEEG /
    saveMEG, nlin() = EEG rows + MEG rows."""

    def __init__(self, geom, dipoles, hm, h2em, h2mm, ds2mm, precision="fp64"):
        n, nv = geom.size, geom.n_vertex_unknowns
        ne = h2em.nlin()
        M = h2mm.array()
        ns = M.shape[0]
        torch = _t()
        F = hm._factor(precision)
        Y = torch.cat([_eeg_rhs(geom, h2em).view(ne, n),
                       to_device_rows(np.asfortranarray(M.T), n, nv).view(ns, n)], dim=0).contiguous().view(-1)
        F.solve(Y, ne + ns)
        D = _dsm_device(geom, dipoles, "")
        nd = len(_dipoles_array(dipoles))
        Ge = _gemm_tn(Y[:ne * n], D, n, ne, nd)
        Gm = _to_dev(np.asfortranarray(ds2mm.array()).ravel(order="F"))
        check(lib().omh_gemm_f64(1, 0, ns, nd, n, 1.0, _ptr(Y[ne * n:]), n, _ptr(D), n, 1.0, _ptr(Gm), ns,
                                 _stream()))
        self.EEGleadfield = Matrix(gain_to_numpy(Ge, ne, nd))
        self.MEGleadfield = Matrix(gain_to_numpy(Gm, ns, nd))

    def saveEEG(self, fname):
        self.EEGleadfield.save(fname)

    def saveMEG(self, fname):
        self.MEGleadfield.save(fname)

    def nlin(self):
        return self.EEGleadfield.nlin() + self.MEGleadfield.nlin()


# ---------------------------------------------------------------------------
# Logging (openmeeg/_utils.py): this library writes no log; the level is kept
# so that code that sets and reads it runs unchanged.
# ---------------------------------------------------------------------------
_log_level = "warning"


def set_log_level(level):
    global _log_level
    _log_level = str(level).lower()


def get_log_level():
    return _log_level


class use_log_level:
    """Context manager of the reference's log level (openmeeg._utils)."""

    def __init__(self, level):
        self.level, self.old = level, None

    def __enter__(self):
        self.old = get_log_level()
        set_log_level(self.level)

    def __exit__(self, *exc):
        set_log_level(self.old)
        return False


# ---------------------------------------------------------------------------
# Device pipeline (the path the benchmark times)
# ---------------------------------------------------------------------------
def gain_eeg_pipeline(geom, dipoles, electrodes, precision="fp64", domain_name="", integrator=None,
                      dsm_integrator=None, max_refine=10, refine_tol=1e-8, far_field=None):
    """HeadMat, DipSourceMat and Head2EEGMat assembly, the HeadMat factor and
    the adjoint gain G = (H^-1 E^T)^T DSM, all on the device.  Returns
    (G as a device tensor n_e x n_dip in column-major storage, info dict).
    Mixed precision refines until the estimated relative error of H^-1 E^T
    is below refine_tol (default 1e-8, a hundredth of the 1e-6 the mixed
    path must meet against FP64) and, unless far_field is False, assembles
    well-separated pairs by Gauss quadrature (far_field defaults to True for
    the mixed paths, False for 'fp64')."""
    if far_field is None:
        far_field = precision not in ("fp64", "fp64_lu")
    torch = _t()
    n = geom.size
    times = {}
    ev = lambda: torch.cuda.Event(enable_timing=True)
    marks = [("start", ev())]
    marks[-1][1].record()
    H = _head_mat_device(geom, integrator, far_field)
    marks.append(("headmat", ev())); marks[-1][1].record()
    # the source matrix on a side stream, beside the factor and the solve (its
    # FP64 VALU work fills the CUs their latency-bound launches leave idle);
    # joined before the gain product
    main = torch.cuda.current_stream()
    side = _side_stream()
    side.wait_stream(main)
    with torch.cuda.stream(side):
        d0, d1 = ev(), ev()
        d0.record()
        D = _dsm_device(geom, dipoles, domain_name, dsm_integrator)
        d1.record()
    h2em = Head2EEGMat(geom, electrodes)
    Y = _eeg_rhs(geom, h2em)
    marks.append(("h2em", ev())); marks[-1][1].record()
    F = _head_factor(geom, H, precision, integrator, far_field)
    marks.append(("factor", ev())); marks[-1][1].record()
    F.solve(Y, h2em.nlin(), max_iter=max_refine, tol=refine_tol)
    marks.append(("solve", ev())); marks[-1][1].record()
    main.wait_stream(side)
    D.record_stream(main)
    nd = len(_dipoles_array(dipoles))
    G = _gemm_tn(Y, D, n, h2em.nlin(), nd)
    marks.append(("gain", ev())); marks[-1][1].record()
    torch.cuda.synchronize()
    for (a, ea), (b, eb) in zip(marks[:-1], marks[1:]):
        times[b] = ea.elapsed_time(eb) / 1e3
    times["dsm"] = d0.elapsed_time(d1) / 1e3   # on the side stream, overlapped
    info = dict(times=times, refine_iterations=F.iterations, refine_rel_update=F.rel_update,
                refine_fell_back_fp64=F.fell_back, pivoted_lu=F.pivoted, n=n,
                n_electrodes=h2em.nlin(), n_dipoles=nd, dsm_overlapped=True)
    del F, H, D, Y
    return G, info


def gain_eeg_meg_pipeline(geom, dipoles, electrodes, squids, precision="fp64", domain_name="", integrator=None,
                          dsm_integrator=None, max_refine=10, refine_tol=1e-8, far_field=None):
    """GainEEGMEGadjoint (gain.h:105-128) on the device: one factor of the
    HeadMat, one solve with [E^T | M^T] (the Head2EEGMat and Head2MEGMat
    rows), then G_EEG = Y_E^T DSM and G_MEG = Y_M^T DSM + DipSource2MEG.
    Returns (G_EEG, G_MEG as device tensors, column-major, info)."""
    if far_field is None:
        far_field = precision not in ("fp64", "fp64_lu")
    torch = _t()
    n = geom.size
    ev = lambda: torch.cuda.Event(enable_timing=True)
    marks = [("start", ev())]
    marks[-1][1].record()
    H = _head_mat_device(geom, integrator, far_field)
    marks.append(("headmat", ev())); marks[-1][1].record()
    D = _dsm_device(geom, dipoles, domain_name, dsm_integrator)
    marks.append(("dsm", ev())); marks[-1][1].record()
    h2em = Head2EEGMat(geom, electrodes)
    ne, ns = h2em.nlin(), squids.getNumberOfSensors()
    M = _head2meg_device(geom, squids)                      # ns x n, triangles-first columns
    nd = len(_dipoles_array(dipoles))
    S2 = _ds2meg_device(dipoles, squids)                     # ns x nd
    Y = torch.cat([_eeg_rhs(geom, h2em).view(ne, n), M[:ns * n].view(n, ns).t()], dim=0).contiguous().view(-1)
    marks.append(("sensors", ev())); marks[-1][1].record()
    F = _head_factor(geom, H, precision, integrator, far_field)
    marks.append(("factor", ev())); marks[-1][1].record()
    F.solve(Y, ne + ns, max_iter=max_refine, tol=refine_tol)
    marks.append(("solve", ev())); marks[-1][1].record()
    Ge = _gemm_tn(Y[:ne * n], D, n, ne, nd)
    Gm = S2[:ns * nd].clone()
    check(lib().omh_gemm_f64(1, 0, ns, nd, n, 1.0, _ptr(Y[ne * n:]), n, _ptr(D), n, 1.0, _ptr(Gm), ns, _stream()))
    marks.append(("gain", ev())); marks[-1][1].record()
    torch.cuda.synchronize()
    times = {b: ea.elapsed_time(eb) / 1e3 for (a, ea), (b, eb) in zip(marks[:-1], marks[1:])}
    info = dict(times=times, refine_iterations=F.iterations, refine_fell_back_fp64=F.fell_back, n=n,
                n_electrodes=ne, n_squids=ns, n_dipoles=nd)
    del F, H, D, Y
    return Ge, Gm, info


def gain_eeg_surf_pipeline(geom, mesh, electrodes, squids=None, precision="fp64", integrator=None, max_refine=10,
                           refine_tol=1e-8, far_field=None):
    """The surface-source gains on the device: SurfSourceMat of the source
    mesh, the HeadMat factor, one solve with the Head2EEGMat rows (and the
    Head2MEGMat rows when squids are given), then G_EEG = Y_E^T SSM
    (= GainEEG(H^-1, SSM, H2EM), gain.h:69-76) and G_MEG = Y_M^T SSM +
    SurfSource2MEGMat (= GainMEG, gain.h:60-67).  Returns (G_EEG, G_MEG or
    None, info); gains are device tensors, column-major (sensors x source
    vertices)."""
    if far_field is None:
        far_field = precision not in ("fp64", "fp64_lu")
    if isinstance(mesh, (str, os.PathLike)):
        mesh = Mesh(mesh)
    torch = _t()
    n, nsrc = geom.size, mesh.nb_vertices()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    marks = [("start", ev())]
    marks[-1][1].record()
    H = _head_mat_device(geom, integrator, far_field)
    marks.append(("headmat", ev())); marks[-1][1].record()
    D = _ssm_device(geom, mesh, integrator, ORDER_TRIANGLES_FIRST, far_field)
    marks.append(("ssm", ev())); marks[-1][1].record()
    h2em = Head2EEGMat(geom, electrodes)
    ne = h2em.nlin()
    Y = _eeg_rhs(geom, h2em)
    ns = 0
    if squids is not None:
        ns = squids.getNumberOfSensors()
        M = _head2meg_device(geom, squids)                   # ns x n, triangles-first columns
        S2 = _ss2meg_device(mesh, squids)                    # ns x nsrc
        Y = torch.cat([Y.view(ne, n), M[:ns * n].view(n, ns).t()], dim=0).contiguous().view(-1)
    marks.append(("sensors", ev())); marks[-1][1].record()
    F = _head_factor(geom, H, precision, integrator, far_field)
    marks.append(("factor", ev())); marks[-1][1].record()
    F.solve(Y, ne + ns, max_iter=max_refine, tol=refine_tol)
    marks.append(("solve", ev())); marks[-1][1].record()
    Ge = _gemm_tn(Y[:ne * n], D, n, ne, nsrc)
    Gm = None
    if ns:
        Gm = S2[:ns * nsrc].clone()
        check(lib().omh_gemm_f64(1, 0, ns, nsrc, n, 1.0, _ptr(Y[ne * n:]), n, _ptr(D), n, 1.0, _ptr(Gm), ns,
                                 _stream()))
    marks.append(("gain", ev())); marks[-1][1].record()
    torch.cuda.synchronize()
    times = {b: ea.elapsed_time(eb) / 1e3 for (a, ea), (b, eb) in zip(marks[:-1], marks[1:])}
    info = dict(times=times, refine_iterations=F.iterations, n=n, n_electrodes=ne, n_squids=ns,
                n_sources=nsrc)
    del F, H, D, Y
    return Ge, Gm, info


def gain_internal_pot_pipeline(geom, dipoles, points, precision="fp64", domain_name="", integrator=None,
                               dsm_integrator=None, max_refine=10, refine_tol=1e-8, far_field=None):
    """GainInternalPot on the device in the adjoint form: one factor of the
    HeadMat, one solve with the Surf2VolMat rows, then G = Y^T DSM +
    DipSource2InternalPotMat.  Returns (G as a device tensor, column-major
    kept points x dipoles, kept point indices, info)."""
    if far_field is None:
        far_field = precision not in ("fp64", "fp64_lu")
    torch = _t()
    n = geom.size
    pts = _points_array(points)
    H = _head_mat_device(geom, integrator, far_field)
    D = _dsm_device(geom, dipoles, domain_name, dsm_integrator)
    M, kept = _surf2vol_device(geom, pts)
    ld, nk = max(1, len(pts)), len(kept)
    nd = len(_dipoles_array(dipoles))
    S2 = DipSource2InternalPotMat(geom, dipoles, pts, domain_name).array()
    # M is (points x n, leading dimension ld): its first nk rows, transposed,
    # are the right-hand sides (n x nk column-major, triangles-first rows)
    Y = M[:ld * n].view(n, ld)[:, :nk].t().contiguous().view(-1)
    F = _head_factor(geom, H, precision, integrator, far_field)
    F.solve(Y, nk, max_iter=max_refine, tol=refine_tol)
    G = _to_dev(np.asfortranarray(S2).ravel(order="F"))
    check(lib().omh_gemm_f64(1, 0, nk, nd, n, 1.0, _ptr(Y), n, _ptr(D), n, 1.0, _ptr(G), nk, _stream()))
    torch.cuda.synchronize()
    info = dict(refine_iterations=F.iterations, n=n, n_points=nk, n_dipoles=nd)
    del F, H, D, Y
    return G, kept, info


def gain_to_numpy(G, ne, nd):
    return G.cpu().numpy().reshape(nd, ne).T.copy(order="F")


# ---------------------------------------------------------------------------
# File IO: TrivialBinIO (.bin, TrivialBinIO.H:107-137) and ascii (.txt)
# ---------------------------------------------------------------------------
def save_matrix(fname, a):
    """Matrix::save: .mat (MAT 7.3), .bin or .txt by suffix (MathsIO.C:71-81)."""
    from . import io as omio
    omio.write_full(fname, np.asarray(a, dtype=np.float64))


def load_matrix(fname):
    """Matrix(filename): the format identified from the content."""
    from . import io as omio
    kind, a = omio.read(fname)
    if kind != "full":
        raise ValueError("%s holds a %s matrix, not a full one" % (fname, kind))
    return a if a.ndim == 2 else a.reshape(-1, 1).copy(order="F")


def save_symmatrix(fname, packed):
    from . import io as omio
    omio.write_symmetric(fname, packed)


def load_symmatrix(fname):
    from . import io as omio
    kind, p = omio.read(fname)
    if kind != "symmetric":
        raise ValueError("%s holds a %s matrix, not a symmetric one" % (fname, kind))
    return p
