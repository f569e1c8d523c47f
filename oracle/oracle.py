"""CPU oracle drivers (TEST INFRASTRUCTURE ONLY).

ctypes bindings to oracle/liboracle.so (the C restatement of the integration
kernels, om_oracle.c) plus the drivers that restate, in the reference's
order, HeadMat (assembleHeadMat.cpp:74-135), DipSourceMat
(assembleSourceMat.cpp:66-104), Head2EEGMat (assembleSensors.cpp:25-39,
danielsson.cpp:107-129), Head2MEGMat / DipSource2MEGMat
(assembleSensors.cpp:69-156) and the gains (gain.h:52-89).  Dense linear algebra
(the reference's LAPACK DSPTRF/DSPTRI/DSPTRS, symmatrix.h:259-297,
symmatrix.cpp:133-152) is restated with numpy/scipy LAPACK (Bunch-Kaufman
dsytrf/dsytrs, the unpacked storage form of the same factorisation).

Only tests/, the smoke() check and bench.py's cpu_baseline leg may import
this module; the product path never does.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

from .geometry import Geometry, K, read_tri

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i64 = C.c_int64
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int64)


class _Geom(C.Structure):
    _fields_ = [("nv", i64), ("verts", dp), ("nt", i64), ("tris", ip),
                ("tri_index", ip), ("vert_index", ip), ("nm", i64),
                ("tri_off", ip), ("mv_off", ip), ("mv_ids", ip),
                ("adj_off", ip), ("adj_tri", ip), ("barrier", ip)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_mesh_pair.argtypes = [C.POINTER(_Geom), i64, i64, C.c_int, C.c_uint, C.c_double, C.c_double,
                                    C.c_double, C.c_double, dp]
        L.orc_mesh_pair.restype = C.c_int
        L.orc_deflate_mesh.argtypes = [C.POINTER(_Geom), i64, C.c_double, dp]
        L.orc_packed_get.argtypes = [dp, i64, i64]
        L.orc_packed_get.restype = C.c_double
        L.orc_dsm_mesh.argtypes = [C.POINTER(_Geom), i64, dp, C.c_int, C.c_uint, C.c_double,
                                   C.c_double, C.c_double, C.c_int, dp]
        L.orc_mesh_solid_angle.argtypes = [C.POINTER(_Geom), i64, dp]
        L.orc_mesh_solid_angle.restype = C.c_double
        L.orc_dist_point_mesh.argtypes = [C.POINTER(_Geom), i64, dp, ip, dp]
        L.orc_dist_point_mesh.restype = C.c_double
        L.orc_operator_s.argtypes = [dp, dp]
        L.orc_operator_s.restype = C.c_double
        L.orc_p1slp.argtypes = [dp, dp, dp]
        L.orc_solid_angle.argtypes = [dp, dp]
        L.orc_solid_angle.restype = C.c_double
        L.orc_integrate_d_pair.argtypes = [dp, dp, C.c_int, dp]
        L.orc_surf2vol_mesh.argtypes = [C.POINTER(_Geom), i64, dp, C.c_double, C.c_double, C.c_int, dp]
        L.orc_integrate_pair_adaptive.argtypes = [dp, dp, C.c_int, C.c_int, C.c_uint, C.c_double, dp]
        L.orc_integrate_s_pair.argtypes = [dp, dp, C.c_int]
        L.orc_integrate_s_pair.restype = C.c_double
        L.orc_sample_rows.argtypes = [C.POINTER(_Geom), i64, i64, C.c_int, i64, i64, C.c_int, dp]
        L.orc_sample_rows.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a, t=dp):
    return a.ctypes.data_as(t)


class _Tables:
    """Flattened geometry handed to the C kernels."""

    def __init__(self, g):
        self.g = g
        self.verts = np.ascontiguousarray(np.array(g.vertices, dtype=np.float64).reshape(-1, 3))
        self.tris = np.ascontiguousarray(np.concatenate([m.tris for m in g.meshes]).astype(np.int64))
        self.tri_off = np.cumsum([0] + [len(m.tris) for m in g.meshes]).astype(np.int64)
        self.tri_index = np.concatenate([m.tri_index for m in g.meshes]).astype(np.int64)
        self.vert_index = np.ascontiguousarray(g.vert_index, dtype=np.int64)
        self.mv_off = np.cumsum([0] + [len(m.verts) for m in g.meshes]).astype(np.int64)
        self.mv_ids = np.concatenate([np.asarray(m.verts, dtype=np.int64) for m in g.meshes])
        adj_off = [0]
        adj = []
        for mi, m in enumerate(g.meshes):
            t0 = self.tri_off[mi]
            lists = {}
            for k, t in enumerate(m.tris):          # mesh.h:252-257
                for v in t:
                    lists.setdefault(int(v), []).append(int(t0 + k))
            for v in m.verts:
                l = lists.get(int(v), [])
                adj.extend(l)
                adj_off.append(adj_off[-1] + len(l))
        self.adj_off = np.asarray(adj_off, dtype=np.int64)
        self.adj_tri = np.asarray(adj, dtype=np.int64)
        self.barrier = np.asarray([int(m.current_barrier) for m in g.meshes], dtype=np.int64)
        self.s = _Geom(len(self.verts), _p(self.verts), len(self.tris), _p(self.tris, ip),
                       _p(self.tri_index, ip), _p(self.vert_index, ip), len(g.meshes),
                       _p(self.tri_off, ip), _p(self.mv_off, ip), _p(self.mv_ids, ip),
                       _p(self.adj_off, ip), _p(self.adj_tri, ip), _p(self.barrier, ip))


class OracleLib:
    """Adapter giving oracle/geometry.py access to the C solid angle."""

    def mesh_solid_angle(self, g, mid, p):
        t = tables(g)
        p = np.ascontiguousarray(p, dtype=np.float64)
        return lib().orc_mesh_solid_angle(C.byref(t.s), mid, _p(p))


def tables(g):
    t = getattr(g, "_tables", None)
    if t is None:
        t = _Tables(g)
        g._tables = t
    return t


def _reset_tables(g):
    g._tables = None


def load_geometry(geom_file, cond_file):
    """Geometry(geom, cond) (geometry.h:67-69)."""
    g = Geometry.from_files(geom_file, cond_file, _LazyLib())
    _reset_tables(g)
    return g


class _LazyLib(OracleLib):
    def mesh_solid_angle(self, g, mid, p):
        # the geometry may still be under construction: build tables for the
        # mesh triangles only (indices are not needed for solid angles)
        m = g.meshes[mid]
        verts = np.array(g.vertices, dtype=np.float64)
        P = verts[m.tris]                       # (T,3,3)
        return _solid_angle_sum(P, np.asarray(p, dtype=np.float64))


def _solid_angle_sum(P, p):
    L = lib()
    s = 0.0
    for tri in P:
        t9 = np.ascontiguousarray(tri.reshape(9))
        s += L.orc_solid_angle(_p(t9), _p(np.ascontiguousarray(p)))
    return s


def nested_geometry(meshes, conductivity):
    """make_nested_geometry (_make_geometry.py:131-211)."""
    g = Geometry.nested(meshes, conductivity, _LazyLib())
    _reset_tables(g)
    return g


# ---------------------------------------------------------------------------
# HeadMat
# ---------------------------------------------------------------------------
def head_mat(g, order=3, levels=0, tol=0.0):
    """Details::HeadMatrix + deflate (assembleHeadMat.cpp:28-60,74-115) with
    Integrator(order, levels, tol) (levels > 0: adaptive pair integrals,
    integrator.h:44-93).  Returns the packed upper SymMatrix data (length
    n(n+1)/2)."""
    t = tables(g)
    n = g.size
    H = np.zeros(n * (n + 1) // 2, dtype=np.float64)
    L = lib()
    for (m1, m2, o, cS, cN, cD) in g.pairs:
        rc = L.orc_mesh_pair(C.byref(t.s), m1, m2, order, levels, tol, cS, cN, cD, _p(H))
        if rc != 0:
            raise MemoryError("oracle mesh pair")
    for part in g.parts:
        nbv = 0
        i_first = 0
        for mid in part:
            m = g.meshes[mid]
            if m.outermost:
                nbv += len(m.verts)
                if i_first == 0:
                    i_first = int(g.vert_index[m.verts[0]])
        coef = L.orc_packed_get(_p(H), i_first, i_first) / nbv
        for mid in part:
            if g.meshes[mid].outermost:
                L.orc_deflate_mesh(C.byref(t.s), mid, coef, _p(H))
    return H


def unpack(Hp, n):
    """Packed upper (symmatrix.h:52-62) -> dense symmetric."""
    iu = np.triu_indices(n)
    # packed index i + j(j+1)/2 for i<=j: order columns j, rows i<=j
    j = np.repeat(np.arange(n), np.arange(1, n + 1))
    i = np.concatenate([np.arange(k + 1) for k in range(n)])
    M = np.zeros((n, n))
    M[i, j] = Hp
    M[j, i] = Hp
    return M


def pack(M):
    n = M.shape[0]
    j = np.repeat(np.arange(n), np.arange(1, n + 1))
    i = np.concatenate([np.arange(k + 1) for k in range(n)])
    return M[i, j].copy()


# ---------------------------------------------------------------------------
# DipSourceMat
# ---------------------------------------------------------------------------
def dip_source_mat(g, dipoles, domain_name="", order=3, levels=10, tol=0.001):
    """SourceMatrix<Dipole> (assembleSourceMat.cpp:66-104) with the default
    adaptive Integrator(3,10,0.001) of assembleSourceMat.h:27-29."""
    t = tables(g)
    L = lib()
    dipoles = np.asarray(dipoles, dtype=np.float64)
    if dipoles.ndim != 2 or dipoles.shape[1] != 6:
        raise ValueError("dipoles must have 6 columns")
    n = g.size
    out = np.zeros((n, len(dipoles)), dtype=np.float64, order="F")
    ol = OracleLib()
    for s, d6 in enumerate(dipoles):
        dk = g.domain_index(domain_name) if domain_name else g.domain_of_point(d6[:3], ol)
        dom = g.domains[dk]
        if dom.cond == 0.0:
            continue
        col = np.zeros(n)
        d6c = np.ascontiguousarray(d6)
        for iface, inside in dom.boundaries:
            factorD = -K if inside else K
            for mid, o in iface.omeshes:
                coeffD = factorD * o
                m = g.meshes[mid]
                L.orc_dsm_mesh(C.byref(t.s), mid, _p(d6c), order, levels, tol, coeffD,
                               coeffD / dom.cond, int(not m.current_barrier), _p(col))
        out[:, s] = col
    return out


# ---------------------------------------------------------------------------
# Head2EEGMat
# ---------------------------------------------------------------------------
def head2eeg(g, electrodes):
    """Head2EEGMat (assembleSensors.cpp:25-39) with dist_point_geom
    (danielsson.cpp:107-129).  Returns (cols (n_e,3), vals (n_e,3)): the
    three vertex unknowns of the closest triangle and their weights."""
    t = tables(g)
    L = lib()
    E = np.asarray(electrodes, dtype=np.float64)
    cols = np.zeros((len(E), 3), dtype=np.int64)
    vals = np.zeros((len(E), 3))
    for e, p in enumerate(E):
        pc = np.ascontiguousarray(p)
        distmin = np.finfo(np.float64).max
        best_tri = -1
        alphas = np.zeros(3)
        for dom in g.domains:
            if dom.cond != 0.0:
                continue
            for iface, _ in dom.boundaries:
                # dist_point_interface: first strict minimum over its meshes
                dmin_i = np.finfo(np.float64).max
                tri_i = -1
                al_i = np.zeros(3)
                for mid, _ in iface.omeshes:
                    tri = C.c_int64(-1)
                    al = np.zeros(3)
                    dd = L.orc_dist_point_mesh(C.byref(t.s), mid, _p(pc), C.byref(tri), _p(al))
                    if dd < dmin_i:
                        dmin_i, tri_i, al_i = dd, tri.value, al
                # the reference overwrites alphas with every interface's result
                alphas = al_i
                if dmin_i < distmin:
                    distmin, best_tri = dmin_i, tri_i
        vids = t.tris[best_tri]
        cols[e] = g.vert_index[vids]
        vals[e] = alphas
    return cols, vals


def h2em_dense(g, cols, vals):
    E = np.zeros((len(cols), g.size))
    for e in range(len(cols)):
        for k in range(3):
            E[e, cols[e, k]] = vals[e, k]
    return E


# ---------------------------------------------------------------------------
# Gains
# ---------------------------------------------------------------------------
def gain_eeg(H, dsm, E):
    """GainEEG (gain.h:69-76): (E * H^-1) * D, with H^-1 applied through a
    symmetric-indefinite LAPACK solve."""
    import scipy.linalg as sla
    Y = sla.solve(H, E.T, assume_a="sym")
    return Y.T @ dsm


def read_dipoles(path):
    return np.loadtxt(path, ndmin=2)


def read_sensors(path):
    """Sensors::load (sensors.cpp:36-143): optional label column then x y z."""
    rows = []
    with open(path) as f:
        for ln in f:
            tok = ln.split()
            if not tok or tok[0].startswith("#"):
                continue
            rows.append(tok)
    labeled = not any(r[0].count(".") == 1 for r in rows)   # any numeric first token (sensors.cpp:67-69)
    data = [r[1:] if labeled else r for r in rows]
    return np.array(data, dtype=np.float64)[:, :3]


# ---------------------------------------------------------------------------
# MEG: Head2MEGMat, DipSource2MEGMat (assembleSensors.cpp:69-156), GainMEG
# (gain.h:60-67)
# ---------------------------------------------------------------------------
MAG_FACTOR = 1e-7   # constants.h:22
_EDGE = ((1, 2), (2, 0), (0, 1))   # Triangle::indices, triangle.h:154


def read_squids(path):
    """Sensors::load (sensors.cpp:36-143) of MEG integration points
    (label x y z dx dy dz w): positions, orientations, weights and the sensor
    index of each point (labels in order of first appearance)."""
    rows = []
    with open(path) as f:
        for ln in f:
            tok = ln.split()
            if tok and not tok[0].startswith("#"):
                rows.append(tok)
    labeled = not any(r[0].count(".") == 1 for r in rows)   # any numeric first token (sensors.cpp:67-69)
    names = [r[0] for r in rows] if labeled else [str(i) for i in range(len(rows))]
    mat = np.array([r[1:] if labeled else r for r in rows], dtype=np.float64)
    order = {}
    sensor = np.array([order.setdefault(nm, len(order)) for nm in names], dtype=np.int64)
    w = mat[:, 6] if mat.shape[1] == 7 else np.ones(len(mat))
    return mat[:, :3].copy(), mat[:, 3:6].copy(), w.copy(), sensor


def _weights_matrix(weights, sensor):
    W = np.zeros((int(sensor.max()) + 1, len(sensor)))
    W[sensor, np.arange(len(sensor))] = weights
    return W


def conductivity_jump(g, m):
    """Geometry::conductivity_jump (geometry.h:188-194)."""
    return sum(d.cond * d.mesh_orientation(m) for d in g.domains)


def _ferguson(verts, tris, pts, coeff, F):
    """F[i, :, v] += coeff x the Ferguson operator of the triangles tris at
    point i for vertex v (operators.h:40-59: sum over the triangles T around
    v of (B - A) S_T(x) / (2 |T|), A and B the edge opposite v)."""
    L = lib()
    for t in tris:
        T = verts[t]
        area = 0.5 * np.linalg.norm(np.cross(T[1] - T[0], T[2] - T[0]))
        t9 = np.ascontiguousarray(T.reshape(9))
        for i, x in enumerate(pts):
            sv = L.orc_operator_s(_p(t9), _p(np.ascontiguousarray(x))) / (2 * area) * coeff
            for k in range(3):
                F[i, :, t[k]] += (T[_EDGE[k][1]] - T[_EDGE[k][0]]) * sv


def head2meg(g, pts, dirs, weights, sensor):
    """Head2MEGMat (assembleSensors.cpp:69-99): the Ferguson matrix of every
    mesh (assemble_ferguson, assembleFerguson.cpp:18-30, coefficient
    MagFactor x the mesh's conductivity jump; Details::operatorFerguson,
    operators.h:40-59: sum over the triangles T around a vertex of
    (B - A) S_T(x) / (2 |T|), A and B the edge opposite the vertex),
    projected on each point's direction and summed into sensors."""
    pts = np.asarray(pts, dtype=np.float64)
    dirs = np.asarray(dirs, dtype=np.float64)
    verts = np.array(g.vertices, dtype=np.float64)
    F = np.zeros((len(pts), 3, len(verts)))
    for mid, m in enumerate(g.meshes):
        coeff = MAG_FACTOR * conductivity_jump(g, mid)
        if coeff != 0.0:
            _ferguson(verts, m.tris, pts, coeff, F)
    mat = np.zeros((len(pts), g.size))
    dn = np.linalg.norm(dirs, axis=1)
    for v in range(len(verts)):
        if g.vert_index[v] >= 0:
            mat[:, g.vert_index[v]] = (F[:, :, v] * dirs).sum(axis=1) / dn
    return _weights_matrix(weights, sensor) @ mat


def dip_source2meg(dipoles, pts, dirs, weights, sensor):
    """DipSource2MEGMat (assembleSensors.cpp:133-156): MagFactor (q x (x - r))
    . d / (|x - r|^3 |d|), summed into sensors."""
    dip = np.asarray(dipoles, dtype=np.float64)
    pts = np.asarray(pts, dtype=np.float64)
    dirs = np.asarray(dirs, dtype=np.float64)
    diff = pts[:, None, :] - dip[None, :, :3]
    nd = np.linalg.norm(diff, axis=2)
    f = np.cross(np.broadcast_to(dip[None, :, 3:], diff.shape), diff) / (nd ** 3)[..., None]
    mat = (f * dirs[:, None, :]).sum(axis=2) * MAG_FACTOR / np.linalg.norm(dirs, axis=1)[:, None]
    return _weights_matrix(weights, sensor) @ mat


def gain_meg(H, dsm, h2mm, ds2mm):
    """GainMEG (gain.h:60-67): Source2MEG + (Head2MEG H^-1) SourceMat."""
    import scipy.linalg as sla
    Y = sla.solve(H, h2mm.T, assume_a="sym")
    return ds2mm + Y.T @ dsm


def surf_source2meg(src_pts, src_tris, pts, dirs, weights, sensor):
    """SurfSource2MEGMat (assembleSensors.cpp:102-124): the Ferguson operator
    of the source mesh with coefficient 1 (no MagFactor) at each point,
    projected on the point's direction, one column per source-mesh vertex
    (local index, Mesh::generate_indices mesh.cpp:266-272), summed into
    sensors."""
    verts = np.asarray(src_pts, dtype=np.float64)
    pts = np.asarray(pts, dtype=np.float64)
    dirs = np.asarray(dirs, dtype=np.float64)
    F = np.zeros((len(pts), 3, len(verts)))
    _ferguson(verts, np.asarray(src_tris, dtype=np.int64), pts, 1.0, F)
    mat = (F * dirs[:, :, None]).sum(axis=1) / np.linalg.norm(dirs, axis=1)[:, None]
    return _weights_matrix(weights, sensor) @ mat


# ---------------------------------------------------------------------------
# SurfSourceMat
# ---------------------------------------------------------------------------
class _SourceTables:
    """The head's tables with the source mesh appended as one more mesh
    (a current barrier, so that only its N and D blocks against a head mesh
    are formed).  Its vertices take the unknown indices n .. n + nsrc - 1, so
    that the rectangular block (head unknown, source vertex) lies in the
    upper half of an (n + nsrc) packed matrix."""

    def __init__(self, g, src_pts, src_tris):
        from .geometry import Mesh
        nvh, n = len(g.vertices), g.size
        ns = len(src_pts)
        sm = Mesh("source")
        sm.verts = list(range(nvh, nvh + ns))
        sm.tris = np.asarray(src_tris, dtype=np.int64) + nvh
        sm.current_barrier = True
        sm.tri_index = np.arange(len(sm.tris), dtype=np.int64) + n + ns   # never written
        self.vertices = list(g.vertices) + [tuple(float(c) for c in p) for p in src_pts]
        self.meshes = list(g.meshes) + [sm]
        self.vert_index = np.r_[np.asarray(g.vert_index, dtype=np.int64), n + np.arange(ns, dtype=np.int64)]


def surf_source_mat(g, src_pts, src_tris, order=3, stats=None):
    """SurfSourceMat (assembleSourceMat.cpp:22-56) with the non-adaptive
    Integrator(order, 0, .): the source mesh lies in the domain of its first
    vertex (conductivity sigma); for each boundary (interface, inside) of that
    domain and each oriented mesh of the interface, coeffN = (inside ? K :
    -K) x orientation,
      N block (head vertices x source vertices) with coefficient coeffN, from
        an S block of coefficient 1 (operators.h:372, 428-437),
      D block (head triangles x source vertices) with coefficient
        -coeffN / sigma (operators.h:439-443).
    The D block of a current-barrier head mesh is left out (its triangles
    have no rows in the n x nsrc matrix; the reference would index past the
    matrix there).  Returns the n x nsrc matrix (reference unknown order);
    stats (a dict), when given, receives the seconds spent in the C pair
    assembly ("pair_seconds"), without the Python table set-up."""
    import time
    src_pts = np.asarray(src_pts, dtype=np.float64)
    src_tris = np.asarray(src_tris, dtype=np.int64)
    n, ns = g.size, len(src_pts)
    ext = _SourceTables(g, src_pts, src_tris)
    t = _Tables(ext)
    dom = g.domains[g.domain_of_point(src_pts[0], OracleLib())]
    Lc = -1.0 / dom.cond
    M = len(g.meshes)
    N = n + ns
    H = np.zeros(N * (N + 1) // 2)
    L = lib()
    t0 = time.perf_counter()
    for iface, inside in dom.boundaries:
        factorN = K if inside else -K
        for mid, o in iface.omeshes:
            coeffN = factorN * o
            # the source mesh is a barrier: S goes to a temporary, N with
            # coeffN, D(head mesh rows, source columns) unless the head mesh
            # is a barrier, no D* (orc_mesh_pair's branches)
            if L.orc_mesh_pair(C.byref(t.s), mid, M, order, 0, 0.0, 0.0, coeffN, coeffN * Lc, _p(H)) != 0:
                raise MemoryError("oracle mesh pair")
    if stats is not None:
        stats["pair_seconds"] = time.perf_counter() - t0
    out = np.zeros((n, ns))
    for c in range(ns):
        j = n + c
        out[:, c] = H[np.arange(n) + j * (j + 1) // 2]
    return out


# ---------------------------------------------------------------------------
# Internal potentials
# ---------------------------------------------------------------------------
def _conducting_points(g, points):
    """Points in a domain of non-zero conductivity, in input order, and their
    domains (the others are dropped, assembleHeadMat.cpp:331-340)."""
    ol = OracleLib()
    kept, doms = [], []
    for i, p in enumerate(np.asarray(points, dtype=np.float64)):
        d = g.domain_of_point(p, ol)
        if g.domains[d].cond != 0.0:
            kept.append(i)
            doms.append(d)
    return np.asarray(kept, dtype=np.int64), doms


def surf2vol(g, points):
    """Surf2VolMat (assembleHeadMat.cpp:325-363) with PartialBlock
    (operators.h:302-331): for a point x of domain Omega and each mesh of
    Omega's boundaries, coeff = K x the boundary's orientation of the mesh,
      row(x) at the mesh's vertices += -coeff x P1SingleLayerPotential_T(x)
        (summed over the triangles T of the vertex),
      row(x) at the mesh's triangles  = coeff / sigma_Omega x OperatorS_T(x)
        (not for a current-barrier mesh).
    Returns (matrix of the kept points x n, indices of the kept points)."""
    L = lib()
    t = tables(g)
    pts = np.asarray(points, dtype=np.float64)
    kept, doms = _conducting_points(g, pts)
    out = np.zeros((len(kept), g.size))
    for row, (i, d) in enumerate(zip(kept, doms)):
        dom = g.domains[d]
        x = np.ascontiguousarray(pts[i])
        orow = out[row]
        for iface, inside in dom.boundaries:
            for mid, o in iface.omeshes:
                m = g.meshes[mid]
                coeff = (o if inside else -o) * K
                L.orc_surf2vol_mesh(C.byref(t.s), mid, _p(x), -coeff, coeff / dom.cond, int(not m.current_barrier),
                                    _p(orow))
    return out, kept


def dip_source2internal_pot(g, dipoles, points, domain_name=""):
    """DipSource2InternalPotMat (assembleSourceMat.cpp:154-186): K / sigma x
    the dipole's potential q . (x - r) / |x - r|^3 at the kept points of the
    dipole's domain, zero elsewhere."""
    ol = OracleLib()
    pts = np.asarray(points, dtype=np.float64)
    kept, doms = _conducting_points(g, pts)
    dip = np.asarray(dipoles, dtype=np.float64)
    out = np.zeros((len(kept), len(dip)))
    for j, d6 in enumerate(dip):
        dk = g.domain_index(domain_name) if domain_name else g.domain_of_point(d6[:3], ol)
        coeff = K / g.domains[dk].cond
        for row, (i, d) in enumerate(zip(kept, doms)):
            if d == dk:
                diff = pts[i] - d6[:3]
                out[row, j] += coeff * np.dot(d6[3:], diff) / np.linalg.norm(diff) ** 3
    return out


def _closest_outer(g, p):
    """dist_point_geom (danielsson.cpp:107-129): (global triangle, interface)
    closest to p on the interfaces of zero-conductivity domains."""
    t = tables(g)
    L = lib()
    pc = np.ascontiguousarray(p, dtype=np.float64)
    distmin, best, best_if = np.finfo(np.float64).max, -1, None
    for dom in g.domains:
        if dom.cond != 0.0:
            continue
        for iface, _ in dom.boundaries:
            dmin_i, tri_i = np.finfo(np.float64).max, -1
            for mid, _ in iface.omeshes:
                tri = C.c_int64(-1)
                al = np.zeros(3)
                dd = L.orc_dist_point_mesh(C.byref(t.s), mid, _p(pc), C.byref(tri), _p(al))
                if dd < dmin_i:
                    dmin_i, tri_i = dd, tri.value
            if dmin_i < distmin:
                distmin, best, best_if = dmin_i, tri_i, iface
    return best, best_if


def eit_source_mat(g, positions, radii=None, order=3):
    """EITSourceMat (assembleSourceMat.cpp:109-152) with the injection
    triangles of Sensors::findInjectionTriangles (sensors.cpp:167-220):
    column e = sum over electrode e's injection triangles t of w x row t of
    the transfer matrix, built from every mesh pair (m1, m2) whose first
    mesh m1 (t's mesh) is a current barrier:
      D block with coefficient K x orientation (rows m1's triangles, columns
        m2's vertices),
      m1 == m2: the P0-P1 identity with -1/2 (|t|/3 at t's vertices),
      m1 != m2: S block with -K x orientation x sigma_inv(m1, m2).
    w = 1/|t| for a zero radius (an intensity), else pi r^2 over the area of
    the electrode's injection triangles."""
    t = tables(g)
    L = lib()
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    rad = np.zeros(len(pos)) if radii is None else np.asarray(radii, dtype=np.float64)
    verts = np.array(g.vertices, dtype=np.float64)
    tmesh = np.concatenate([np.full(len(m.tris), i) for i, m in enumerate(g.meshes)])
    tri9 = lambda k: np.ascontiguousarray(verts[t.tris[k]].reshape(9))
    area = lambda k: 0.5 * np.linalg.norm(np.cross(*(verts[t.tris[k][1:]] - verts[t.tris[k][0]])))
    n = g.size
    out = np.zeros((n, len(pos)))
    r3 = np.zeros(3)
    for e, p in enumerate(pos):
        t0, iface = _closest_outer(g, p)
        tris = [t0]
        if rad[e] != 0.0:
            if area(t0) < 4 * math.pi * rad[e] ** 2:
                edges = {}
                for mid, _ in iface.omeshes:
                    for k in range(t.tri_off[mid], t.tri_off[mid + 1]):
                        for a in range(3):
                            ed = tuple(sorted((t.tris[k][a], t.tris[k][(a + 1) % 3])))
                            edges.setdefault(ed, []).append(k)
                seen, stack = {t0}, [t0]
                while stack:
                    k = stack.pop()
                    if np.linalg.norm(verts[t.tris[k]].mean(axis=0) - p) < rad[e]:
                        if k != t0:
                            tris.append(k)
                        for a in range(3):
                            for u in edges[tuple(sorted((t.tris[k][a], t.tris[k][(a + 1) % 3])))]:
                                if u != k and u not in seen:
                                    seen.add(u)
                                    stack.append(u)
            w_e = math.pi * rad[e] ** 2 / sum(area(k) for k in tris)
        for k in tris:
            m1 = int(tmesh[k])
            if not g.meshes[m1].current_barrier:
                continue
            w = 1.0 / area(k) if rad[e] == 0.0 else w_e
            row = np.zeros(g.nb_parameters)
            for (pm1, m2, o, cS, cN, cD) in g.pairs:
                if pm1 != m1:
                    continue
                for j in range(t.tri_off[m2], t.tri_off[m2 + 1]):
                    L.orc_integrate_d_pair(_p(tri9(k)), _p(tri9(j)), order, _p(r3))
                    for a in range(3):
                        row[g.vert_index[t.tris[j][a]]] += K * o * r3[a]
                if m2 == m1:
                    for a in range(3):
                        row[g.vert_index[t.tris[k][a]]] += -0.5 * area(k) / 3.0
                else:
                    for j in range(t.tri_off[m2], t.tri_off[m2 + 1]):
                        row[t.tri_index[j]] = -K * o * g.sigma_inv(m1, m2) * \
                            L.orc_integrate_s_pair(_p(tri9(k)), _p(tri9(j)), order)
            out[:, e] += row[:n] * w
    return out


def gain_internal_pot(H, dsm, h2ip, ds2ip):
    """GainInternalPot (gain.h:138-144): Source2IP + Head2IP H^-1 SourceMat."""
    import scipy.linalg as sla
    return ds2ip + h2ip @ sla.solve(H, dsm, assume_a="sym")


# ---------------------------------------------------------------------------
# CPU baseline (bench.py's cpu_baseline leg)
# ---------------------------------------------------------------------------

def pair_counts(g):
    """(S pairs, D pairs) of the HeadMat of geometry g: every communicating
    mesh pair's S block (upper triangle of a diagonal block) and its D / D*
    blocks (none for a current-barrier row mesh); assembleHeadMat.cpp:83-108
    with the block rules of operators.h:180-197,363-384."""
    n_s = n_d = 0.0
    for (m1, m2, o, cS, cN, cD) in g.pairs:
        n1, n2 = len(g.meshes[m1].tris), len(g.meshes[m2].tris)
        b1, b2 = g.meshes[m1].current_barrier, g.meshes[m2].current_barrier
        n_s += n1 * (n1 + 1) / 2.0 if m1 == m2 else n1 * n2
        if not b1:
            n_d += n1 * n2
        if m1 != m2 and not b2:
            n_d += n1 * n2
    return n_s, n_d


def time_pipeline(meshes, cond, dip, el, domain_name="Brain"):
    """Runs the oracle's whole EEG gain pipeline and times each stage (wall
    clock, all threads the process has): HeadMat (C, OpenMP), DipSourceMat
    (adaptive Integrator(3,10,0.001)), Head2EEGMat, the factorisation of the
    full symmetric matrix (LAPACK dsytrf, Bunch-Kaufman), the solve with the
    electrode rows as right-hand sides (dsytrs; GainEEGadjoint's linsolve,
    gain.h:78-89) and the gain product.  Returns (gain, stage seconds, g)."""
    import time
    import scipy.linalg.lapack as la
    st = {}
    t0 = time.time()
    g = nested_geometry(meshes, cond)
    Hp = head_mat(g)
    st["headmat"] = time.time() - t0
    H = unpack(Hp, g.size)
    del Hp
    t0 = time.time()
    D = dip_source_mat(g, dip, domain_name)
    st["dsm"] = time.time() - t0
    t0 = time.time()
    cols, vals = head2eeg(g, el)
    E = h2em_dense(g, cols, vals)
    st["h2em"] = time.time() - t0
    t0 = time.time()
    lu, piv, info = la.dsytrf(H, lower=0, overwrite_a=1)
    st["factor"] = time.time() - t0
    if info != 0:
        raise RuntimeError("dsytrf info %d" % info)
    t0 = time.time()
    Y, info = la.dsytrs(lu, piv, np.asfortranarray(E.T), lower=0)
    st["solve"] = time.time() - t0
    t0 = time.time()
    G = Y.T @ D
    st["gain"] = time.time() - t0
    return G, st, g


def extrapolate_stages(st, g_from, nd_from, ne_from, g_to, nd_to, ne_to):
    """Scales measured oracle stage times to a larger workload: HeadMat by
    the S + D pair count, DipSourceMat by dipoles x innermost triangles,
    Head2EEGMat by electrodes x scalp triangles, the factorisation by n^3,
    the solve by n^2 x electrodes, the gain by n_e x n x n_dip.  Returns the
    scaled stage times."""
    sf, df = pair_counts(g_from)
    stt, dt_ = pair_counts(g_to)
    # the measured HeadMat time splits over S and D like the pair counts
    asm = st["headmat"] * (stt + dt_) / (sf + df)
    tin_f, tin_t = len(g_from.meshes[0].tris), len(g_to.meshes[0].tris)
    tout_f, tout_t = len(g_from.meshes[-1].tris), len(g_to.meshes[-1].tris)
    n_f, n_t = float(g_from.size), float(g_to.size)
    return dict(headmat=asm, dsm=st["dsm"] * (nd_to * tin_t) / (nd_from * tin_f),
                h2em=st["h2em"] * (ne_to * tout_t) / (ne_from * tout_f),
                factor=st["factor"] * (n_t / n_f) ** 3, solve=st["solve"] * (n_t / n_f) ** 2 * ne_to / ne_from,
                gain=st["gain"] * (ne_to * n_t * nd_to) / (ne_from * n_f * nd_from))
