"""The mesh-by-mesh factor order (Geometry.factor_order) on the host: the
order and the reordered block structure, the pivot signs of a pivot-free
LDL^T of the oracle's HeadMat taken in that order, and the argument checks
of the reordering entry points (no device needed)."""
import ctypes as C

import numpy as np
import pytest

FOUR = ((0.8, 0.84, 0.9, 1.0), (0.33, 1.79, 0.0042, 0.33))
THREE = ((0.87, 0.92, 1.0), (1.0, 0.0125, 1.0))


def _geometry(k, spec):
    import openmeeg_amd as om
    from openmeeg_amd import spheres as S
    meshes = S.nested_sphere_meshes(k, spec[0])
    return om.make_nested_geometry(meshes, spec[1]), meshes


def _kinds(g):
    """(kind, mesh) of each segment of factor_segments()."""
    starts, _ = g.factor_segments()
    nt, nv = g._nt, g._nv
    out = {}
    for m in range(g.n_meshes):
        inf = g.mesh_info(m)
        if inf["ntris"] > 0 and not inf["current_barrier"]:
            out[inf["tri_first"] - nv] = ("T", m)
        if inf["nverts"] > 0:
            out[inf["vert_first"] + nt] = ("V", m)
    return [out[int(a)] for a in starts]


def _update_flops(starts, hfill, n, group=1024):
    """Trailing-update flops of a right-looking factorisation that updates,
    per group of columns, only the rows L's blocks reach (L's pattern by
    symbolic elimination over the segments)."""
    ns = len(starts)
    L = np.tril((hfill != 0) | (hfill != 0).T | np.eye(ns, dtype=bool))
    L = L | L.T
    for k in range(ns):
        nb = [i for i in range(k + 1, ns) if L[i, k]]
        for i in nb:
            for j in nb:
                L[i, j] = True
    ends = np.append(starts[1:], n)
    seg = np.searchsorted(starts, np.arange(n), side="right") - 1
    f = 0.0
    for c0 in range(0, n, group):
        c1 = min(n, c0 + group)
        ss = set(seg[c0:c1].tolist())
        rows = sum(max(0, ends[s] - max(starts[s], c1)) for s in range(ns) if any(L[s, q] for q in ss))
        f += rows * (rows + 1) * (c1 - c0)
    return f


@pytest.mark.parametrize("spec", [FOUR, THREE])
def test_factor_order_is_mesh_by_mesh(spec):
    """Each mesh's triangles, then its vertices, innermost mesh first; the
    reordered pattern is block tridiagonal over the meshes, and its
    block-restricted trailing updates take fewer flops than in the
    triangles-first order (a quarter fewer at 4 layers)."""
    from openmeeg_amd.api import _reorder_structure
    g, _ = _geometry(4, spec)
    n = g.size
    starts, hfill = g.factor_segments()
    order = g.factor_order()
    kinds = _kinds(g)
    nm = len(spec[0])
    expect = []
    for m in range(nm):
        if ("T", m) in kinds:
            expect.append(("T", m))
        expect.append(("V", m))
    assert [kinds[q] for q in order] == expect
    fstarts, fh = _reorder_structure((starts, hfill), order, n)
    lens = np.append(starts[1:], n) - starts
    assert np.array_equal(np.diff(np.append(fstarts, n)), lens[order])
    mesh = [kinds[q][1] for q in order]
    for q in range(len(order)):
        for p in range(q):
            if fh[q, p]:
                assert abs(mesh[q] - mesh[p]) <= 1, (q, p)
    f_tri = _update_flops(starts, hfill, n, group=8)     # groups small against the segments, as at full size
    f_mesh = _update_flops(fstarts, fh, n, group=8)
    assert f_mesh < f_tri
    if nm == 4:
        assert f_mesh < 0.85 * f_tri


def _ldl(A, nb=64):
    """Pivot-free LDL^T (float64, blocked): unit lower L and the pivots d."""
    A = A.copy()
    n = A.shape[0]
    d = np.zeros(n)
    for k0 in range(0, n, nb):
        k1 = min(n, k0 + nb)
        for j in range(k0, k1):
            d[j] = A[j, j]
            l = A[j + 1:k1, j] / d[j]
            A[j + 1:k1, j + 1:k1] -= np.outer(l, A[j + 1:k1, j])
            A[j + 1:k1, j] = l
        if k1 < n:
            L11 = np.tril(A[k0:k1, k0:k1], -1) + np.eye(k1 - k0)
            X = np.linalg.solve(L11, A[k0:k1, k1:]).T          # A21 L11^-T
            L21 = X / d[k0:k1]
            A[k1:, k1:] -= L21 @ X.T
            A[k1:, k0:k1] = L21
    return np.tril(A, -1) + np.eye(n), d


@pytest.mark.parametrize("spec", [FOUR, THREE])
def test_factor_order_pivot_signs_on_oracle_headmat(spec):
    """The oracle's HeadMat (reference order) taken in the factor order has a
    pivot-free LDL^T whose pivots are + on every triangle segment and - on
    every vertex segment (the pattern the device factorisation checks), and
    that factor solves the system to the rounding of the order it replaces."""
    from oracle import oracle as O
    from openmeeg_amd.api import _segment_signs
    g, meshes = _geometry(3, spec)
    n = g.size
    H = O.unpack(O.head_mat(O.nested_geometry(meshes, spec[1])), n)
    p = g.tri_first_index(np.arange(n))
    Htf = np.empty_like(H)
    Htf[np.ix_(p, p)] = H
    starts, hfill = g.factor_segments()
    order = g.factor_order()
    ends = np.append(starts[1:], n)
    idx = np.concatenate([np.arange(starts[s], ends[s]) for s in order])
    Hm = Htf[np.ix_(idx, idx)]
    L, d = _ldl(Hm)
    sg = _segment_signs((starts, hfill), order, n - g.n_vertex_unknowns)
    expect = np.concatenate([np.full(ends[s] - starts[s], sg[q]) for q, s in enumerate(order)])
    assert np.array_equal(np.sign(d), expect)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    y = np.linalg.solve(L, b[idx]) / d
    x = np.empty(n)
    x[idx] = np.linalg.solve(L.T, y)
    ref = np.linalg.solve(Htf, b)
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    # triangles-first gives the same pivots' signs pattern (+ then -)
    _, d_tf = _ldl(Htf)
    assert np.array_equal(np.sign(d_tf), np.where(np.arange(n) < n - g.n_vertex_unknowns, 1.0, -1.0))


def test_reorder_arguments_are_checked():
    """The reordering entry points refuse a segment order that is not a
    permutation, and per-segment signs other than +-1, before touching the
    device."""
    from openmeeg_amd import _lib
    L = _lib.lib()
    dummy = C.c_void_p(16)               # never dereferenced: the check fails first
    st = np.array([0, 10, 20], dtype=np.int64)
    sp = st.ctypes.data_as(C.c_void_p)
    hfill = np.zeros((3, 3), dtype=np.uint8)
    hp = hfill.ctypes.data_as(C.c_void_p)
    for bad in ([0, 1, 1], [0, 1, 3], [2, -1, 0]):
        o = np.array(bad, dtype=np.int32)
        op = o.ctypes.data_as(C.c_void_p)
        assert L.omh_sym_reorder(dummy, 30, C.c_void_p(32), 30, 4, 30, 3, sp, op, C.c_void_p(0)) == 2
        assert b"permutation" in L.omh_last_error()
        assert L.omh_reorder_rows_f64(dummy, 30, C.c_void_p(32), 30, 30, 2, 3, sp, op, 0, C.c_void_p(0)) == 2
        assert b"permutation" in L.omh_last_error()
        it, rel = C.c_int32(), C.c_double()
        assert L.omh_mixed_solve_reordered(dummy, 30, 30, dummy, 30, dummy, dummy, 30, dummy, 30, 4, 5, 1e-8, 3, sp,
                                           hp, op, C.byref(it), C.byref(rel), C.c_void_p(0)) == 2
        assert b"permutation" in L.omh_last_error()
    o = np.array([0, 2, 1], dtype=np.int32)
    assert L.omh_sym_reorder(dummy, 30, dummy, 30, 4, 30, 3, sp, o.ctypes.data_as(C.c_void_p), C.c_void_p(0)) == 2
    assert b"must not be" in L.omh_last_error()
    sg = np.array([1, 0, -1], dtype=np.int8)
    first = C.c_int64()
    assert L.omh_factor_sign_mismatch_segments(dummy, 30, 4, 3, sp, sg.ctypes.data_as(C.c_void_p), C.byref(first),
                                               C.c_void_p(0)) == 2
    assert b"signs" in L.omh_last_error()
