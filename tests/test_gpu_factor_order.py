"""Factorisation in the mesh-by-mesh segment order on the device: the
reordering kernels against numpy, factors of the reordered HeadMat against
the triangles-first ones (the same solutions to the refinement's and FP64
rounding), and the fallback when a reordered factorisation breaks down.
Needs a HIP GPU.

Tolerances: the reordering kernels are copies (bitwise; FP32 = numpy's
rounding of the FP64 entry); solutions of the two orders agree to 1e-10
relative (FP64 factors and refined mixed factors alike; the gain bar of the
FP64 path), residuals to 1e-10."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd
    return openmeeg_amd


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("nseg", [5, 12])
def test_sym_reorder_and_rows_match_numpy(om, nseg):
    """omh_sym_reorder's lower triangle (FP64 and FP32) is the symmetric
    segment permutation of the full matrix, its upper triangle untouched
    (5 segments: block copies between segment pairs; 12: more pairs than the
    block table holds, the per-row map); omh_reorder_rows_f64 gathers and
    scatters rows as numpy indexing does."""
    import torch
    from openmeeg_amd import _lib
    L = _lib.lib()
    n = 1000
    if nseg == 5:
        starts = np.array([0, 300, 520, 700, 910], dtype=np.int64)
        order = np.array([1, 3, 0, 4, 2], dtype=np.int32)
    else:
        starts = np.array([0, 40, 130, 200, 330, 400, 520, 610, 700, 777, 910, 950], dtype=np.int64)
        order = np.random.default_rng(2).permutation(nseg).astype(np.int32)
    ends = np.append(starts[1:], n)
    idx = np.concatenate([np.arange(starts[s], ends[s]) for s in order])
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n))
    A = A + A.T
    H = torch.from_numpy(A.copy()).cuda().reshape(-1)        # symmetric: row- and column-major alike
    ref = A[np.ix_(idx, idx)]
    lower = np.tril(np.ones((n, n), dtype=bool))               # [i, j], i >= j
    sp, op = starts.ctypes.data_as(C.c_void_p), order.ctypes.data_as(C.c_void_p)
    for elem, dt in ((8, torch.float64), (4, torch.float32)):
        D = torch.full((n * n,), float("nan"), dtype=dt, device="cuda")
        _lib.check(L.omh_sym_reorder(_p(H), n, _p(D), n, elem, n, len(starts), sp, op, C.c_void_p(0)))
        torch.cuda.synchronize()
        got = D.cpu().numpy().reshape(n, n).T                  # column-major -> [i, j]
        want = ref.astype(np.float64 if elem == 8 else np.float32)
        assert np.array_equal(got[lower], want[lower]), elem
        assert np.isnan(got[~lower]).all(), elem
        print("sym_reorder elem %d: lower bitwise, upper untouched" % elem, flush=True)
    ncol = 7
    Bm = rng.standard_normal((n, ncol))
    B = torch.from_numpy(np.asfortranarray(Bm).ravel(order="F").copy()).cuda()
    G = torch.empty_like(B)
    _lib.check(L.omh_reorder_rows_f64(_p(B), n, _p(G), n, n, ncol, len(starts), sp, op, 0, C.c_void_p(0)))
    R = torch.empty_like(B)
    _lib.check(L.omh_reorder_rows_f64(_p(G), n, _p(R), n, n, ncol, len(starts), sp, op, 1, C.c_void_p(0)))
    torch.cuda.synchronize()
    Gm = G.cpu().numpy().reshape(ncol, n).T
    assert np.array_equal(Gm, Bm[idx])
    assert torch.equal(R, B)


def _four_layers(om, k=12):
    from openmeeg_amd import spheres as S
    return om.make_nested_geometry(S.nested_sphere_meshes(k, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))


@pytest.mark.parametrize("precision", ["mixed16", "mixed", "fp64"])
def test_factor_order_solutions_agree(om, precision):
    """4 layers of 1442 vertices (n = 14408): the factor of the mesh-by-mesh
    reordered HeadMat solves the system as the triangles-first factor does
    (FP64: 1e-10; refined mixed factors: 1e-10), residual 1e-10; the FP64
    reordered factorisation works on a copy and leaves the matrix intact."""
    import torch
    from openmeeg_amd.api import _head_mat_device, factor
    g = _four_layers(om)
    n = g.size
    structure, order = g.factor_segments(), g.factor_order()
    assert order is not None
    npos = n - g.n_vertex_unknowns
    H = _head_mat_device(g, far_field=precision != "fp64")
    H0 = H.clone()
    torch.manual_seed(7)
    B = torch.randn(n * 24, dtype=torch.float64, device="cuda")
    if precision == "fp64":
        F1 = factor(H, n, "fp64", structure=structure, n_positive=npos, order=order)
        assert F1.order is not None and not hasattr(F1, "reorder_breakdown")
        assert torch.equal(H, H0)                      # factored a reordered copy
        X1 = F1.solve(B.clone(), 24)
        del F1
        F0 = factor(H, n, "fp64", structure=structure, n_positive=npos)
        X0 = F0.solve(B.clone(), 24)
        del F0
    else:
        F1 = factor(H, n, precision, structure=structure, n_positive=npos, order=order)
        assert F1.order is not None and not hasattr(F1, "reorder_breakdown")
        X1 = F1.solve(B.clone(), 24)
        it1 = F1.iterations
        del F1
        F0 = factor(H, n, precision, structure=structure, n_positive=npos)
        X0 = F0.solve(B.clone(), 24)
        it0 = F0.iterations
        del F0
        print("refinement passes: reordered %d, triangles-first %d" % (it1, it0), flush=True)
    diff = float(torch.linalg.norm(X1 - X0) / torch.linalg.norm(X0))
    Hm = H0.view(n, n)                                  # symmetric: the transpose is the matrix
    res = float(torch.linalg.norm(B.view(24, n).t() - Hm @ X1.view(24, n).t()) / torch.linalg.norm(B))
    print("n=%d %s: reordered vs triangles-first %.3e, residual %.3e" % (n, precision, diff, res), flush=True)
    assert diff < 1e-10
    assert res < 1e-10


def test_factor_order_breakdown_falls_back(om):
    """A reordered factorisation whose pivot signs do not match is redone in
    the matrix's own order (and, when that breaks down as well, by the
    pivoted LU): forced from the host by declaring every unknown positive
    (n_positive = n), which every vertex pivot contradicts.  The solution is
    still the FP64 one."""
    import torch
    from openmeeg_amd.api import _head_mat_device, factor
    g = _four_layers(om, 8)
    n = g.size
    structure, order = g.factor_segments(), g.factor_order()
    H = _head_mat_device(g, far_field=True)
    H0 = H.clone()
    F = factor(H, n, "mixed16", structure=structure, n_positive=n, order=order)
    print("reordered: %s; own order: %s" % (getattr(F, "reorder_breakdown", None), getattr(F, "breakdown", None)),
          flush=True)
    assert F.reorder_breakdown.startswith("pivot sign")
    assert F.order is None and F.pivoted
    torch.manual_seed(9)
    B = torch.randn(n * 4, dtype=torch.float64, device="cuda")
    X = F.solve(B.clone(), 4)
    res = float(torch.linalg.norm(B.view(4, n).t() - H0.view(n, n) @ X.view(4, n).t()) / torch.linalg.norm(B))
    assert res < 1e-10


def test_fp64_factor_order_without_room_factors_in_place(om, monkeypatch):
    """Without room for a reordered FP64 copy (forced from the host), the
    FP64 factorisation keeps the matrix's own order and factors it in place;
    the solution is the reordered factor's (1e-10)."""
    import torch
    from openmeeg_amd import api
    from openmeeg_amd.api import _head_mat_device, factor
    g = _four_layers(om, 8)
    n = g.size
    structure, order = g.factor_segments(), g.factor_order()
    npos = n - g.n_vertex_unknowns
    H = _head_mat_device(g)
    H0 = H.clone()
    torch.manual_seed(3)
    B = torch.randn(n * 8, dtype=torch.float64, device="cuda")
    F1 = factor(H.clone(), n, "fp64", structure=structure, n_positive=npos, order=order)
    assert F1.order is not None
    X1 = F1.solve(B.clone(), 8)
    del F1
    monkeypatch.setattr(api, "_room_for", lambda nbytes: False)
    F0 = factor(H, n, "fp64", structure=structure, n_positive=npos, order=order)
    assert F0.order is None and not torch.equal(H, H0)      # factored in place
    X0 = F0.solve(B.clone(), 8)
    diff = float(torch.linalg.norm(X1 - X0) / torch.linalg.norm(X0))
    print("fp64 without room for the reordered copy: in place, solutions agree to %.3e" % diff, flush=True)
    assert diff < 1e-10
