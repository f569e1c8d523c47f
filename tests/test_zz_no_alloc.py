"""The allocation-free mixed solve (omh_mixed_solve_noalloc: its refinement
buffers in the library's fixed device array) against omh_mixed_solve, both
through the C ABI, on two calls in a row with the second smaller: Head3
(n = 4486, 64 right-hand sides), then Head1 (n = 286, 5 right-hand sides).
Same factor, same right-hand sides: the solutions agree to 1e-14 (relative
Frobenius; the kernels are the same, so they are expected bitwise equal)
with the same number of corrections.  A call beyond the array's 2^22
elements of n x nrhs is refused with invalid argument before any device
work.  Needs a HIP GPU."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

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


def _split_factor(L, H, n):
    import torch
    L32 = torch.empty(n * n, dtype=torch.float32, device="cuda")
    L.check(L.lib().omh_convert_f64_f32(_p(H), n, _p(L32), n, n, n, 1, C.c_void_p(0)))
    work = torch.empty(int(L.lib().omh_factor_workspace_bytes(n, 4)), dtype=torch.uint8, device="cuda")
    info = C.c_int32()
    L.check(L.lib().omh_sym_factor_f32_split(_p(L32), n, n, _p(work), C.byref(info), C.c_void_p(0)))
    assert info.value == 0
    return L32, work


def _solve(L, fn, H, L32, work, B, X, n, r):
    it, rel = C.c_int32(), C.c_double()
    rc = fn(_p(H), n, n, _p(L32), n, _p(work), _p(B), n, _p(X), n, r, 10, 1e-12, C.byref(it), C.byref(rel),
            C.c_void_p(0))
    return rc, it.value, rel.value


def test_noalloc_solve_matches_mixed_solve(om):
    import torch
    from openmeeg_amd import _lib as L
    from openmeeg_amd.api import _head_mat_device
    cases = [("Head3", os.path.join(GOLDEN, "head3", "Head3.geom"), os.path.join(GOLDEN, "head3", "Head3.cond"), 64),
             ("Head1", os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"), 5)]
    for name, gf, cf, r in cases:
        g = om.read_geometry(gf, cf)
        n = g.size
        print("no-alloc solve: %s n=%d nrhs=%d" % (name, n, r), flush=True)
        H = _head_mat_device(g)
        L32, work = _split_factor(L, H, n)
        B = torch.from_numpy(np.random.default_rng(r).standard_normal(n * r)).cuda()
        X1 = torch.full((n * r,), np.nan, dtype=torch.float64, device="cuda")
        X2 = torch.full((n * r,), np.nan, dtype=torch.float64, device="cuda")
        rc1, it1, rel1 = _solve(L, L.lib().omh_mixed_solve, H, L32, work, B, X1, n, r)
        rc2, it2, rel2 = _solve(L, L.lib().omh_mixed_solve_noalloc, H, L32, work, B, X2, n, r)
        torch.cuda.synchronize()
        assert rc1 == 0 and rc2 == 0, L.lib().omh_last_error()
        x1, x2 = X1.cpu().numpy(), X2.cpu().numpy()
        Hh = H.view(n, n).cpu().numpy()
        res = rel_fro(Hh @ x2.reshape(r, n).T, B.cpu().numpy().reshape(r, n).T)
        print("  %d / %d corrections, rel update %.2e / %.2e, |X1 - X2| %.3e, bitwise %s, residual %.2e"
              % (it1, it2, rel1, rel2, rel_fro(x2, x1), np.array_equal(x1, x2), res), flush=True)
        assert it1 == it2
        assert rel_fro(x2, x1) < 1e-14
        assert res < 1e-12
    # beyond the fixed array: refused on the host (buffers sized for the call)
    g = om.read_geometry(*cases[0][1:3])
    n, r = g.size, 1000
    assert n * r > (1 << 22)
    H = _head_mat_device(g)
    L32, work = _split_factor(L, H, n)
    B = torch.zeros(n * r, dtype=torch.float64, device="cuda")
    X = torch.zeros(n * r, dtype=torch.float64, device="cuda")
    rc, _, _ = _solve(L, L.lib().omh_mixed_solve_noalloc, H, L32, work, B, X, n, r)
    assert rc == 2 and b"fixed buffer" in L.lib().omh_last_error()
