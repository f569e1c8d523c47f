"""The pivoted fallback factorisation (FP64 LU with partial pivoting,
omh_lu_factor_f64 / omh_lu_solve_f64) and the breakdown detection that
switches to it.

* LU against numpy's LAPACK solve on random non-symmetric matrices (sizes
  that are and are not multiples of the 64-column panel, more right-hand
  sides than unknowns): 1e-12 relative Frobenius.
* The Head3 HeadMat in the reference's own unknown order (vertices first):
  its vertex block is only semi-definite (Appendix facts of the symmetric
  BEM: one null direction per inner mesh), so the pivot-free factorisation
  must not be trusted there.  factor() sees the pivot-sign pattern broken,
  falls back to LU, and the solve matches the oracle's (LAPACK on the
  oracle's HeadMat) to 1e-10; the same matrix in triangles-first order
  keeps the pivot-free factor.
* The gain pipeline with the pivoted factor forced (precision 'fp64_lu')
  against the oracle's Head3 gain: 1e-10.
Needs a HIP GPU."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd
    return openmeeg_amd


@pytest.mark.parametrize("n,nrhs", [(700, 9), (512, 3), (50, 120)])
def test_lu_matches_lapack(om, n, nrhs):
    import torch
    from openmeeg_amd.api import _lu_factor
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs))
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    Bd = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    print("LU n=%d nrhs=%d" % (n, nrhs), flush=True)
    F = _lu_factor(Ad, n)
    F.solve(Bd, nrhs)
    X = Bd.cpu().numpy().reshape(nrhs, n).T
    ref = np.linalg.solve(A, B)
    err = rel_fro(X, ref)
    ipiv = F.ipiv.cpu().numpy()
    print("  vs numpy %.3e, rows interchanged %d" % (err, int((ipiv != np.arange(n)).sum())), flush=True)
    assert err < 1e-12
    assert np.all(ipiv >= np.arange(n)) and np.all(ipiv < n)


def test_fallback_on_the_reference_order(om, c0_oracle):
    import torch
    from openmeeg_amd import _lib
    from openmeeg_amd.api import _dev_f64, _ptr, _stream, factor
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    n, nv = g.size, g.n_vertex_unknowns
    nt = n - nv
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, 5))
    ref = np.linalg.solve(c["H"], B)             # oracle HeadMat, reference order
    print("Head3 n=%d (vertices first, %d vertex unknowns)" % (n, nv), flush=True)
    H = _dev_f64(n * n)
    _lib.check(_lib.lib().omh_head_mat_ex(g._h, 3, 0, _ptr(H), n, _lib.ORDER_REFERENCE, _stream()))
    Hcopy = H.clone()
    F = factor(H, n, "fp64", n_positive=nt, rebuild=lambda: Hcopy.clone())
    Bd = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    F.solve(Bd, 5)
    X = Bd.cpu().numpy().reshape(5, n).T
    err = rel_fro(X, ref)
    print("  reference order: %s (%s), solve vs oracle %.3e" % (F.precision, getattr(F, "breakdown", "-"), err),
          flush=True)
    assert F.pivoted and err < 1e-10
    # triangles-first: the pivot-free factor passes its sign check
    from openmeeg_amd.api import _head_mat_device
    Ht = _head_mat_device(g)
    Ft = factor(Ht, n, "fp64", n_positive=nt, rebuild=lambda: None)
    assert not Ft.pivoted and Ft.precision == "fp64"


def test_pipeline_with_pivoted_factor(om, c0_oracle):
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    G, info = om.gain_eeg_pipeline(g, c["dip"], c["el"], precision="fp64_lu", domain_name="Brain")
    G = om.gain_to_numpy(G, len(c["el"]), len(c["dip"]))
    err = rel_fro(G, c["G"])
    print("Head3 gain with the pivoted LU vs oracle %.3e" % err, flush=True)
    assert info["pivoted_lu"] and err < 1e-10
