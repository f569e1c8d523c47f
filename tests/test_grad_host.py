"""Host logic of the conductivity gradient (EEGForward._sigma_gradient) on
the CPU: the block decomposition of dH/dsigma (S blocks ~ sum 1/sigma, N
blocks ~ sum sigma, deflation constant, 1/sigma on the potential rows of the
source matrix) is fed with the oracle's HeadMat / source matrix / EEG
interpolation and compared with central finite differences of the oracle's
gain.  The device GEMM is replaced by a torch CPU product of the same
column-major operands, so the test exercises the product's block ranges,
offsets and coefficient ratios, and nothing of the GPU."""
import ctypes as C

import numpy as np
import pytest

import openmeeg_amd as om
from openmeeg_amd import differentiable as DF
from openmeeg_amd import spheres as S
from oracle import oracle as O

torch = pytest.importorskip("torch")

RADII = (0.87, 0.92, 1.0)
LAYER_NAMES = ("Brain", "Skull", "Scalp")


def _cpu_gemm(ta, tb, m, n, k, A, lda, B, ldb, Cout, ldc, a_off=0, b_off=0, alpha=1.0, beta=0.0):
    Am = A.as_strided((k, m) if ta else (m, k), (1, lda), a_off)
    Bm = B.as_strided((n, k) if tb else (k, n), (1, ldb), b_off)
    Cm = Cout.as_strided((m, n), (1, ldc))
    P = alpha * ((Am.t() if ta else Am) @ (Bm.t() if tb else Bm))
    # beta = 0: C is not read (as the device GEMM; the buffers come from
    # torch.empty and may hold NaN patterns)
    Cm.copy_(P if beta == 0.0 else P + beta * Cm)


def _gain(meshes, layer_sigma, dip, E):
    og = O.nested_geometry(meshes, layer_sigma)
    H = O.unpack(O.head_mat(og), og.size)
    return O.gain_eeg(H, O.dip_source_mat(og, dip, "", levels=0), E), og, H


def test_sigma_gradient_host_logic(monkeypatch):
    meshes = S.nested_sphere_meshes(3, RADII)
    layer_sigma = np.array([1.0, 0.0125, 1.0])
    rng = np.random.default_rng(5)
    dip = S.random_dipoles(rng, 4, 0.75)
    el = S.sphere_electrodes(rng, 10)
    g = om.make_nested_geometry(meshes, tuple(layer_sigma))
    og = O.nested_geometry(meshes, tuple(layer_sigma))
    cols, vals = O.head2eeg(og, el)
    E = O.h2em_dense(og, cols, vals)
    G0, og, H = _gain(meshes, tuple(layer_sigma), dip, E)
    D = O.dip_source_mat(og, dip, "", levels=0)
    R = rng.standard_normal(G0.shape)
    n, nv = g.size, g.n_vertex_unknowns
    perm = np.r_[np.arange(nv, n), np.arange(nv)]        # triangles-first order
    Hd, Dd, Ed = H[np.ix_(perm, perm)], D[perm], E[:, perm]
    Y = np.linalg.solve(Hd, Ed.T)
    W = Y @ R
    Z = np.linalg.solve(Hd, Dd @ R.T)

    flat = lambda a: torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy())
    monkeypatch.setattr(DF, "_gemm", _cpu_gemm)
    monkeypatch.setattr(DF, "_dev_f64", lambda k: torch.empty(int(k), dtype=torch.float64))
    monkeypatch.setattr(DF, "_t", lambda: torch)
    model = DF.EEGForward.__new__(DF.EEGForward)
    model.geom, model.ne = g, len(el)
    model._sigma = g.conductivities()
    model._layout()
    names = g.domain_names()
    doms = (C.c_int32 * len(dip))(*([names.index("Brain")] * len(dip)))
    st = dict(H=flat(Hd), Y=flat(Y), D=flat(Dd), sigma=model._sigma.copy(), nd=len(dip), doms=doms)
    grad = model._sigma_gradient(st, flat(W), flat(Z))

    h = 1e-5
    for k, name in enumerate(LAYER_NAMES):
        s1, s2 = layer_sigma.copy(), layer_sigma.copy()
        s1[k] += h * layer_sigma[k]
        s2[k] -= h * layer_sigma[k]
        fd = ((_gain(meshes, tuple(s1), dip, E)[0] - _gain(meshes, tuple(s2), dip, E)[0]) * R).sum() \
            / (2 * h * layer_sigma[k])
        got = grad[names.index(name)]
        assert abs(got - fd) <= 1e-6 * abs(fd), (name, got, fd)
    assert grad[names.index("Air")] == 0.0
