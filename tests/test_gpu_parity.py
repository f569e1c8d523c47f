"""Parity of the HIP path (through the C ABI) with the reference's goldens
and the CPU oracle.  Needs a HIP GPU.

Tolerances (relative Frobenius unless noted):
  * HeadMat / DipSourceMat entries vs goldens or oracle: 1e-13 (FP64, device
    math library vs glibc: a few ulp per entry);
  * Head2EEGMat: exact indices, weights to 1e-14;
  * inverse and gains (FP64 path): 1e-10, the bar the reference's gain must
    meet;
  * mixed-precision gain vs the FP64 GPU gain: 1e-6.
"""
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


@pytest.fixture(scope="module")
def head1(om):
    g = om.read_geometry(os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"))
    from oracle import oracle as O
    dip = O.read_dipoles(os.path.join(GOLDEN, "head1", "Head1.dip"))
    sens = om.Sensors()
    sens.load(os.path.join(GOLDEN, "head1", "Head1.patches"))
    return g, dip, sens


@pytest.fixture(scope="module")
def nna1(om):
    """HeadNNa1 of the reference's regression suite: non-nested, the north and
    south meshes share the cut mesh's vertices."""
    from oracle import oracle as O
    d = os.path.join(GOLDEN, "headnna1")
    gf, cf = os.path.join(d, "HeadNNa1.geom"), os.path.join(d, "HeadNNa1.cond")
    dip = O.read_dipoles(os.path.join(d, "HeadNNa1.dip"))
    sens = om.Sensors()
    sens.load(os.path.join(d, "HeadNNa1.patches"))
    el = np.loadtxt(os.path.join(d, "HeadNNa1.patches"), usecols=(1, 2, 3))
    return om.read_geometry(gf, cf), O.load_geometry(gf, cf), dip, sens, el


def test_nonnested_nna1(om, nna1, monkeypatch):
    """HeadNNa1 (shared vertices): the FP64 HeadMat against the oracle's
    (1e-13); the far-field assembly (fused pass: shared vertices in its
    vertex gathers and the N blocks' atomic accumulation) against the
    per-operator kernels (1e-13 of the largest entry), symmetric; the FP64
    adjoint gain against the oracle's (1e-10)."""
    import torch
    from oracle import oracle as O
    from openmeeg_amd.api import _head_mat_device
    g, og, dip, sens, el = nna1
    n = g.size
    H = O.unpack(O.head_mat(og), og.size)
    hm = om.HeadMat(g)
    e_h = rel_fro(hm.data(), O.pack(H))
    Hf = _head_mat_device(g, far_field=True).view(n, n)
    monkeypatch.setenv("OMH_FAR_SEPARATE", "1")
    Hs = _head_mat_device(g, far_field=True).view(n, n)
    monkeypatch.delenv("OMH_FAR_SEPARATE")
    e_f = float((Hf - Hs).abs().max() / Hs.abs().max())
    cols, vals = O.head2eeg(og, el)
    Gref = O.gain_eeg(H, O.dip_source_mat(og, dip, ""), O.h2em_dense(og, cols, vals))
    G = om.GainEEGadjoint(g, dip, hm, om.Head2EEGMat(g, sens)).array()
    e_g = rel_fro(G, Gref)
    print("NNa1 n=%d: HeadMat vs oracle %.3e, far fused vs separate %.3e, gain vs oracle %.3e" % (n, e_h, e_f, e_g),
          flush=True)
    assert e_h < 1e-13 and e_f < 1e-13 and e_g < 1e-10
    assert torch.equal(Hf, Hf.t())


def test_meg_head1_matches_golden(om, head1):
    """MEG forward on Head1 against the reference's goldens with
    Head1.squids: Head2MEGMat (Ferguson operator on the device), 
    DipSource2MEGMat, GainMEG (HeadMat inverse) and GainMEGadjoint (stored
    factor)."""
    g, dip, _ = head1
    sq = om.Sensors()
    sq.load(os.path.join(GOLDEN, "head1", "Head1.squids"))
    h2mm = om.Head2MEGMat(g, sq)
    ds2mm = om.DipSource2MEGMat(dip, sq)
    e_h = rel_fro(h2mm.array(), np.load(os.path.join(GOLDEN, "head1", "h2mm.npy")))
    e_s = rel_fro(ds2mm.array(), np.load(os.path.join(GOLDEN, "head1", "ds2mm.npy")))
    hm = om.HeadMat(g)
    dsm = om.DipSourceMat(g, dip, "Brain")
    ref = np.load(os.path.join(GOLDEN, "head1", "dgmm.npy"))
    e_g = rel_fro(om.GainMEG(hm.inverse(), dsm, h2mm, ds2mm).array(), ref)
    e_a = rel_fro(om.GainMEGadjoint(g, dip, hm, h2mm, ds2mm).array(), ref)
    sens = om.Sensors()
    sens.load(os.path.join(GOLDEN, "head1", "Head1.patches"))
    both = om.GainEEGMEGadjoint(g, dip, hm, om.Head2EEGMat(g, sens), h2mm, ds2mm)
    e_be = rel_fro(both.EEGleadfield.array(), np.load(os.path.join(GOLDEN, "head1", "dgem.npy")))
    e_bm = rel_fro(both.MEGleadfield.array(), ref)
    print("Head1 MEG: h2mm %.3e, ds2mm %.3e, GainMEG %.3e, GainMEGadjoint %.3e, GainEEGMEGadjoint %.3e / %.3e"
          % (e_h, e_s, e_g, e_a, e_be, e_bm), flush=True)
    assert e_h < 1e-10 and e_s < 1e-14 and e_g < 1e-10 and e_a < 1e-10 and e_be < 1e-10 and e_bm < 1e-10
    assert both.nlin() == 32 + h2mm.nlin()


def test_meg_nonnested_nna1(om, nna1):
    """Head2MEGMat and the MEG gain on the non-nested HeadNNa1 (meshes that
    share vertices: every mesh's Ferguson term lands in the shared vertex
    columns) against the oracle."""
    from oracle import oracle as O
    g, og, dip, _, _ = nna1
    path = os.path.join(GOLDEN, "headnna1", "HeadNNa1.squids")
    sq = om.Sensors()
    sq.load(path)
    pts, dirs, w, sens = O.read_squids(path)
    Mref = O.head2meg(og, pts, dirs, w, sens)
    Sref = O.dip_source2meg(dip, pts, dirs, w, sens)
    h2mm = om.Head2MEGMat(g, sq)
    ds2mm = om.DipSource2MEGMat(dip, sq)
    e_h = rel_fro(h2mm.array(), Mref)
    H = O.unpack(O.head_mat(og), og.size)
    Gref = O.gain_meg(H, O.dip_source_mat(og, dip, ""), Mref, Sref)
    e_g = rel_fro(om.GainMEGadjoint(g, dip, om.HeadMat(g), h2mm, ds2mm).array(), Gref)
    print("NNa1 MEG: h2mm vs oracle %.3e, gain %.3e" % (e_h, e_g), flush=True)
    assert e_h < 1e-12 and e_g < 1e-10


def test_eeg_meg_pipeline(om, c0_oracle):
    """The combined device pipeline (GainEEGMEGadjoint, gain.h:105-128: one
    factor, one solve with the EEG and MEG rows) on the Head3 head with
    80 tilted magnetometers: its EEG gain equals gain_eeg_pipeline's (1e-12),
    its MEG gain GainMEGadjoint's from the reference-API matrices (1e-10);
    mixed16 within 1e-6 of FP64.  Independent of the reference: radial
    magnetometers around a spherical head see no field from the volume
    currents (Sarvas), so their MEG gain is the dipoles' own field up to the
    discretisation (here 5%)."""
    from openmeeg_amd import spheres as S
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    rng = np.random.default_rng(11)
    p, o, w = S.sphere_squids(rng, 80)
    sq = om.Sensors(p, orientations=o, weights=w)
    dip, el = c["dip"], c["el"]
    ne, ns, nd = len(el), 80, len(dip)
    Ge, Gm, info = om.gain_eeg_meg_pipeline(g, dip, el, sq, precision="fp64", domain_name="Brain")
    Ge, Gm = om.gain_to_numpy(Ge, ne, nd), om.gain_to_numpy(Gm, ns, nd)
    G0, _ = om.gain_eeg_pipeline(g, dip, el, precision="fp64", domain_name="Brain")
    e_eeg = rel_fro(Ge, om.gain_to_numpy(G0, ne, nd))
    h2mm, ds2mm = om.Head2MEGMat(g, sq), om.DipSource2MEGMat(dip, sq)
    Gref = om.GainMEGadjoint(g, dip, om.HeadMat(g), h2mm, ds2mm).array()
    e_meg = rel_fro(Gm, Gref)
    _, Gm16, info16 = om.gain_eeg_meg_pipeline(g, dip, el, sq, precision="mixed16", domain_name="Brain")
    e_mix = rel_fro(om.gain_to_numpy(Gm16, ns, nd), Gm)
    pr, orad, _ = S.sphere_squids(np.random.default_rng(12), 40, tilt=0.0)
    sr = om.Sensors(pr, orientations=orad)
    _, Gr, _ = om.gain_eeg_meg_pipeline(g, dip, el, sr, precision="fp64", domain_name="Brain")
    Gr = om.gain_to_numpy(Gr, 40, nd)
    e_rad = rel_fro(Gr, om.DipSource2MEGMat(dip, sr).array())
    print("EEG+MEG c0: EEG vs EEG pipeline %.3e, MEG vs GainMEGadjoint %.3e, mixed16 vs fp64 %.3e (%d "
          "refinements), radial sensors vs dipole field %.3e" % (e_eeg, e_meg, e_mix, info16["refine_iterations"],
                                                                  e_rad), flush=True)
    assert e_eeg < 1e-12 and e_meg < 1e-10 and e_mix < 1e-6 and e_rad < 0.05


def test_mfma_lane_layout(om):
    """D(i,j) of one 16x16x4 block sits in lane j + 16*(i % 4), element
    i // 4, for the FP64 instruction and in lane j + 16*(i // 4), element
    i % 4, for the FP32 one (the layouts the GEMM epilogue uses)."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    out = torch.zeros(768, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().omh_probe_mfma_layout(C.c_void_p(out.data_ptr()), C.c_void_p(0)))
    torch.cuda.synchronize()
    i = np.arange(16)
    A = 4.0 * i[:, None] + np.arange(4)[None, :] + 1.0
    B = (np.arange(4)[:, None] + 1.0) * (i[None, :] % 5 + 1.0)
    ref = A @ B
    o = out.cpu().numpy()
    rows = {"f64": lambda lane, r: (lane >> 4) + 4 * r, "f32": lambda lane, r: 4 * (lane >> 4) + r}
    for part, kind in ((o[:256], "f64"), (o[256:512], "f32")):
        raw = part.reshape(64, 4)
        got = np.zeros((16, 16))
        for lane in range(64):
            for r in range(4):
                got[rows[kind](lane, r), lane & 15] = raw[lane, r]
        np.testing.assert_array_equal(got, ref, err_msg=kind)
    # FP16 16x16x32: operands A(i, k) in lane i + 16 (k // 8), element k % 8
    # and B(k, j) in lane j + 16 (k // 8), element k % 8; output as FP32
    k32 = np.arange(32)
    A16 = (i[:, None] + 2 * k32[None, :]) % 7 - 3.0
    B16 = (3 * k32[:, None] + i[None, :]) % 5 - 2.0
    raw = o[512:].reshape(64, 4)
    got = np.zeros((16, 16))
    for lane in range(64):
        for r in range(4):
            got[rows["f32"](lane, r), lane & 15] = raw[lane, r]
    np.testing.assert_array_equal(got, A16 @ B16, err_msg="f16")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_matches_numpy(om, dtype, ta, tb):
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    rng = np.random.default_rng(10 + 2 * ta + tb)
    m, n, k = 131, 77, 203
    npdt = np.float64 if dtype == "f64" else np.float32
    tdt = torch.float64 if dtype == "f64" else torch.float32
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(npdt)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(npdt)
    Cm = rng.standard_normal((m, n)).astype(npdt)
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dB = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    dC = torch.from_numpy(np.asfortranarray(Cm).ravel(order="F").copy()).cuda()
    fn = _lib.lib().omh_gemm_f64 if dtype == "f64" else _lib.lib().omh_gemm_f32
    _lib.check(fn(ta, tb, m, n, k, 1.5, C.c_void_p(dA.data_ptr()), A.shape[0], C.c_void_p(dB.data_ptr()), B.shape[0],
                  -0.5, C.c_void_p(dC.data_ptr()), m, C.c_void_p(0)))
    torch.cuda.synchronize()
    ref = 1.5 * (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64) - 0.5 * Cm
    got = dC.cpu().numpy().reshape(n, m).T
    assert rel_fro(got, ref) < (1e-14 if dtype == "f64" else 2e-6)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ta", [0, 1])
def test_gemm_large_tile_order(om, dtype, ta):
    """Products with more than 1024 interior tiles run in the XCD-aware,
    band-swizzled tile order (plus partial edge strips)."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    rng = np.random.default_rng(31 + ta)
    m, n, k = 3001, 2003, 40
    npdt = np.float64 if dtype == "f64" else np.float32
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(npdt)
    B = rng.standard_normal((k, n)).astype(npdt)
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dB = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    dC = torch.zeros(m * n, dtype=dA.dtype, device="cuda")
    fn = _lib.lib().omh_gemm_f64 if dtype == "f64" else _lib.lib().omh_gemm_f32
    _lib.check(fn(ta, 0, m, n, k, 1.0, C.c_void_p(dA.data_ptr()), A.shape[0], C.c_void_p(dB.data_ptr()), k,
                  0.0, C.c_void_p(dC.data_ptr()), m, C.c_void_p(0)))
    torch.cuda.synchronize()
    ref = (A.T if ta else A).astype(np.float64) @ B.astype(np.float64)
    got = dC.cpu().numpy().reshape(n, m).T
    assert rel_fro(got, ref) < (1e-14 if dtype == "f64" else 2e-6)


@pytest.mark.parametrize("shape", [(2200, 2100, 100), (2304, 2048, 256), (4100, 256, 10242)])
def test_gemm_f64_dma_tiles(om, monkeypatch, shape):
    """FP64 A B^T on 128 x 128 tiles fed by direct global-to-LDS loads (the
    kernel of the FP64 trailing update and of the structured residual) over
    the interior, the 64 x 64 tiles on the strips around it: against numpy
    (1e-14) and bitwise against the 64 x 64 tiles throughout
    (OMH_GEMM_F64_DMA=0: the same MFMAs over k in the same order).  k a
    multiple of the 16-wide k-slice and not (a zero-padded last slice),
    partial edge tiles, a residual-like 256-column output."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    m, n, k = shape
    rng = np.random.default_rng(m + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    Cm = rng.standard_normal((m, n))
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dB = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    out = {}
    for mode in ("dma", "tiles64"):
        if mode == "tiles64":
            monkeypatch.setenv("OMH_GEMM_F64_DMA", "0")
        dC = torch.from_numpy(np.asfortranarray(Cm).ravel(order="F").copy()).cuda()
        _lib.check(_lib.lib().omh_gemm_f64(0, 1, m, n, k, -1.0, C.c_void_p(dA.data_ptr()), m,
                                           C.c_void_p(dB.data_ptr()), n, 1.0, C.c_void_p(dC.data_ptr()), m,
                                           C.c_void_p(0)))
        torch.cuda.synchronize()
        out[mode] = dC.cpu().numpy().reshape(n, m).T
        monkeypatch.delenv("OMH_GEMM_F64_DMA", raising=False)
    err = rel_fro(out["dma"], Cm - A @ B.T)
    print("%s: vs numpy %.2e, bitwise the 64 x 64 tiles: %s" % (shape, err, np.array_equal(out["dma"],
                                                                                             out["tiles64"])),
          flush=True)
    assert err < 1e-14
    assert np.array_equal(out["dma"], out["tiles64"])


def test_fp64_factor_dma_tiles_bitwise(om, monkeypatch):
    """The FP64 factorisation's trailing updates (lower products) on the
    128 x 128 DMA tiles give the factor bitwise, with and without the block
    structure, and the refined mixed solution through the structured
    residual (its products with a transposed solution) is bitwise the same
    as with the 64 x 64 tiles: 3 layers of 2562 vertices (n = 17926)."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    n = g.size
    H = _head_mat_device(g)
    for structure in (g.factor_segments(), None):
        L = {}
        for mode in ("dma", "tiles64"):
            if mode == "tiles64":
                monkeypatch.setenv("OMH_GEMM_F64_DMA", "0")
            F = factor(H.clone(), n, "fp64", structure=structure)
            L[mode] = F.L.view(n, n).t().tril().clone()
            del F
            monkeypatch.delenv("OMH_GEMM_F64_DMA", raising=False)
        print("n=%d structure %s: FP64 factor max |diff| %.3e" % (n, structure is not None,
                                                                  float((L["dma"] - L["tiles64"]).abs().max())),
              flush=True)
        assert torch.equal(L["dma"], L["tiles64"])
        del L
    Hf = _head_mat_device(g, far_field=True)
    B = torch.from_numpy(np.random.default_rng(3).standard_normal(n * 64)).cuda()
    F = factor(Hf, n, "mixed16", structure=g.factor_segments())
    X = {}
    for mode in ("dma", "tiles64"):
        if mode == "tiles64":
            monkeypatch.setenv("OMH_GEMM_F64_DMA", "0")
        X[mode] = B.clone()
        F.solve(X[mode], 64, max_iter=10, tol=1e-10)
        torch.cuda.synchronize()
        monkeypatch.delenv("OMH_GEMM_F64_DMA", raising=False)
    print("refined solution max |diff| %.3e after %d iterations" % (float((X["dma"] - X["tiles64"]).abs().max()),
                                                                    F.iterations), flush=True)
    assert torch.equal(X["dma"], X["tiles64"])


def test_fp64_diag_inverse_narrow_bitwise(om, monkeypatch):
    """The FP64 diagonal-block inverse on 16-column workgroups (a third of
    the LDS of the 32-column ones, so that it runs beside the trailing
    update) sums every entry in the same order: factor and workspace
    (diagonal inverses, signs) bitwise those of the 32-column kernel.  3
    layers of 2562 vertices (n = 17926: a last diagonal block of 6
    columns), with the block structure."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    n = g.size
    H = _head_mat_device(g)
    out = {}
    for mode in ("narrow", "wide"):
        if mode == "wide":
            monkeypatch.setenv("OMH_DIAG_INV64_WIDE", "1")
        F = factor(H.clone(), n, "fp64", structure=g.factor_segments())
        torch.cuda.synchronize()
        out[mode] = (F.L.view(n, n).t().tril().clone(), F.work.clone())
        del F
        monkeypatch.delenv("OMH_DIAG_INV64_WIDE", raising=False)
    print("n=%d: factor max |diff| %.3e, workspace equal %s" % (
        n, float((out["narrow"][0] - out["wide"][0]).abs().max()), bool(torch.equal(out["narrow"][1], out["wide"][1]))),
        flush=True)
    assert torch.equal(out["narrow"][0], out["wide"][0])
    assert torch.equal(out["narrow"][1], out["wide"][1])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape", [(128, 320, 40000), (70, 133, 9001)])
def test_gemm_split_k_matches_numpy(om, dtype, ta, tb, shape):
    """Outputs of few tiles with a long inner dimension (the gain product
    Y^T D has 256 x 5000 outputs over 102408 unknowns) run split-K: chunks
    of the inner dimension in one launch, summed in order.  Full tiles and
    partial edge tiles, aligned and odd leading dimensions, beta != 0."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    m, n, k = shape
    rng = np.random.default_rng(7 + 4 * ta + 2 * tb + m)
    npdt = np.float64 if dtype == "f64" else np.float32
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(npdt)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(npdt)
    Cm = rng.standard_normal((m, n)).astype(npdt)
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dB = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    dC = torch.from_numpy(np.asfortranarray(Cm).ravel(order="F").copy()).cuda()
    fn = _lib.lib().omh_gemm_f64 if dtype == "f64" else _lib.lib().omh_gemm_f32
    for rep in range(2):   # the second call reuses the grown workspace
        dC.copy_(torch.from_numpy(np.asfortranarray(Cm).ravel(order="F").copy()))
        _lib.check(fn(ta, tb, m, n, k, 1.5, C.c_void_p(dA.data_ptr()), A.shape[0], C.c_void_p(dB.data_ptr()),
                      B.shape[0], -0.5, C.c_void_p(dC.data_ptr()), m, C.c_void_p(0)))
        torch.cuda.synchronize()
        ref = 1.5 * (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64) - 0.5 * Cm
        got = dC.cpu().numpy().reshape(n, m).T
        err = rel_fro(got, ref)
        print("%s ta=%d tb=%d %s: %.2e" % (dtype, ta, tb, shape, err), flush=True)
        assert err < (1e-14 if dtype == "f64" else 2e-6)


def test_n_blocks_chunked(om, monkeypatch):
    """The N blocks are assembled in column chunks (1 GiB of per-triangle
    sums each, several chunks from about 2200 vertices per surface on); with
    the chunk width capped at 37 columns the c0 HeadMat (642 vertices per
    surface, 18 chunks per block) is bitwise the single-chunk one."""
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device
    g = om.make_nested_geometry(S.nested_sphere_meshes(8, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    H1 = _head_mat_device(g).cpu().numpy()
    monkeypatch.setenv("OMH_N_CHUNK_COLS", "37")
    H2 = _head_mat_device(g).cpu().numpy()
    assert np.array_equal(H1, H2)


def test_sparse_transpose_dense(om):
    """Right-hand sides of the adjoint solve (linsolve's rows, gain.h:28-49):
    the dense transpose of a 3-entries-per-row sparse matrix, with repeated
    rows adding and rows outside [0, n) ignored, into a padded leading
    dimension whose padding rows must stay zero."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    rng = np.random.default_rng(5)
    n, ne, ld = 1000, 77, 1003
    rows = rng.integers(0, n, size=(ne, 3)).astype(np.int64)
    rows[5, 1] = rows[5, 0]          # repeated row: the values add
    rows[9, 2] = n + 4               # outside: ignored
    vals = rng.standard_normal((ne, 3))
    ref = np.zeros((n, ne))
    for e in range(ne):
        for k in range(3):
            if rows[e, k] < n:
                ref[rows[e, k], e] += vals[e, k]
    dr = torch.from_numpy(rows.ravel().copy()).cuda()
    dv = torch.from_numpy(vals.ravel().copy()).cuda()
    out = torch.full((ld * ne,), 7.0, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().omh_sparse_transpose_dense(C.c_void_p(dr.data_ptr()), C.c_void_p(dv.data_ptr()), ne, n,
                                                     C.c_void_p(out.data_ptr()), ld, C.c_void_p(0)))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(ne, ld).T
    np.testing.assert_array_equal(got[:n], ref)
    assert np.all(got[n:] == 7.0)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed16"])
def test_factor_solve_residual_large(om, precision):
    """HeadMat of the 2562-vertex 3-layer head (n = 17926): the trailing
    updates use the L2-aware lower-triangle order.  Residual of H X = B for
    random right-hand sides."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    n = g.size
    H = _head_mat_device(g)
    Href = H.clone().view(n, n)
    rng = np.random.default_rng(4)
    B = torch.from_numpy(rng.standard_normal(n * 8)).cuda()
    F = factor(H, n, precision)
    X = B.clone()
    F.solve(X, 8)
    R = B.view(8, n) - (X.view(8, n) @ Href)          # H symmetric: rows of X^T H
    res = float(torch.linalg.norm(R) / torch.linalg.norm(B))
    print("n=%d %s residual %.3e" % (n, precision, res), flush=True)
    assert res < 1e-12


@pytest.mark.parametrize("layers,precision", [(3, "mixed16"), (4, "mixed16"), (4, "mixed"), (4, "fp64")])
def test_block_structured_factor_and_solve(om, monkeypatch, layers, precision):
    """The factorisations told the HeadMat's block structure
    (Geometry.factor_segments) skip trailing-update work that only multiplies
    structural zeros: the factor must be bitwise the dense one (the skipped
    tiles would add exact zeros).  The structured solves skip L's and H's
    zero blocks: the same solution (FP64 triangular solves: the same sums;
    refined: the residual's sums in another order, 1e-12).  3 layers of 2562
    vertices (n = 17926) and 4 layers of 1442 vertices (n = 14408); a
    pattern that declares a non-zero block zero is refused."""
    import ctypes as C
    import torch
    from openmeeg_amd import spheres as S, _lib
    from openmeeg_amd.api import _head_mat_device, factor
    monkeypatch.setenv("OMH_SYRK_CLOCK", "0")   # k slices in order: bitwise comparable
    if layers == 3:
        g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    else:
        g = om.make_nested_geometry(S.nested_sphere_meshes(12, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    starts, hfill = g.factor_segments()
    assert hfill.sum() < hfill.shape[0] * (hfill.shape[0] - 1) // 2   # some blocks of H are zero
    H = _head_mat_device(g, far_field=precision != "fp64")
    inplace = precision == "fp64"
    Fd = factor(H.clone() if inplace else H, n, precision)
    Fs = factor(H.clone() if inplace else H, n, precision, structure=(starts, hfill))
    Ld = Fd.L.view(n, n).t().tril()       # lower triangle (column-major storage)
    Ls = Fs.L.view(n, n).t().tril()
    print("n=%d %s segments %s, zero blocks of H %d: factor max |diff| %.3e" % (
        n, precision, starts.tolist(), int((np.tril(np.ones_like(hfill), -1) - hfill).sum()),
        float((Ld - Ls).abs().max())), flush=True)
    assert torch.equal(Ld, Ls)
    del Ld, Ls
    torch.manual_seed(5)
    B = torch.randn(n * 24, dtype=torch.float64, device="cuda")
    Xd, Xs = Fd.solve(B.clone(), 24), Fs.solve(B.clone(), 24)
    diff = float(torch.linalg.norm(Xd - Xs) / torch.linalg.norm(Xd))
    print("solve: structured vs dense %.3e (bitwise %s), refinement steps %d / %d" % (
        diff, bool(torch.equal(Xd, Xs)), Fd.iterations, Fs.iterations), flush=True)
    assert diff < (1e-15 if precision == "fp64" else 1e-12)
    del Fd, Fs, Xd, Xs
    bad = hfill.copy()
    s, t = np.argwhere(np.tril(hfill, -1))[0]
    bad[s, t] = 0                          # a non-zero block of H declared zero
    L32 = torch.empty(n * n, dtype=torch.float32, device="cuda")
    check = _lib.check
    check(_lib.lib().omh_convert_f64_f32(C.c_void_p(H.data_ptr()), n, C.c_void_p(L32.data_ptr()), n, n, n, 1,
                                         C.c_void_p(0)))
    work = torch.empty(int(_lib.lib().omh_factor_workspace_bytes(n, 4)), dtype=torch.uint8, device="cuda")
    info = C.c_int32()
    rc = _lib.lib().omh_sym_factor_f32_split_blocks(
        C.c_void_p(L32.data_ptr()), n, n, C.c_void_p(work.data_ptr()), len(starts),
        np.ascontiguousarray(starts).ctypes.data_as(C.c_void_p), np.ascontiguousarray(bad).ctypes.data_as(C.c_void_p),
        C.byref(info), C.c_void_p(0))
    assert rc == 2
    assert b"declared zero" in _lib.lib().omh_last_error()


def test_split_factor_tiles_agree(om, monkeypatch):
    """The split-FP16 factorisation's main update on 256 x 256 tiles (direct
    global-to-LDS operand loads, 8 waves) and on 128 x 128 tiles
    (the default) sums every entry's products in the same order: the
    factors must be bitwise equal.  4 layers of 2562 vertices (n = 25608), so
    that the early groups run on 256-row tiles, with and without the block
    structure."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    H = _head_mat_device(g, far_field=True)
    monkeypatch.setenv("OMH_SYRK_CLOCK", "0")   # k slices in order: bitwise comparable
    for structure in (None, g.factor_segments()):
        monkeypatch.setenv("OMH_SYRK_TILE", "256")
        F256 = factor(H, n, "mixed16", structure=structure)
        monkeypatch.delenv("OMH_SYRK_TILE")
        L256 = F256.L.view(n, n).t().tril().clone()
        del F256
        F128 = factor(H, n, "mixed16", structure=structure)
        L128 = F128.L.view(n, n).t().tril()
        print("n=%d structure %s: max |L256 - L128| %.3e" % (n, structure is not None,
                                                             float((L256 - L128).abs().max())), flush=True)
        assert torch.equal(L256, L128)
        del F128, L128, L256


def test_split_factor_pipelined_tile_bitwise(om, monkeypatch):
    """The software-pipelined 128 x 128 split-update tile (operand slices by
    direct global-to-LDS loads, fragments double-buffered in registers) and
    the register-staged one (OMH_SYRK_PP=0) sum every entry's products in the
    same order: bitwise equal factors, with and without the block structure,
    on 4 layers of 2562 vertices (n = 25608)."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    H = _head_mat_device(g, far_field=True)
    monkeypatch.setenv("OMH_SYRK_CLOCK", "0")   # k slices in order: bitwise comparable
    for structure in (None, g.factor_segments()):
        monkeypatch.setenv("OMH_SYRK_PP", "0")
        F0 = factor(H, n, "mixed16", structure=structure)
        monkeypatch.delenv("OMH_SYRK_PP")
        L0 = F0.L.view(n, n).t().tril().clone()
        del F0
        F1 = factor(H, n, "mixed16", structure=structure)
        L1 = F1.L.view(n, n).t().tril()
        print("n=%d structure %s: max |L_pp - L_staged| %.3e" % (n, structure is not None,
                                                                 float((L1 - L0).abs().max())), flush=True)
        assert torch.equal(L0, L1)
        del F1, L1, L0


def test_split_factor_wide_tiles_bitwise(om, monkeypatch):
    """The main split update on 256 x 128 tiles (8 waves, one workgroup per
    CU, a ring of three operand slices in LDS; the row pair of a tile may
    come from two intervals of the block structure and the last pair may
    have one unit) sums every entry's products in the same order as the
    128 x 128 tiles (OMH_SYRK_WIDE=0): bitwise equal factors, with and
    without the block structure, on 4 layers of 2562 vertices (n = 25608,
    an odd number of 128-row units below the first groups)."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    H = _head_mat_device(g, far_field=True)
    monkeypatch.setenv("OMH_SYRK_CLOCK", "0")   # k slices in order: bitwise comparable
    for structure in (g.factor_segments(), None):
        monkeypatch.setenv("OMH_SYRK_WIDE", "0")
        F0 = factor(H, n, "mixed16", structure=structure)
        monkeypatch.delenv("OMH_SYRK_WIDE")
        L0 = F0.L.view(n, n).t().tril().clone()
        del F0
        F1 = factor(H, n, "mixed16", structure=structure)
        L1 = F1.L.view(n, n).t().tril()
        print("n=%d structure %s: max |L_wide - L_128| %.3e" % (n, structure is not None,
                                                               float((L1 - L0).abs().max())), flush=True)
        assert torch.equal(L0, L1)
        del F1, L1, L0


def test_split_factor_slice_clock(om, monkeypatch):
    """The split update's slice clock changes only the order in which each
    tile sums its k slices (the tiles in flight start at the clock's slice
    and wrap around): the factor with the clock equals the factor without it
    to FP32 rounding (1e-5 of the largest entry of L, well below the factor's
    own split-FP16 error), and the refined solutions of both agree to 1e-12.
    4 layers of 2562 vertices (n = 25608) with the block structure."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    H = _head_mat_device(g, far_field=True)
    B = torch.from_numpy(np.random.default_rng(8).standard_normal(n * 4)).cuda()
    out = {}
    for clock in ("1", "0"):
        monkeypatch.setenv("OMH_SYRK_CLOCK", clock)
        F = factor(H, n, "mixed16", structure=g.factor_segments())
        X = B.clone()
        F.solve(X, 4, tol=1e-12)
        out[clock] = (F.L.view(n, n).t().tril().clone(), X)
        del F
    dL = float((out["1"][0] - out["0"][0]).abs().max() / out["0"][0].abs().max())
    dX = float(torch.linalg.norm(out["1"][1] - out["0"][1]) / torch.linalg.norm(out["0"][1]))
    print("n=%d slice clock vs in-order k slices: max |dL| %.3e, solutions %.3e" % (n, dL, dX), flush=True)
    assert dL < 1e-5 and dX < 1e-12


def test_split_solve_products(om):
    """The mixed refinement's triangular solves run the products of their
    full 1024-column groups as split-FP16 MFMA (k_solve_split: the FP32
    factor over a power-of-two scale per row or column, W's group rows over
    one per column, three FP16 products, FP32 sums).  One solve of the
    refinement (max_iter 1: no correction yet) against the FP32-product solve
    of the same factor (omh_sym_solve_f32), both against the refined
    solution: the split solve's error is at most 4x the FP32 solve's (plus
    1e-7), and the refined solution's residual is < 1e-12.  3 layers of 2562
    vertices (n = 17926: 17 full groups and a partial one, whose products
    stay FP32) and 200 right-hand sides (a full and a partial 128-column
    tile)."""
    import ctypes as C
    import torch
    from openmeeg_amd import spheres as S, _lib
    from openmeeg_amd.api import _head_mat_device, factor
    g = om.make_nested_geometry(S.nested_sphere_meshes(16, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    n, r = g.size, 200
    H = _head_mat_device(g, far_field=True)
    F = factor(H, n, "mixed16")
    B = torch.from_numpy(np.random.default_rng(21).standard_normal(n * r)).cuda()

    def mixed(max_iter, tol):
        X = torch.full((n * r,), np.nan, dtype=torch.float64, device="cuda")
        it, rel = C.c_int32(), C.c_double()
        rc = _lib.lib().omh_mixed_solve(C.c_void_p(H.data_ptr()), n, n, C.c_void_p(F.L.data_ptr()), n,
                                        C.c_void_p(F.work.data_ptr()), C.c_void_p(B.data_ptr()), n,
                                        C.c_void_p(X.data_ptr()), n, r, max_iter, tol, C.byref(it), C.byref(rel),
                                        C.c_void_p(0))
        assert rc in (0, _lib.STATUS_NOT_CONVERGED), _lib.lib().omh_last_error()
        return X, rc, it.value

    X1, rc1, it1 = mixed(1, 1e-12)          # one split solve, then stop
    Xr, rcr, itr = mixed(10, 1e-12)          # refined
    assert rc1 == _lib.STATUS_NOT_CONVERGED and rcr == 0
    W = B.to(torch.float32).clone()
    _lib.check(_lib.lib().omh_sym_solve_f32(C.c_void_p(F.L.data_ptr()), n, n, C.c_void_p(F.work.data_ptr()),
                                            C.c_void_p(W.data_ptr()), n, r, C.c_void_p(0)))
    torch.cuda.synchronize()
    nr = torch.linalg.norm(Xr)
    e_split = float(torch.linalg.norm(X1 - Xr) / nr)
    e_fp32 = float(torch.linalg.norm(W.double() - Xr) / nr)
    Hm = H.view(n, n)
    res = float(torch.linalg.norm(B.view(r, n) - Xr.view(r, n) @ Hm) / torch.linalg.norm(B))
    print("n=%d r=%d: one solve, error vs refined: split products %.3e, FP32 products %.3e; refined in %d "
          "corrections, residual %.3e" % (n, r, e_split, e_fp32, itr, res), flush=True)
    assert e_split <= 4 * e_fp32 + 1e-7
    assert res < 1e-12


def test_headmat_head1_matches_golden(om, head1):
    g, _, _ = head1
    hm = om.HeadMat(g)
    assert hm.nlin() == hm.ncol() == g.size == 286
    assert rel_fro(hm.data(), np.load(os.path.join(GOLDEN, "head1", "hm.npy"))) < 1e-13


def test_dsm_head1_matches_golden(om, head1):
    g, dip, _ = head1
    ref = np.load(os.path.join(GOLDEN, "head1", "dsm.npy"))
    for name in ("Brain", ""):
        dsm = om.DipSourceMat(g, dip, name)
        assert (dsm.nlin(), dsm.ncol()) == (286, 6)
        assert rel_fro(dsm.array(), ref) < 1e-13


def test_h2em_head1_matches_golden(om, head1):
    import scipy.sparse as sp
    g, _, sens = head1
    h2em = om.Head2EEGMat(g, sens)
    d, ir, jc = (np.load(os.path.join(GOLDEN, "head1", "h2em_%s.npy" % k)) for k in ("data", "ir", "jc"))
    ref = sp.csc_matrix((d, ir, jc), shape=(32, 286)).toarray()
    got = h2em.array()
    assert np.array_equal(got != 0, ref != 0)
    assert np.abs(got - ref).max() < 1e-14


def test_inverse_and_gain_head1_match_golden(om, head1):
    g, dip, sens = head1
    hm = om.HeadMat(g)
    hminv = hm.inverse()
    ref_inv = np.load(os.path.join(GOLDEN, "head1", "hm_inv.npy"))
    assert rel_fro(hminv.data(), ref_inv) < 1e-12
    dsm = om.DipSourceMat(g, dip, "Brain")
    h2em = om.Head2EEGMat(g, sens)
    ref = np.load(os.path.join(GOLDEN, "head1", "dgem.npy"))
    assert rel_fro(om.GainEEG(hminv, dsm, h2em).array(), ref) < 1e-11
    assert rel_fro(om.GainEEGadjoint(g, dip, hm, h2em).array(), ref) < 1e-11
    # SymMatrix.invert in place equals inverse (test_python.py:126-130)
    hm2 = om.HeadMat(g)
    hm2.invert()
    assert rel_fro(hm2.data(), hminv.data()) < 1e-14


def test_solvelin_head1(om, head1):
    g, dip, _ = head1
    hm = om.HeadMat(g)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((g.size, 7))
    X = hm.solveLin(om.Matrix(np.asfortranarray(B))).array()
    H = hm.array()
    assert rel_fro(H @ X, B) < 1e-12


def test_headmat_c0_matches_oracle(om, c0_oracle):
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    hm = om.HeadMat(g)
    from oracle import oracle as O
    assert rel_fro(hm.data(), O.pack(c["H"])) < 1e-13


def test_dsm_h2em_c0_match_oracle(om, c0_oracle):
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    assert rel_fro(om.DipSourceMat(g, c["dip"], "Brain").array(), c["D"]) < 1e-13
    h2em = om.Head2EEGMat(g, c["el"])
    assert np.array_equal(h2em.cols, c["cols"])
    assert np.abs(h2em.vals - c["vals"]).max() < 1e-14


def test_inverse_c0_matches_oracle(om, c0_oracle):
    """SymMatrix::inverse / invert (symmatrix.h:259-297, om_minverser) on
    Head3 against the inverse of the oracle's HeadMat (LAPACK): 1e-10
    relative Frobenius, in the reference's packed storage and order."""
    from oracle import oracle as O
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    hm = om.HeadMat(g)
    inv = hm.inverse()
    ref = np.linalg.inv(c["H"])
    e = rel_fro(inv.data(), O.pack(ref))
    hm.invert()
    e_inplace = rel_fro(hm.data(), O.pack(ref))
    print("Head3 inverse vs oracle %.3e (in place %.3e)" % (e, e_inplace), flush=True)
    assert e < 1e-10 and e_inplace < 1e-10


def test_gain_c0_fp64_and_mixed(om, c0_oracle):
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    G64, info64 = om.gain_eeg_pipeline(g, c["dip"], c["el"], precision="fp64", domain_name="Brain")
    G64 = om.gain_to_numpy(G64, len(c["el"]), len(c["dip"]))
    assert rel_fro(G64, c["G"]) < 1e-10
    for prec in ("mixed", "mixed16"):
        Gm, infom = om.gain_eeg_pipeline(g, c["dip"], c["el"], precision=prec, domain_name="Brain")
        Gm = om.gain_to_numpy(Gm, len(c["el"]), len(c["dip"]))
        err = rel_fro(Gm, G64)
        print("%s vs fp64 %.3e after %d refinements" % (prec, err, infom["refine_iterations"]), flush=True)
        assert err < 1e-6
        assert 1 <= infom["refine_iterations"] <= 10


def test_headmat_far_field_close_to_analytic(om, c0_oracle):
    """The mixed path's far-field assembly (Gauss quadrature on both
    triangles of pairs at least 4 diameters apart) against the reference's
    analytic integrals on the Head3 head: entries agree to 1e-10 of the
    matrix scale."""
    import torch
    from openmeeg_amd.api import _head_mat_device
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    n = g.size
    Ha = _head_mat_device(g).view(n, n)
    Hf = _head_mat_device(g, far_field=True).view(n, n)
    d = float(torch.linalg.norm(Hf - Ha) / torch.linalg.norm(Ha))
    m = float((Hf - Ha).abs().max() / Ha.abs().max())
    changed = float(((Hf - Ha) != 0).double().mean())
    print("far field vs analytic: rel fro %.3e, max %.3e, entries changed %.3f" % (d, m, changed), flush=True)
    assert d < 1e-10 and m < 1e-10
    assert changed > 0.5            # most pairs are far


@pytest.mark.parametrize("layers", [3, 4])
def test_far_field_fused_pass(om, monkeypatch, layers):
    """The fused far-field pass (S, D and D* of a mesh pair from one set of
    Gauss-point distances, near and 16-point pairs through a queue) against
    the per-operator far-field kernels: the same integrals summed in another
    order (1e-13 of the largest entry).  In stripes of 37 row triangles (the
    vertex gathers then add partial sums over several stripes) the same to
    1e-13.  3 layers of 642 vertices with a barrier scalp; 4 layers of 362."""
    import torch
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _head_mat_device
    if layers == 3:
        g = om.make_nested_geometry(S.nested_sphere_meshes(8, (0.87, 0.92, 1.0)), (1.0, 0.0125, 1.0))
    else:
        g = om.make_nested_geometry(S.nested_sphere_meshes(6, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    Hf = _head_mat_device(g, far_field=True).view(n, n)
    monkeypatch.setenv("OMH_FAR_SEPARATE", "1")
    Hs = _head_mat_device(g, far_field=True).view(n, n)
    monkeypatch.delenv("OMH_FAR_SEPARATE")
    monkeypatch.setenv("OMH_PAIR_STRIPE", "37")
    Hc = _head_mat_device(g, far_field=True).view(n, n)
    monkeypatch.delenv("OMH_PAIR_STRIPE")
    scale = float(Hs.abs().max())
    d_sep = float((Hf - Hs).abs().max()) / scale
    d_str = float((Hc - Hf).abs().max()) / scale
    print("n=%d fused vs separate kernels %.3e, striped vs one stripe %.3e" % (n, d_sep, d_str), flush=True)
    assert d_sep < 1e-13 and d_str < 1e-13
    assert torch.equal(Hf, Hf.t()) and torch.equal(Hc, Hc.t())


def test_gain_c0_vs_analytic_sphere(om, c0_oracle):
    """Independent of the reference: the BEM gain on the Head3 spheres
    against Zhang's series (RDM/MAG, tests/analytic tolerances for Head3:
    0.03)."""
    from oracle import analytic as A
    c = c0_oracle
    g = om.read_geometry(*c["geom_files"])
    G, _ = om.gain_eeg_pipeline(g, c["dip"], c["el"], precision="fp64", domain_name="Brain")
    G = om.gain_to_numpy(G, len(c["el"]), len(c["dip"]))
    ref = A.sphere_gain(c["el"], c["dip"], c["cond"], c["radii"])
    r = np.linalg.norm(c["dip"][:, :3], axis=1)
    shallow = np.where(r < 0.7)[0]
    assert np.median([A.rdm(G[:, j], ref[:, j]) for j in shallow]) < 0.03
    assert np.median([A.mag(G[:, j], ref[:, j]) for j in shallow]) < 0.03


def test_factor_breakdown_is_reported(om):
    """A vanishing pivot is reported (forced from the host: a zero matrix)."""
    import ctypes as C
    import torch
    from openmeeg_amd import _lib
    n = 300
    A = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    work = torch.empty(int(_lib.lib().omh_factor_workspace_bytes(n, 8)), dtype=torch.uint8, device="cuda")
    info = C.c_int32()
    _lib.check(_lib.lib().omh_sym_factor_f64(C.c_void_p(A.data_ptr()), n, n, C.c_void_p(work.data_ptr()),
                                             C.byref(info), C.c_void_p(0)))
    assert info.value == 1


@pytest.mark.parametrize("levels,tol", [(2, 1e-3), (1, 0.0)])
def test_adaptive_headmat_matches_oracle(om, head1, levels, tol):
    """HeadMat(geo, Integrator(3, levels, tol)) (assemble.h:32 with the
    adaptive recursion of integrator.h:69-93 on every S and D pair
    integral) on Head1 against the oracle's: 1e-13 relative Frobenius; it
    differs from the non-adaptive HeadMat."""
    from oracle import oracle as O
    g, _, _ = head1
    og = O.load_geometry(os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"))
    ref = O.head_mat(og, 3, levels, tol)
    hm = om.HeadMat(g, om.Integrator(3, levels, tol))
    e = rel_fro(hm.data(), ref)
    d0 = rel_fro(hm.data(), np.load(os.path.join(GOLDEN, "head1", "hm.npy")))
    print("adaptive HeadMat (levels %d, tol %g) vs oracle %.3e; vs non-adaptive %.3e" % (levels, tol, e, d0), flush=True)
    assert e < 1e-13 and d0 > 1e-9


def test_refinement_scale_cache_bitwise(om, monkeypatch):
    """The refinement's later solves read the split products' row and column
    scales the first one formed (they depend on the factor only): the
    refined solution is bitwise the one that forms them in every solve.  4
    layers of 1442 vertices (n = 14408), mesh-by-mesh factor order, block
    structure."""
    import torch
    from openmeeg_amd.api import _head_mat_device, factor
    from openmeeg_amd import spheres as S
    g = om.make_nested_geometry(S.nested_sphere_meshes(12, (0.8, 0.84, 0.9, 1.0)), (0.33, 1.79, 0.0042, 0.33))
    n = g.size
    H = _head_mat_device(g, far_field=True)
    F = factor(H, n, "mixed16", structure=g.factor_segments(), n_positive=n - g.n_vertex_unknowns,
               order=g.factor_order())
    B = torch.from_numpy(np.random.default_rng(4).standard_normal(n * 48)).cuda()
    X = {}
    for mode in ("cache", "none"):
        if mode == "none":
            monkeypatch.setenv("OMH_SOLVE_SCALE_CACHE", "0")
        X[mode] = B.clone()
        F.solve(X[mode], 48, max_iter=10, tol=1e-12)
        torch.cuda.synchronize()
        monkeypatch.delenv("OMH_SOLVE_SCALE_CACHE", raising=False)
    print("refined with cached scales vs formed each time: max |diff| %.3e after %d passes" % (
        float((X["cache"] - X["none"]).abs().max()), F.iterations), flush=True)
    assert F.iterations >= 2
    assert torch.equal(X["cache"], X["none"])
