"""The benchmark configurations at their full sizes on the GPU.

* The 4-layer 10242-vertex head (radii 0.80/0.84/0.90/1.0, conductivities
  0.33/1.79/0.0042/0.33, n = 102408) with 5000 dipoles and 256 electrodes:
  the mixed-precision gains (split-FP16 and FP32 factors with FP64
  refinement) within 1e-6 of the plain FP64 gain (relative Frobenius), and
  the FP64 factor's solve residual ||B - H X|| / ||B|| below 1e-12.
* The 2562-vertex three-sphere head (n = 17926) with 2000 grid dipoles and
  128 electrodes: the FP64 gain within 1e-10 of the oracle's, written into
  tests/golden/c1 by tests/golden/make_oracle_gains.py.
* The 32 perturbed heads (642 to 2562 vertices per surface, 200-2000
  dipoles, 64-256 electrodes) in one batch: the FP64 batch against the
  single-head FP64 pipeline (1e-12) and, for the largest head, against the
  oracle's gain of 200 of its dipoles (1e-10, tests/golden/c4); the mixed
  batch's graph replays bitwise equal to its eager pass and within 1e-6 of
  FP64.
Needs a HIP GPU (about 200 GB of device memory for the first test).
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
def c2(om):
    from openmeeg_amd import spheres as S
    meshes, cond, dip, el = S.make_config("c2")
    return om.make_nested_geometry(meshes, cond), dip, el


def test_c2_mixed_gains_within_1e6_of_fp64(om, c2):
    import torch
    g, dip, el = c2
    ne, nd = len(el), len(dip)
    print("c2: n=%d, %d dipoles, %d electrodes: fp64 gain" % (g.size, nd, ne), flush=True)
    G64, i64 = om.gain_eeg_pipeline(g, dip, el, precision="fp64", domain_name="Brain")
    G64 = om.gain_to_numpy(G64, ne, nd)
    errs = {}
    for prec in ("mixed16", "mixed"):
        print("c2: %s gain" % prec, flush=True)
        G, info = om.gain_eeg_pipeline(g, dip, el, precision=prec, domain_name="Brain")
        errs[prec] = (rel_fro(om.gain_to_numpy(G, ne, nd), G64), info["refine_iterations"],
                      info["refine_fell_back_fp64"], sum(info["times"].values()))
        del G
        torch.cuda.empty_cache()
    print("c2 vs fp64 (%.2f s): %s" % (sum(i64["times"].values()),
                                       {k: "%.2e after %d solves, fell back %s, %.2f s" % v for k, v in errs.items()}),
          flush=True)
    assert np.isfinite(G64).all() and np.abs(G64).max() > 0
    for prec, (err, its, fell_back, _) in errs.items():
        assert err < 1e-6, prec
        assert not fell_back and 1 <= its <= 10, prec


def test_c2_fp64_factor_residual(om, c2):
    import torch
    from openmeeg_amd.api import _head_mat_device, factor
    g, _, _ = c2
    n = g.size
    print("c2: FP64 HeadMat, factor and solve of 8 right-hand sides, n=%d" % n, flush=True)
    H = _head_mat_device(g)
    Href = H.clone().view(n, n)
    B = torch.from_numpy(np.random.default_rng(6).standard_normal(n * 8)).cuda()
    F = factor(H, n, "fp64", structure=g.factor_segments())
    X = B.clone()
    F.solve(X, 8)
    R = B.view(8, n) - X.view(8, n) @ Href            # H symmetric: rows of X^T H
    res = float(torch.linalg.norm(R) / torch.linalg.norm(B))
    print("c2 FP64 residual %.3e" % res, flush=True)
    assert res < 1e-12
    del F, H, Href, X, R
    torch.cuda.empty_cache()


def test_c1_fp64_gain_matches_oracle_fixture(om):
    from openmeeg_amd import spheres as S
    d = os.path.join(GOLDEN, "c1")
    meshes, cond, dip, el = S.make_config("c1")
    # the fixture was made from the same generated inputs
    assert np.array_equal(dip, np.load(os.path.join(d, "dipoles.npy")))
    assert np.array_equal(el, np.load(os.path.join(d, "electrodes.npy")))
    chk = np.array([[np.asarray(p).sum(), float(len(t))] for p, t in meshes])
    np.testing.assert_allclose(chk, np.load(os.path.join(d, "mesh_check.npy")), rtol=1e-14)
    ref = np.load(os.path.join(d, "gain.npy"))
    g = om.make_nested_geometry(meshes, cond)
    print("c1: n=%d fp64 gain vs the oracle's" % g.size, flush=True)
    G, _ = om.gain_eeg_pipeline(g, dip, el, precision="fp64", domain_name="Brain")
    G = om.gain_to_numpy(G, len(el), len(dip))
    err = rel_fro(G, ref)
    Gm, info = om.gain_eeg_pipeline(g, dip, el, precision="mixed16", domain_name="Brain")
    errm = rel_fro(om.gain_to_numpy(Gm, len(el), len(dip)), ref)
    print("c1 gain vs oracle: fp64 %.3e, mixed16 %.3e (%d solves)" % (err, errm, info["refine_iterations"]),
          flush=True)
    assert G.shape == ref.shape == (128, 2000)
    assert err < 1e-10 and errm < 1e-6


@pytest.fixture(scope="module")
def c4_heads(om):
    from openmeeg_amd import spheres as S
    heads = S.make_batched_heads(32, seed=4)
    return heads, [(om.make_nested_geometry(m, c), dip, el) for (m, c, dip, el) in heads]


def test_c4_batch_of_32_heads(om, c4_heads):
    raw, heads = c4_heads
    sizes = [g.size for g, _, _ in heads]
    print("c4: 32 heads, n %d..%d, FP64 batch" % (min(sizes), max(sizes)), flush=True)
    b64 = om.BatchedGain(heads, precision="fp64")
    b64.run(graph=False)
    G64 = [b64.gain(i) for i in range(len(heads))]
    worst_single = 0.0
    for i, (g, dip, el) in enumerate(heads):
        Gs, _ = om.gain_eeg_pipeline(g, dip, el, precision="fp64")
        worst_single = max(worst_single, rel_fro(G64[i], om.gain_to_numpy(Gs, len(el), len(dip))))
    # the largest head against the oracle's gain of 200 of its dipoles
    d = os.path.join(GOLDEN, "c4")
    h = int(np.load(os.path.join(d, "head.npy"))[0])
    cols = np.load(os.path.join(d, "columns.npy"))
    chk = np.array([[np.asarray(p).sum(), float(len(t))] for p, t in raw[h][0]])
    np.testing.assert_allclose(chk, np.load(os.path.join(d, "mesh_check.npy")), rtol=1e-14)
    assert sizes[h] == max(sizes)
    e_oracle = rel_fro(G64[h][:, cols], np.load(os.path.join(d, "gain_columns.npy")))
    print("c4 FP64 batch: vs single-head pipeline %.3e (worst head), head %d (n=%d) vs oracle %.3e"
          % (worst_single, h, sizes[h], e_oracle), flush=True)
    del b64
    print("c4: mixed batch, eager pass then graph replays", flush=True)
    bm = om.BatchedGain(heads, precision="mixed")
    assert not bm.run(graph=True)
    eager = [bm.gain(i) for i in range(len(heads))]
    replays_equal = True
    for _ in range(2):
        assert bm.run(graph=True)
        replays_equal &= all(np.array_equal(bm.gain(i), eager[i]) for i in range(len(heads)))
    worst_mixed = max(rel_fro(eager[i], G64[i]) for i in range(len(heads)))
    print("c4 mixed batch: graph replays equal eager %s, worst head vs FP64 %.3e, config %s"
          % (replays_equal, worst_mixed, bm.config()), flush=True)
    assert worst_single < 1e-12 and e_oracle < 1e-10
    assert replays_equal and worst_mixed < 1e-6
