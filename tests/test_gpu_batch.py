"""Batched gains (BatchedGain) on the GPU.

Heads of uneven size (perturbed 3-layer spheres) go through one batch; each
head's gain must equal the single-head pipeline's (FP64: 1e-12, the same
arithmetic in another launch grouping), the oracle's (1e-10), and in mixed
precision the FP64 gain (1e-6).  Graph replays must reproduce the eager pass
exactly, with other GPU work between the replays, and tuning must not change
the results (beyond the refinement tolerance when it switches between the
FP32 and the split-FP16 trailing update).
"""
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


def _heads(om, freqs, seed):
    from openmeeg_amd import spheres as S
    rng = np.random.default_rng(seed)
    out = []
    for f in freqs:
        meshes = S.nested_sphere_meshes(f, (0.87, 0.92, 1.0), rng=rng, amplitude=0.03)
        g = om.make_nested_geometry(meshes, (1.0, 0.0125, 1.0))
        dip = S.random_dipoles(rng, int(rng.integers(20, 60)), 0.78)
        el = S.sphere_electrodes(rng, int(rng.integers(16, 40)))
        out.append((g, dip, el, meshes))
    return out


@pytest.fixture(scope="module")
def heads(om):
    return _heads(om, (3, 5, 4), 21)


@pytest.fixture(scope="module")
def fp64_batch(om, heads):
    b = om.BatchedGain([h[:3] for h in heads], precision="fp64")
    assert not b.run(graph=True)            # the first pass is eager
    return b, [b.gain(i) for i in range(len(heads))]


def test_batch_fp64_matches_single_head_pipeline(om, heads, fp64_batch):
    b, G = fp64_batch
    for i, (g, dip, el, _) in enumerate(heads):
        Gs, _ = om.gain_eeg_pipeline(g, dip, el, precision="fp64")
        Gs = om.gain_to_numpy(Gs, len(el), len(dip))
        err = rel_fro(G[i], Gs)
        print("head %d n=%d: batch vs pipeline %.3e" % (i, g.size, err), flush=True)
        assert G[i].shape == (len(el), len(dip))
        assert err < 1e-12
        assert b.stats(i)["factor_info"] == 0


def test_batch_fp64_matches_oracle(heads, fp64_batch):
    from oracle import oracle as O
    _, G = fp64_batch
    g, dip, el, meshes = heads[0]
    og = O.nested_geometry(meshes, (1.0, 0.0125, 1.0))
    H = O.unpack(O.head_mat(og), og.size)
    cols, vals = O.head2eeg(og, el)
    Gref = O.gain_eeg(H, O.dip_source_mat(og, dip), O.h2em_dense(og, cols, vals))
    err = rel_fro(G[0], Gref)
    print("head 0 vs oracle %.3e" % err, flush=True)
    assert err < 1e-10


def test_batch_mixed_graph_replays(om, heads, fp64_batch):
    import torch
    _, G64 = fp64_batch
    b = om.BatchedGain([h[:3] for h in heads], precision="mixed")
    assert not b.run(graph=True)
    eager = [b.gain(i) for i in range(len(heads))]
    cfg = b.config()
    print("mixed batch config before capture", cfg, flush=True)
    assert 1 <= cfg["refine_steps"] <= 10
    for i in range(len(heads)):
        # refinement stops at an estimated error of 1e-8 (last update times
        # its contraction ratio, as the single-head gain pipelines): the gain
        # is well inside the 1e-6 bar
        assert rel_fro(eager[i], G64[i]) < 1e-7
        st = b.stats(i)
        # the last correction's relative size: the refinement stopped on its
        # contraction estimate, (rel / prev) rel <= 1e-8
        assert st["factor_info"] == 0 and st["rel_update"] <= 1e-4
    for rep in range(3):
        x = torch.randn(512, 512, device="cuda", dtype=torch.float64)
        y = (x @ x).sum().item()                 # unrelated work between replays
        assert np.isfinite(y)
        assert b.run(graph=True)
        for i in range(len(heads)):
            np.testing.assert_array_equal(b.gain(i), eager[i])
    cfg = b.config()
    print("after capture", cfg, flush=True)
    assert cfg["graph_nodes"] > cfg["grouped_launches"]


def test_batch_tune_keeps_results(om, heads, fp64_batch):
    _, G64 = fp64_batch
    b = om.BatchedGain([h[:3] for h in heads], precision="mixed")
    b.run(graph=False)
    before = [b.gain(i) for i in range(len(heads))]
    split_before = b.config()["split16"]
    res = b.tune()
    print("tune", res, b.config(), flush=True)
    assert res and all(t > 0 for _, _, _, t in res)
    assert {r[0] for r in res if not r[2]} == {64, 128}
    assert any(r[2] for r in res)                # the split-FP16 candidate
    assert not b.run(graph=True)                 # eager again after tuning
    assert b.run(graph=True)
    same = b.config()["split16"] == split_before
    for i in range(len(heads)):
        # the same update kernel: the same sums; the other one: both refined
        assert rel_fro(b.gain(i), before[i]) < (1e-12 if same else 1e-7)
        assert rel_fro(b.gain(i), G64[i]) < 1e-6


def test_batch_uneven_heads(om):
    """Six heads of the batched-heads generator (642 to 1442 vertices per
    surface, 200 to 2000 dipoles, 64 to 256 electrodes) against the
    single-head mixed pipeline."""
    from openmeeg_amd import spheres as S
    hs = S.make_batched_heads(n_heads=6, seed=9, freqs=(8, 10, 12))
    heads = [(om.make_nested_geometry(m, c), d, e) for m, c, d, e in hs]
    b = om.BatchedGain(heads, precision="mixed")
    b.run(graph=False)
    assert b.run(graph=True)
    for i, (g, d, e) in enumerate(heads):
        Gs, _ = om.gain_eeg_pipeline(g, d, e, precision="fp64")
        err = rel_fro(b.gain(i), om.gain_to_numpy(Gs, len(e), len(d)))
        print("head %d n=%d n_dip=%d n_el=%d: %.3e" % (i, g.size, len(d), len(e), err), flush=True)
        assert err < 1e-6
