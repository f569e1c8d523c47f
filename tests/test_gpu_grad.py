"""Gradients of the EEG gain (EEGForward) on the GPU.

The forward gain is checked against the oracle (same non-adaptive source
rule, 1e-10).  The adjoint gradients of a fixed linear functional
L(G) = sum(R * G) with respect to dipole positions, moments and domain
conductivities are checked against central finite differences of the GPU
forward: moments exactly (G is linear in them, 1e-9), positions and
conductivities with steps of 1e-5 (relative), where the truncation error is
O(h^2) and the tolerance 1e-6 relative to the gradient norm.  Mixed and FP64
gradients agree to 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAYERS = ((0.87, 0.92, 1.0), (1.0, 0.0125, 1.0))


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd as om
    from openmeeg_amd import spheres as S
    meshes = S.nested_sphere_meshes(4, LAYERS[0])
    rng = np.random.default_rng(11)
    dip = S.random_dipoles(rng, 6, 0.78)
    el = S.sphere_electrodes(rng, 16)
    g = om.make_nested_geometry(meshes, LAYERS[1])
    R = rng.standard_normal((len(el), len(dip)))
    return om, meshes, dip, el, g, R


def _loss_and_grads(model, dip, R, sigma=None):
    import torch
    pos = torch.tensor(dip[:, :3], requires_grad=True)
    mom = torch.tensor(dip[:, 3:], requires_grad=True)
    sig = torch.tensor(model.conductivities() if sigma is None else sigma, requires_grad=True)
    G = model.gain(pos, mom, sig)
    L = (G.cpu() * torch.tensor(R)).sum()
    L.backward()
    return float(L.detach()), pos.grad.numpy().copy(), mom.grad.numpy().copy(), sig.grad.numpy().copy(), G


def _loss(model, dip, R, sigma):
    import torch
    with torch.no_grad():
        G = model.gain(torch.tensor(dip[:, :3]), torch.tensor(dip[:, 3:]), torch.tensor(sigma))
    return float((G.cpu().numpy() * R).sum())


def test_forward_matches_oracle(setup):
    om, meshes, dip, el, g, R = setup
    from oracle import oracle as O
    model = om.EEGForward(g, el, precision="fp64")
    _, _, _, _, G = _loss_and_grads(model, dip, R)
    og = O.nested_geometry(meshes, LAYERS[1])
    H = O.unpack(O.head_mat(og), og.size)
    cols, vals = O.head2eeg(og, el)
    Gref = O.gain_eeg(H, O.dip_source_mat(og, dip, "", levels=0), O.h2em_dense(og, cols, vals))
    G = G.detach().cpu().numpy()
    err = np.linalg.norm(G - Gref) / np.linalg.norm(Gref)
    print("forward gain vs oracle %.3e" % err, flush=True)
    assert err < 1e-10


def test_gradients_match_finite_differences(setup):
    om, meshes, dip, el, g, R = setup
    model = om.EEGForward(g, el, precision="fp64")
    sig0 = model.conductivities()
    L0, gp, gm, gs, _ = _loss_and_grads(model, dip, R)
    h = 1e-5
    fp = np.zeros_like(gp)
    for s in range(len(dip)):
        for c in range(3):
            d1, d2 = dip.copy(), dip.copy()
            d1[s, c] += h
            d2[s, c] -= h
            fp[s, c] = (_loss(model, d1, R, sig0) - _loss(model, d2, R, sig0)) / (2 * h)
    fm = np.zeros_like(gm)
    for s in range(len(dip)):
        for c in range(3):
            d1, d2 = dip.copy(), dip.copy()
            d1[s, 3 + c] += 1.0
            d2[s, 3 + c] -= 1.0
            fm[s, c] = (_loss(model, d1, R, sig0) - _loss(model, d2, R, sig0)) / 2.0
    fs = np.zeros_like(gs)
    for d in range(len(sig0)):
        if sig0[d] == 0:
            continue
        hs = h * sig0[d]
        s1, s2 = sig0.copy(), sig0.copy()
        s1[d] += hs
        s2[d] -= hs
        fs[d] = (_loss(model, dip, R, s1) - _loss(model, dip, R, s2)) / (2 * hs)
    ep = np.linalg.norm(gp - fp) / np.linalg.norm(fp)
    em = np.linalg.norm(gm - fm) / np.linalg.norm(fm)
    es = np.abs(gs - fs) / np.abs(fs).max()
    print("position %.3e moment %.3e conductivity %s" % (ep, em, es), flush=True)
    assert em < 1e-9
    assert ep < 1e-6
    assert es.max() < 1e-6
    assert gs[list(sig0).index(0.0)] == 0.0


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
def test_mixed_gradients_match_fp64(setup, precision):
    om, meshes, dip, el, g, R = setup
    a = _loss_and_grads(om.EEGForward(g, el, precision="fp64"), dip, R)
    b = _loss_and_grads(om.EEGForward(g, el, precision=precision), dip, R)
    for x, y in zip(a[1:4], b[1:4]):
        assert np.linalg.norm(x - y) / np.linalg.norm(x) < 1e-6


def test_moved_conductivity_gradient(setup):
    """Gradients at conductivities away from the geometry's initial ones
    (the model changes the geometry's conductivities and its coefficients)."""
    om, meshes, dip, el, g, R = setup
    model = om.EEGForward(g, el, precision="fp64")
    sig = model.conductivities()
    live = sig > 0
    sig[live] *= np.array([1.3, 0.7, 1.1])[: live.sum()]
    _, _, _, gs, _ = _loss_and_grads(model, dip, R, sig)
    h = 1e-5
    for d in np.flatnonzero(live):
        s1, s2 = sig.copy(), sig.copy()
        s1[d] += h * sig[d]
        s2[d] -= h * sig[d]
        fd = (_loss(model, dip, R, s1) - _loss(model, dip, R, s2)) / (2 * h * sig[d])
        assert abs(gs[d] - fd) < 1e-6 * np.abs(gs).max()


def test_c3_fit_gradient_matches_fd(setup):
    """The 128-electrode, 4096-dipole fit on the 2562-vertex 3-layer head:
    L = 0.5 |G(p, q, sigma) - T|^2 with targets T at SNR 10, from a start
    perturbed off the true dipoles and conductivities.  Central differences
    with h = 1e-6 (relative for conductivities) on sampled parameters agree
    with the adjoint gradient to 1e-5; the mixed-precision gradient agrees
    with the FP64 one to 1e-6."""
    om = setup[0]
    import torch
    from openmeeg_amd import spheres as S
    meshes, cond, dip, el = S.make_config("c3")
    g = om.make_nested_geometry(meshes, cond)
    rng = np.random.default_rng(33)
    model = om.EEGForward(g, el, precision="fp64")
    sig_true = model.conductivities()
    with torch.no_grad():
        Gt = model.gain(torch.tensor(dip[:, :3]), torch.tensor(dip[:, 3:]), torch.tensor(sig_true))
    T = torch.tensor(S.noisy_targets(Gt.cpu().numpy(), rng, 10.0))
    p0 = dip[:, :3] + 0.01 * rng.standard_normal((len(dip), 3))
    r = np.linalg.norm(p0, axis=1, keepdims=True)
    p0 = np.where(r > 0.82, p0 * 0.82 / r, p0)
    q0 = dip[:, 3:] * 1.1
    s0 = sig_true.copy()
    s0[s0 > 0] *= np.array([1.05, 0.95, 1.02])

    def loss(p, q, s):
        G = model.gain(p, q, s)
        return 0.5 * ((G.cpu() - T) ** 2).sum()

    def grads(m):
        model_ = m
        p = torch.tensor(p0, requires_grad=True)
        q = torch.tensor(q0, requires_grad=True)
        s = torch.tensor(s0, requires_grad=True)
        G = model_.gain(p, q, s)
        (0.5 * ((G.cpu() - T) ** 2).sum()).backward()
        return p.grad.numpy(), q.grad.numpy(), s.grad.numpy()

    gp, gq, gs = grads(model)
    h = 1e-6
    idx = rng.choice(len(dip), 3, replace=False)

    def fd(kind, i, c):
        a = [p0.copy(), q0.copy(), s0.copy()]
        b = [p0.copy(), q0.copy(), s0.copy()]
        step = h * (s0[i] if kind == 2 else 1.0)
        if kind == 2:
            a[2][i] += step
            b[2][i] -= step
        else:
            a[kind][i, c] += step
            b[kind][i, c] -= step
        with torch.no_grad():
            la = loss(*[torch.tensor(x) for x in a])
            lb = loss(*[torch.tensor(x) for x in b])
        return float(la - lb) / (2 * step)

    checks = [(0, i, int(rng.integers(3)), gp) for i in idx] + [(1, i, int(rng.integers(3)), gq) for i in idx[:2]]
    for kind, i, c, gr in checks:
        f = fd(kind, i, c)
        print("param %d dipole %d comp %d: adjoint %.10e fd %.10e" % (kind, i, c, gr[i, c], f), flush=True)
        assert abs(gr[i, c] - f) <= 1e-5 * abs(f) + 1e-12 * np.abs(gr).max()
    for d in np.flatnonzero(s0 > 0):
        f = fd(2, d, 0)
        print("conductivity %d: adjoint %.10e fd %.10e" % (d, gs[d], f), flush=True)
        assert abs(gs[d] - f) <= 1e-5 * abs(f)
    mp, mq, ms = grads(om.EEGForward(g, el, precision="mixed"))
    for x, y in ((gp, mp), (gq, mq), (gs, ms)):
        assert np.linalg.norm(x - y) / np.linalg.norm(x) < 1e-6
