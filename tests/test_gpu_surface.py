"""Surface sources on the HIP path: SurfSourceMat, SurfSource2MEGMat, the
surface EEG / MEG gains and the overlap check, against the reference's
Head1 goldens and the CPU oracle.  Needs a HIP GPU.

Tolerances (relative Frobenius): SurfSourceMat entries vs goldens or the
oracle 1e-13 (FP64, a few ulp per entry); SurfSource2MEGMat 1e-10 (the
oracle's own agreement with the golden); gains 1e-10; mixed precision vs
FP64 1e-6."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu

H1 = os.path.join(GOLDEN, "head1")


@pytest.fixture(scope="module")
def om():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd
    return openmeeg_amd


def _golden(name):
    return np.load(os.path.join(H1, name + ".npy"))


def test_surf_sources_head1_match_golden(om):
    """Head1 with the source mesh Head1.tri: SurfSourceMat, SurfSource2MEGMat,
    GainEEG / GainMEG from the HeadMat inverse, the device pipeline (one
    factor, one solve with the EEG and MEG rows) and the EEG forward of
    Head1.src against the reference's goldens."""
    g = om.read_geometry(os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond"))
    mesh = om.Mesh(os.path.join(H1, "Head1.tri"))
    print("Head1 surface sources: SurfSourceMat, SurfSource2MEGMat, gains", flush=True)
    ssm = om.SurfSourceMat(g, mesh)
    e_ssm = rel_fro(ssm.array(), _golden("ssm"))
    sens = om.Sensors()
    sens.load(os.path.join(H1, "Head1.patches"))
    sq = om.Sensors()
    sq.load(os.path.join(H1, "Head1.squids"))
    ss2mm = om.SurfSource2MEGMat(mesh, sq)
    e_s2m = rel_fro(ss2mm.array(), _golden("ss2mm"))
    hm = om.HeadMat(g)
    hminv = hm.inverse()
    h2em = om.Head2EEGMat(g, sens)
    h2mm = om.Head2MEGMat(g, sq)
    e_sgem = rel_fro(om.GainEEG(hminv, ssm, h2em).array(), _golden("sgem"))
    e_sgmm = rel_fro(om.GainMEG(hminv, ssm, h2mm, ss2mm).array(), _golden("sgmm"))
    Ge, Gm, info = om.gain_eeg_surf_pipeline(g, mesh, sens, sq, precision="fp64")
    Ge = om.gain_to_numpy(Ge, 32, 42)
    Gm = om.gain_to_numpy(Gm, sq.getNumberOfSensors(), 42)
    e_pe = rel_fro(Ge, _golden("sgem"))
    e_pm = rel_fro(Gm, _golden("sgmm"))
    src = np.loadtxt(os.path.join(H1, "Head1.src"))
    e_est = rel_fro(Ge @ src, _golden("est_eeg"))
    print("  ssm %.3e ss2mm %.3e sgem %.3e sgmm %.3e pipeline EEG %.3e MEG %.3e est_eeg %.3e"
          % (e_ssm, e_s2m, e_sgem, e_sgmm, e_pe, e_pm, e_est), flush=True)
    assert e_ssm < 1e-13 and e_s2m < 1e-10
    assert e_sgem < 1e-10 and e_sgmm < 1e-10 and e_pe < 1e-10 and e_pm < 1e-10 and e_est < 1e-10


@pytest.fixture(scope="module")
def spheres3(om):
    """3-layer sphere head (162 vertices per surface) and source spheres in
    each conducting layer: the brain (one inside boundary), the skull (an
    outside and an inside boundary) and the scalp (its inside boundary is the
    outermost mesh, a current barrier: no D block)."""
    from oracle import oracle as O
    from openmeeg_amd import spheres as S
    meshes = S.nested_sphere_meshes(4, (0.87, 0.92, 1.0))
    cond = (1.0, 0.0125, 1.0)
    P, T = S.geodesic_sphere(3)
    sources = {"brain": (0.6 * P, T), "skull": (0.90 * P, T), "scalp": (0.96 * P, T)}
    return om.make_nested_geometry(meshes, cond), O.nested_geometry(meshes, cond), sources


@pytest.mark.parametrize("layer", ["brain", "skull", "scalp"])
def test_surf_source_layers_vs_oracle(om, spheres3, layer):
    """SurfSourceMat of a source sphere in each layer against the oracle, in
    the reference unknown order (SurfSourceMat) and the triangles-first
    order (the device pipeline's); the far-field quadrature of well-separated
    pairs within 1e-9 of the exact integrals."""
    import torch
    from oracle import oracle as O
    from openmeeg_amd.api import _ssm_device
    g, og, sources = spheres3
    P, T = sources[layer]
    print("surface source in the %s layer: %d vertices" % (layer, len(P)), flush=True)
    ref = O.surf_source_mat(og, P, T)
    mesh = om.Mesh(P, T)
    got = om.SurfSourceMat(g, mesh).array()
    e_ref = rel_fro(got, ref)
    n, nv = g.size, g.n_vertex_unknowns
    perm = np.r_[np.arange(nv, n), np.arange(nv)]
    D = _ssm_device(g, mesh).view(len(P), n).t().cpu().numpy()
    e_tf = rel_fro(D, ref[perm])
    Df = _ssm_device(g, mesh, far_field=True).view(len(P), n).t().cpu().numpy()
    e_far = rel_fro(Df, ref[perm])
    torch.cuda.synchronize()
    print("  vs oracle %.3e, triangles-first %.3e, far field %.3e" % (e_ref, e_tf, e_far), flush=True)
    assert e_ref < 1e-13 and e_tf < 1e-13 and e_far < 1e-9


def test_surf_source_overlap_is_refused(om, spheres3):
    """A source sphere that crosses the inner interface is refused with
    OverlappingSourceMesh (Geometry::check); one inside the brain is not;
    a source mesh touching the interface (a copy of it) is found too."""
    from openmeeg_amd import spheres as S
    g, _, sources = spheres3
    P, T = S.geodesic_sphere(3)
    print("overlap checks", flush=True)
    crossing = om.Mesh(0.7 * P + np.array([0.3, 0.0, 0.0]), T)
    assert om.mesh_intersects(g, crossing)
    with pytest.raises(om.OverlappingSourceMesh):
        om.SurfSourceMat(g, crossing)
    assert not om.mesh_intersects(g, om.Mesh(*sources["brain"]))
    inner = S.nested_sphere_meshes(4, (0.87,))[0]
    assert om.mesh_intersects(g, om.Mesh(*inner))


def test_surf_gain_pipeline_mixed_and_dipole_limit(om, spheres3):
    """The surface-source pipeline in mixed precision within 1e-6 of FP64
    (EEG and MEG); the FP64 surface EEG gain against the gain of the
    equivalent normal dipole layer (-n dA per unit density, three dipoles
    per source triangle) within the quadrature error of that layer (3%)."""
    from openmeeg_amd import spheres as S
    g, _, sources = spheres3
    P, T = sources["brain"]
    mesh = om.Mesh(P, T)
    rng = np.random.default_rng(3)
    el = S.sphere_electrodes(rng, 40)
    p, o, w = S.sphere_squids(rng, 30)
    sq = om.Sensors(p, orientations=o, weights=w)
    ns = len(P)
    print("surface pipeline fp64 / mixed16", flush=True)
    Ge, Gm, _ = om.gain_eeg_surf_pipeline(g, mesh, el, sq, precision="fp64")
    Ge, Gm = om.gain_to_numpy(Ge, 40, ns), om.gain_to_numpy(Gm, 30, ns)
    Ge16, Gm16, info = om.gain_eeg_surf_pipeline(g, mesh, el, sq, precision="mixed16")
    e_e = rel_fro(om.gain_to_numpy(Ge16, 40, ns), Ge)
    e_m = rel_fro(om.gain_to_numpy(Gm16, 30, ns), Gm)
    dips, W = [], np.zeros((3 * len(T), ns))
    for i, t in enumerate(T):
        nrm = np.cross(P[t[1]] - P[t[0]], P[t[2]] - P[t[0]])
        area = 0.5 * np.linalg.norm(nrm)
        for k in range(3):
            lam = np.full(3, 1.0 / 6.0)
            lam[k] = 2.0 / 3.0
            dips.append(np.r_[lam @ P[t], -nrm / np.linalg.norm(nrm)])
            W[3 * i + k, t] = lam * area / 3.0
    Gd, _ = om.gain_eeg_pipeline(g, np.array(dips), el, precision="fp64", domain_name="Brain")
    e_dip = rel_fro(Ge, om.gain_to_numpy(Gd, 40, len(dips)) @ W)
    print("  mixed16 vs fp64: EEG %.3e MEG %.3e (%d refinements); vs dipole layer %.3e"
          % (e_e, e_m, info["refine_iterations"], e_dip), flush=True)
    assert e_e < 1e-6 and e_m < 1e-6 and e_dip < 0.03
