"""Internal potentials and EIT on the HIP path: Surf2VolMat,
DipSource2InternalPotMat, GainInternalPot and the device pipeline, against
the CPU oracle (Head1 with the reference's internal points, a 3-layer sphere
head with points in every layer) and the exact reproduction of linear
fields; EITSourceMat against the oracle and Helmholtz reciprocity; the
command-line options -H2IPM / -DS2IPM / -EITSM and om_gain -IP / -EITIP.
Needs a HIP GPU.

Tolerances (relative Frobenius): matrices vs the oracle 1e-13, gains 1e-10,
mixed precision vs FP64 1e-6."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu

H1 = os.path.join(GOLDEN, "head1")
POINTS = os.path.join(GOLDEN, "analytic", "eeg_internal_points.txt")


@pytest.fixture(scope="module")
def om():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd
    return openmeeg_amd


def test_internal_potential_head1_vs_oracle(om):
    from oracle import oracle as O
    gf, cf = os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond")
    g, og = om.read_geometry(gf, cf), O.load_geometry(gf, cf)
    pts = np.loadtxt(POINTS, usecols=(0, 1, 2))
    dip = O.read_dipoles(os.path.join(H1, "Head1.dip"))
    print("Head1 internal potentials: %d points, %d dipoles" % (len(pts), len(dip)), flush=True)
    Mref, kept = O.surf2vol(og, pts)
    Sref = O.dip_source2internal_pot(og, dip, pts)
    M = om.Surf2VolMat(g, pts).array()
    S2 = om.DipSource2InternalPotMat(g, dip, pts).array()
    e_m, e_s = rel_fro(M, Mref), rel_fro(S2, Sref)
    H = O.unpack(np.load(os.path.join(H1, "hm.npy")), g.size)
    Gref = O.gain_internal_pot(H, np.load(os.path.join(H1, "dsm.npy")), Mref, Sref)
    hm = om.HeadMat(g)
    dsm = om.DipSourceMat(g, dip, "Brain")
    e_g = rel_fro(om.GainInternalPot(hm.inverse(), dsm, om.Matrix(M), om.Matrix(S2)).array(), Gref)
    Gp, kp, _ = om.gain_internal_pot_pipeline(g, dip, pts, precision="fp64", domain_name="Brain")
    e_p = rel_fro(om.gain_to_numpy(Gp, len(kp), len(dip)), Gref)
    Gm, _, info = om.gain_internal_pot_pipeline(g, dip, pts, precision="mixed16", domain_name="Brain")
    e_x = rel_fro(om.gain_to_numpy(Gm, len(kp), len(dip)), om.gain_to_numpy(Gp, len(kp), len(dip)))
    print("  h2ip %.3e ds2ip %.3e GainInternalPot %.3e pipeline %.3e mixed16 vs fp64 %.3e"
          % (e_m, e_s, e_g, e_p, e_x), flush=True)
    assert list(kp) == list(kept)
    assert e_m < 1e-13 and e_s < 1e-14 and e_g < 1e-10 and e_p < 1e-10 and e_x < 1e-6


def test_internal_potential_layers_and_linear_fields(om):
    """Points in every layer of a 3-layer sphere head (and outside it, where
    they are dropped): Surf2VolMat in both unknown orders and
    DipSource2InternalPotMat against the oracle; on the device, Surf2VolMat
    reproduces a linear potential from its boundary values to rounding in
    the inner domain and in a shell with sigma = 2."""
    import torch
    from oracle import oracle as O
    from openmeeg_amd import spheres as S
    from openmeeg_amd.api import _surf2vol_device
    meshes = S.nested_sphere_meshes(4, (0.5, 0.8, 1.0))
    cond = (1.0, 2.0, 1.0)
    g, og = om.make_nested_geometry(meshes, cond), O.nested_geometry(meshes, cond)
    rng = np.random.default_rng(2)
    d = rng.standard_normal((40, 3))
    pts = d * (rng.random(40) * 1.2 / np.linalg.norm(d, axis=1))[:, None]
    dip = S.random_dipoles(rng, 5, 0.4)
    print("sphere head: %d points, %d dipoles" % (len(pts), len(dip)), flush=True)
    Mref, kept = O.surf2vol(og, pts)
    assert 0 < len(kept) < len(pts)
    e_ref = rel_fro(om.Surf2VolMat(g, pts).array(), Mref)
    n, nv = g.size, g.n_vertex_unknowns
    perm = np.r_[np.arange(nv, n), np.arange(nv)]
    Md, kd = _surf2vol_device(g, pts)
    A = Md.cpu().numpy().reshape(n, len(pts)).T[:len(kd)]
    e_tf = rel_fro(A, Mref[:, perm])
    e_s = rel_fro(om.DipSource2InternalPotMat(g, dip, pts).array(), O.dip_source2internal_pot(og, dip, pts))
    a = np.array([0.3, -0.5, 0.8])
    verts = np.array(og.vertices)
    errs = []
    for rmin, rmax, ms, sigma in ((0.05, 0.45, (0,), 1.0), (0.55, 0.72, (0, 1), 2.0)):
        q = rng.standard_normal((16, 3))
        q *= ((rmin + (rmax - rmin) * rng.random(16)) / np.linalg.norm(q, axis=1))[:, None]
        x = np.zeros(n)
        for mi in ms:
            m = og.meshes[mi]
            for v in m.verts:
                x[og.vert_index[v]] = verts[v] @ a
            for kt, t in enumerate(m.tris):
                P = verts[t]
                nrm = np.cross(P[1] - P[0], P[2] - P[0])
                nrm /= np.linalg.norm(nrm)
                x[m.tri_index[kt]] = sigma * (nrm @ a) * (1 if nrm @ P.mean(axis=0) > 0 else -1)
        u = om.Surf2VolMat(g, q).array() @ x
        errs.append(np.abs(u - q @ a).max())
    torch.cuda.synchronize()
    print("  vs oracle %.3e, triangles-first %.3e, ds2ip %.3e, linear fields %s" % (e_ref, e_tf, e_s, errs),
          flush=True)
    assert e_ref < 1e-13 and e_tf < 1e-13 and e_s < 1e-14 and max(errs) < 1e-13


def test_internal_potential_cli(om, tmp_path):
    """om_assemble -H2IPM / -DS2IPM and om_gain -IP on Head1 give the
    matrices of the Python surface."""
    from openmeeg_amd import io as omio
    BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openmeeg_amd", "bin")
    gf, cf, df = os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond"), os.path.join(H1, "Head1.dip")
    t = lambda f: str(tmp_path / f)

    def run(tool, *args):
        print("run", tool, args[0], flush=True)
        r = subprocess.run([os.path.join(BIN, tool)] + list(args), capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr

    run("om_assemble", "-H2IPM", gf, cf, POINTS, t("h2ip.mat"))
    run("om_assemble", "-DS2IPM", gf, cf, df, POINTS, t("ds2ip.bin"))
    run("om_assemble", "-HM", gf, cf, t("hm.bin"))
    run("om_minverser", t("hm.bin"), t("hminv.bin"))
    run("om_assemble", "-DSM", gf, cf, df, t("dsm.bin"))
    run("om_gain", "-IP", t("hminv.bin"), t("dsm.bin"), t("h2ip.mat"), t("ds2ip.bin"), t("gip.mat"))
    run("om_assemble", "-EITSM", gf, cf, os.path.join(H1, "Head1-EIT.patches"), t("eitsm.mat"))
    run("om_gain", "-EITIP", t("hminv.bin"), t("eitsm.mat"), t("h2ip.mat"), t("geit.bin"))
    g = om.read_geometry(gf, cf)
    pts = np.loadtxt(POINTS, usecols=(0, 1, 2))
    from oracle import oracle as O
    dip = O.read_dipoles(df)
    M = om.Surf2VolMat(g, pts).array()
    S2 = om.DipSource2InternalPotMat(g, dip, pts).array()
    Gp, kp, _ = om.gain_internal_pot_pipeline(g, dip, pts, precision="fp64")
    sens = om.Sensors()
    sens.load(os.path.join(H1, "Head1-EIT.patches"))
    EIT = om.EITSourceMat(g, sens)
    Geit = om.GainEITInternalPot(om.HeadMat(g).inverse(), EIT, om.Matrix(M)).array()
    e = (rel_fro(omio.read(t("h2ip.mat"))[1], M), rel_fro(omio.read(t("ds2ip.bin"))[1], S2),
         rel_fro(omio.read(t("gip.mat"))[1], om.gain_to_numpy(Gp, len(kp), len(dip))),
         rel_fro(omio.read(t("eitsm.mat"))[1], EIT.array()), rel_fro(omio.read(t("geit.bin"))[1], Geit))
    print("  CLI vs Python: h2ip %.3e ds2ip %.3e gain %.3e eitsm %.3e eit gain %.3e" % e, flush=True)
    assert e[0] < 1e-15 and e[1] < 1e-15 and e[2] < 1e-10 and e[3] < 1e-15 and e[4] < 1e-10


def test_eit_source_mat_vs_oracle_and_reciprocity(om):
    """EITSourceMat on the device against the oracle: Head1 with the two
    point electrodes of the reference's EIT validation (intensities), with
    Head1-EIT.patches (radius 1e-3: one injection triangle, weight pi r^2 /
    |t|), and a 642-vertex sphere head with 0.15-radius electrodes (several
    injection triangles found through shared edges); then the reciprocity
    of the reference's EIT validation on the device (dipole potential
    differences vs q . grad of the EIT potential, 1% here, 13% in the
    reference's test)."""
    from oracle import oracle as O
    from openmeeg_amd import spheres as S
    gf, cf = os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond")
    g, og = om.read_geometry(gf, cf), O.load_geometry(gf, cf)
    el = np.array([[0.886556, 0.278249, -0.166667], [0.547922, -0.269672, -0.719889]])
    print("EIT: Head1 point electrodes", flush=True)
    e1 = rel_fro(om.EITSourceMat(g, el).array(), O.eit_source_mat(og, el))
    sens = om.Sensors()
    sens.load(os.path.join(H1, "Head1-EIT.patches"))
    e2 = rel_fro(om.EITSourceMat(g, sens).array(), O.eit_source_mat(og, sens.positions, sens.radii))
    meshes = S.nested_sphere_meshes(8, (0.87, 0.92, 1.0))
    cond = (1.0, 0.0125, 1.0)
    gs, ogs = om.make_nested_geometry(meshes, cond), O.nested_geometry(meshes, cond)
    pos = S.sphere_electrodes(np.random.default_rng(4), 6)
    rad = np.full(6, 0.15)
    es = om.Sensors(pos)
    es.radii = rad
    print("EIT: sphere head, 6 electrodes of radius 0.15", flush=True)
    e3 = rel_fro(om.EITSourceMat(gs, es).array(), O.eit_source_mat(ogs, pos, rad))
    # reciprocity on the device
    dip = O.read_dipoles(os.path.join(H1, "Head1.dip"))
    hm = om.HeadMat(g)
    Hi = hm.inverse().array()
    E = om.Head2EEGMat(g, el).array()
    V = E @ Hi @ om.DipSourceMat(g, dip, "Brain").array()
    diff = V[0] - V[1]
    x = Hi @ (om.EITSourceMat(g, el).array() @ np.array([1.0, -1.0]))
    P = dip[:, :3]
    v0 = om.Surf2VolMat(g, P).array() @ x
    grad = np.zeros((len(dip), 3))
    for i in range(3):
        Q = P.copy()
        Q[:, i] += 1e-6
        grad[:, i] = (om.Surf2VolMat(g, Q).array() @ x - v0) / 1e-6
    rel = np.abs((dip[:, 3:] * grad).sum(axis=1) - diff) / np.abs(diff)
    print("  vs oracle: point %.3e, radius 1e-3 %.3e, sphere r=0.15 %.3e; reciprocity max %.3e"
          % (e1, e2, e3, rel.max()), flush=True)
    assert e1 < 1e-13 and e2 < 1e-13 and e3 < 1e-13 and rel.max() < 0.01


def test_nonnested_nna1_surface_internal_eit(om):
    """HeadNNa1 (non-nested: the NORTH and SOUTH domains share the cut
    mesh's vertices): SurfSourceMat of a small source sphere in NORTH (its
    boundary is north + cut, vertex columns shared between them), Surf2VolMat
    for points in every conducting domain, and EITSourceMat for point and
    0.15-radius electrodes, against the oracle."""
    from oracle import oracle as O
    from openmeeg_amd import spheres as S
    d = os.path.join(GOLDEN, "headnna1")
    gf, cf = os.path.join(d, "HeadNNa1.geom"), os.path.join(d, "HeadNNa1.cond")
    g, og = om.read_geometry(gf, cf), O.load_geometry(gf, cf)
    P, T = S.geodesic_sphere(2)
    P = 0.15 * P + np.array([0.0, 0.0, 0.4])
    print("HeadNNa1: surface source in NORTH, internal points, EIT", flush=True)
    e_ssm = rel_fro(om.SurfSourceMat(g, om.Mesh(P, T)).array(), O.surf_source_mat(og, P, T))
    rng = np.random.default_rng(6)
    q = rng.standard_normal((60, 3))
    q *= (rng.random(60) * 0.95 / np.linalg.norm(q, axis=1))[:, None]
    q = q[np.abs(q[:, 2]) > 0.03]
    Mref, kept = O.surf2vol(og, q)
    e_ip = rel_fro(om.Surf2VolMat(g, q).array(), Mref)
    el = S.sphere_electrodes(np.random.default_rng(8), 4, 0.97)
    e_eit0 = rel_fro(om.EITSourceMat(g, el).array(), O.eit_source_mat(og, el))
    es = om.Sensors(el)
    es.radii = np.full(4, 0.15)
    e_eit1 = rel_fro(om.EITSourceMat(g, es).array(), O.eit_source_mat(og, el, es.radii))
    doms = set(og.domain_of_point(x, O.OracleLib()) for x in q[kept])
    print("  ssm %.3e, surf2vol %.3e (%d points in %d domains), EIT %.3e / %.3e"
          % (e_ssm, e_ip, len(kept), len(doms), e_eit0, e_eit1), flush=True)
    assert len(doms) >= 3
    # SurfSourceMat within 1e-12: its N block sums the edge-vector products
    # in another association than the oracle (the two contractions), and
    # here the source sits close enough for the cancellation in those sums
    # to reach 1.3e-13
    assert e_ssm < 1e-12 and e_ip < 1e-13 and e_eit0 < 1e-13 and e_eit1 < 1e-13
