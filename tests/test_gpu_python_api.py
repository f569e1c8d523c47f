"""The flow of the reference's own Python test (wrapping/python/openmeeg/
tests/test_python.py:22-190) on Head1 against this module: geometry from the
files or from numpy meshes, HeadMat / inverse / invert, DipSourceMat,
Head2EEGMat, GainEEG / GainEEGadjoint, SurfSourceMat with a Mesh and the
surface gains, the MEG matrices and gains (dipoles as an array or a Matrix,
a transposed array refused with "6 columns"), GainEEGMEGadjoint, Forward,
and saving / loading the matrices -- with the reference's shape checks, and
values against the Head1 goldens.  Needs a HIP GPU."""
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


@pytest.mark.parametrize("load_from_numpy", [True, False])
def test_reference_python_flow_head1(om, load_from_numpy, tmp_path):
    from oracle.geometry import read_tri
    geom_file, cond_file = os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond")
    dipole_file, patches_file = os.path.join(H1, "Head1.dip"), os.path.join(H1, "Head1.patches")
    print("reference Python flow, load_from_numpy=%s" % load_from_numpy, flush=True)
    if load_from_numpy:
        meshes = [read_tri(os.path.join(H1, f)) for f in ("cortex.1.tri", "skull.1.tri", "scalp.1.tri")]
        geom = om.make_nested_geometry(meshes, conductivity=(1, 0.0125, 1))
        dipoles = om.Matrix(np.asfortranarray(np.loadtxt(dipole_file)))
        patches = om.Sensors(om.Matrix(np.asfortranarray(np.loadtxt(patches_file, usecols=range(1, 4)))), geom)
    else:
        geom = om.read_geometry(geom_file, cond_file)
        dipoles = om.Matrix()
        dipoles.load(dipole_file)
        patches = om.Sensors()
        patches.load(patches_file)
    assert geom.is_nested()
    n_dipoles = dipoles.nlin()
    sensors = om.Sensors()
    sensors.load(os.path.join(H1, "Head1.squids"))
    n_meg_sensors = sensors.getNumberOfSensors()
    n_eeg_sensors = patches.getNumberOfSensors()

    integrator = om.Integrator(3, 0, 0.005)
    with om.use_log_level("debug"):
        hm = om.HeadMat(geom, integrator)
    hminv = hm.inverse()
    with om.use_log_level("debug"):
        hminv_inplace = om.HeadMat(geom, integrator)
    hminv_inplace.invert()
    np.testing.assert_allclose(om.Matrix(hminv).array(), om.Matrix(hminv_inplace).array())

    dsm = om.DipSourceMat(geom, dipoles, "Brain")
    h2em = om.Head2EEGMat(geom, patches)
    gain_eeg_dip = om.GainEEG(hminv, dsm, h2em)
    gain_adjoint_eeg_dip = om.GainEEGadjoint(geom, dipoles, hm, h2em)
    n_hm_unknowns = geom.nb_parameters() - geom.nb_current_barrier_triangles()
    assert hm.nlin() == hm.ncol() == n_hm_unknowns
    assert hminv.nlin() == hminv.ncol() == n_hm_unknowns
    assert (dsm.nlin(), dsm.ncol()) == (n_hm_unknowns, n_dipoles)
    assert (gain_eeg_dip.nlin(), gain_eeg_dip.ncol()) == (n_eeg_sensors, n_dipoles)
    assert (gain_adjoint_eeg_dip.nlin(), gain_adjoint_eeg_dip.ncol()) == (n_eeg_sensors, n_dipoles)

    mesh = om.Mesh(os.path.join(H1, "Head1.tri"))
    n_dipoles_surf = mesh.vertices().size()
    ssm = om.SurfSourceMat(geom, mesh, integrator)
    gain_eeg_surf = om.GainEEG(hminv, ssm, h2em)
    assert (ssm.nlin(), ssm.ncol()) == (n_hm_unknowns, n_dipoles_surf)
    assert (gain_eeg_surf.nlin(), gain_eeg_surf.ncol()) == (n_eeg_sensors, n_dipoles_surf)

    ss2mm = om.SurfSource2MEGMat(mesh, sensors)
    ds2mm = om.DipSource2MEGMat(dipoles, sensors)
    dip_arr = np.zeros((4, 6))
    dip_arr[:, 2] = [0.02, 0.03, 0.04, 0.05]
    dip_arr[:, 3] = 1.0
    dip_arr = np.asfortranarray(dip_arr)
    np.testing.assert_allclose(om.DipSource2MEGMat(dip_arr, sensors).array(),
                               om.DipSource2MEGMat(om.Matrix(dip_arr), sensors).array())
    with pytest.raises(Exception, match="6 columns"):
        om.DipSource2MEGMat(np.asfortranarray(dip_arr.T), sensors)
    h2mm = om.Head2MEGMat(geom, sensors)
    gain_meg_surf = om.GainMEG(hminv, ssm, h2mm, ss2mm)
    gain_meg_dip = om.GainMEG(hminv, dsm, h2mm, ds2mm)
    gain_adjoint_meg_dip = om.GainMEGadjoint(geom, dipoles, hm, h2mm, ds2mm)
    gain_adjoint_eeg_meg_dip = om.GainEEGMEGadjoint(geom, dipoles, hm, h2em, h2mm, ds2mm)
    assert gain_adjoint_eeg_meg_dip.nlin() == n_meg_sensors + n_eeg_sensors
    assert (ss2mm.nlin(), ss2mm.ncol()) == (n_meg_sensors, n_dipoles_surf)
    assert (ds2mm.nlin(), ds2mm.ncol()) == (n_meg_sensors, n_dipoles)
    assert (h2mm.nlin(), h2mm.ncol()) == (n_meg_sensors, n_hm_unknowns)
    assert (h2em.nlin(), h2em.ncol()) == (n_eeg_sensors, n_hm_unknowns)
    assert (gain_meg_dip.nlin(), gain_meg_dip.ncol()) == (n_meg_sensors, n_dipoles)
    assert (gain_meg_surf.nlin(), gain_meg_surf.ncol()) == (n_meg_sensors, n_dipoles_surf)
    assert (gain_adjoint_meg_dip.nlin(), gain_adjoint_meg_dip.ncol()) == (n_meg_sensors, n_dipoles)
    gain_meg_surf_one_line = om.GainMEG(om.HeadMat(geom, integrator).inverse(), om.SurfSourceMat(geom, mesh, integrator),
                                        om.Head2MEGMat(geom, sensors), om.SurfSource2MEGMat(mesh, sensors))
    np.testing.assert_allclose(gain_meg_surf_one_line.array(), gain_meg_surf.array())

    sources = om.Matrix(os.path.join(H1, "Head1.srcdip"))
    n_times = sources.ncol()
    for G, ns in ((gain_meg_dip, n_meg_sensors), (gain_adjoint_meg_dip, n_meg_sensors),
                  (gain_eeg_dip, n_eeg_sensors), (gain_adjoint_eeg_dip, n_eeg_sensors)):
        est = om.Forward(G, sources, 0.0)
        assert (est.nlin(), est.ncol()) == (ns, n_times)

    hm_fname, ssm_fname = str(tmp_path / "Head1.hm"), str(tmp_path / "Head1.ssm")
    hm.save(hm_fname)
    ssm.save(ssm_fname)
    m1 = om.SymMatrix()
    m1.load(hm_fname)
    m2 = om.Matrix()
    m2.load(ssm_fname)
    assert np.array_equal(m1.data(), hm.data()) and np.array_equal(m2.array(), ssm.array())

    gold = lambda n: np.load(os.path.join(H1, n + ".npy"))
    errs = dict(hm=rel_fro(hm.data(), gold("hm")), dgem=rel_fro(gain_eeg_dip.array(), gold("dgem")),
                dgem_adj=rel_fro(gain_adjoint_eeg_dip.array(), gold("dgem")),
                dgmm=rel_fro(gain_meg_dip.array(), gold("dgmm")), sgem=rel_fro(gain_eeg_surf.array(), gold("sgem")),
                sgmm=rel_fro(gain_meg_surf.array(), gold("sgmm")),
                eegmeg=rel_fro(gain_adjoint_eeg_meg_dip.MEGleadfield.array(), gold("dgmm")))
    print("  vs goldens:", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert errs["hm"] < 1e-13 and max(v for k, v in errs.items() if k != "hm") < 1e-10
