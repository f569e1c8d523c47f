"""The Python surface as a contract: the names the reference module exports
(wrapping/python/openmeeg/__init__.py:15-39) exist here, and each product
they build has the shape its arguments fix and, on Head1, the values of the
reference's goldens.

The shapes are written once as a table over five sizes of a problem: the
HeadMat unknowns N (nb_parameters - nb_current_barrier_triangles,
assembleHeadMat.cpp:77), the EEG electrodes E, the MEG sensors M, the
dipoles D and the vertices V of a source mesh.  The products are built from
two geometries (the .geom/.cond files and numpy meshes through
make_nested_geometry) and must agree with each other, with the goldens and
with the table.  The name check runs on the CPU; the rest needs a HIP GPU.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

H1 = os.path.join(GOLDEN, "head1")

# The reference's exported names that belong to the paths built here.
EXPORTED = ("HeadMat", "Sensors", "Integrator", "Head2EEGMat", "Head2MEGMat", "DipSourceMat", "DipSource2MEGMat",
            "DipSource2InternalPotMat", "EITSourceMat", "GainEEG", "GainMEG", "GainEEGadjoint", "GainMEGadjoint",
            "GainEEGMEGadjoint", "GainInternalPot", "GainEITInternalPot", "Forward", "Surf2VolMat", "SurfSourceMat",
            "SurfSource2MEGMat", "Matrix", "SymMatrix", "make_geometry", "make_nested_geometry", "read_geometry",
            "get_log_level", "set_log_level", "use_log_level")
# Exported by the reference but outside these paths: monopole sources,
# ECoG projection and the sensor-array generator of _make_sensors.py.
NOT_BUILT = ("MonopoleSourceMat", "Head2ECoGMat", "make_sensors")

# product -> (rows, columns) in the sizes N, E, M, D, V
SHAPES = {
    "HeadMat": ("N", "N"),
    "HeadMat.inverse": ("N", "N"),
    "DipSourceMat": ("N", "D"),
    "SurfSourceMat": ("N", "V"),
    "Head2EEGMat": ("E", "N"),
    "Head2MEGMat": ("M", "N"),
    "DipSource2MEGMat": ("M", "D"),
    "SurfSource2MEGMat": ("M", "V"),
    "GainEEG(dipoles)": ("E", "D"),
    "GainEEG(surface)": ("E", "V"),
    "GainEEGadjoint": ("E", "D"),
    "GainMEG(dipoles)": ("M", "D"),
    "GainMEG(surface)": ("M", "V"),
    "GainMEGadjoint": ("M", "D"),
}

# product -> golden file of the reference's Head1 regression data
GOLDENS = {
    "DipSourceMat": "dsm", "SurfSourceMat": "ssm", "Head2MEGMat": "h2mm", "DipSource2MEGMat": "ds2mm",
    "SurfSource2MEGMat": "ss2mm", "GainEEG(dipoles)": "dgem", "GainEEG(surface)": "sgem",
    "GainEEGadjoint": "dgem", "GainMEG(dipoles)": "dgmm", "GainMEG(surface)": "sgmm", "GainMEGadjoint": "dgmm",
}


def test_reference_names_are_exported():
    import openmeeg_amd as om
    missing = [n for n in EXPORTED if not hasattr(om, n)]
    assert not missing, missing
    assert not any(hasattr(om, n) for n in NOT_BUILT)
    assert isinstance(om.get_log_level(), str)


@pytest.fixture(scope="module")
def om():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openmeeg_amd
    return openmeeg_amd


def _geometry(om, source):
    if source == "files":
        return om.read_geometry(os.path.join(H1, "Head1.geom"), os.path.join(H1, "Head1.cond"))
    from oracle.geometry import read_tri
    layers = [read_tri(os.path.join(H1, name)) for name in ("cortex.1.tri", "skull.1.tri", "scalp.1.tri")]
    return om.make_nested_geometry(layers, conductivity=(1, 0.0125, 1))


def _products(om, geom):
    """Every product of the table on Head1, built through the exported names."""
    integ = om.Integrator(3, 0, 0.005)
    dip = np.loadtxt(os.path.join(H1, "Head1.dip"))
    eeg = om.Sensors()
    eeg.load(os.path.join(H1, "Head1.patches"))
    meg = om.Sensors()
    meg.load(os.path.join(H1, "Head1.squids"))
    src = om.Mesh(os.path.join(H1, "Head1.tri"))
    hm = om.HeadMat(geom, integ)
    inv = hm.inverse()
    out = {"HeadMat": hm, "HeadMat.inverse": inv}
    out["DipSourceMat"] = om.DipSourceMat(geom, om.Matrix(np.asfortranarray(dip)), "Brain")
    out["SurfSourceMat"] = om.SurfSourceMat(geom, src, integ)
    out["Head2EEGMat"] = om.Head2EEGMat(geom, eeg)
    out["Head2MEGMat"] = om.Head2MEGMat(geom, meg)
    out["DipSource2MEGMat"] = om.DipSource2MEGMat(dip, meg)
    out["SurfSource2MEGMat"] = om.SurfSource2MEGMat(src, meg)
    out["GainEEG(dipoles)"] = om.GainEEG(inv, out["DipSourceMat"], out["Head2EEGMat"])
    out["GainEEG(surface)"] = om.GainEEG(inv, out["SurfSourceMat"], out["Head2EEGMat"])
    out["GainEEGadjoint"] = om.GainEEGadjoint(geom, dip, hm, out["Head2EEGMat"])
    out["GainMEG(dipoles)"] = om.GainMEG(inv, out["DipSourceMat"], out["Head2MEGMat"], out["DipSource2MEGMat"])
    out["GainMEG(surface)"] = om.GainMEG(inv, out["SurfSourceMat"], out["Head2MEGMat"], out["SurfSource2MEGMat"])
    out["GainMEGadjoint"] = om.GainMEGadjoint(geom, dip, hm, out["Head2MEGMat"], out["DipSource2MEGMat"])
    sizes = dict(N=geom.nb_parameters() - geom.nb_current_barrier_triangles(), E=eeg.getNumberOfSensors(),
                 M=meg.getNumberOfSensors(), D=len(dip), V=src.nb_vertices())
    return out, sizes, dict(dip=dip, eeg=eeg, meg=meg)


@pytest.fixture(scope="module")
def built(om):
    print("Head1 products from the .geom files and from numpy meshes", flush=True)
    return {s: _products(om, _geometry(om, s)) for s in ("files", "arrays")}


@pytest.mark.gpu
def test_shapes_follow_the_table(built):
    for source, (out, sizes, _) in built.items():
        assert set(out) == set(SHAPES)
        for name, (r, c) in SHAPES.items():
            m = out[name]
            assert (m.nlin(), m.ncol()) == (sizes[r], sizes[c]), (source, name)
    n = built["files"][1]
    assert n["N"] == 286 and n["E"] == 32 and n["D"] == 6


@pytest.mark.gpu
def test_values_match_goldens_and_both_geometries_agree(built):
    files, arrays = built["files"][0], built["arrays"][0]
    errs = {}
    for name, gold in GOLDENS.items():
        ref = np.load(os.path.join(H1, gold + ".npy"))
        errs[name] = rel_fro(files[name].array(), ref)
        assert rel_fro(arrays[name].array(), files[name].array()) < 1e-12, name
    errs["HeadMat"] = rel_fro(files["HeadMat"].data(), np.load(os.path.join(H1, "hm.npy")))
    errs["HeadMat.inverse"] = rel_fro(files["HeadMat.inverse"].data(), np.load(os.path.join(H1, "hm_inv.npy")))
    print("Head1 vs goldens:", {k: "%.1e" % v for k, v in errs.items()}, flush=True)
    for name, e in errs.items():
        assert e < (1e-13 if name in ("HeadMat", "DipSourceMat", "SurfSourceMat", "DipSource2MEGMat") else 1e-10), name


@pytest.mark.gpu
def test_combined_lead_fields_and_forward(om, built):
    """GainEEGMEGadjoint stacks the EEG and MEG lead fields of one solve
    (gain.h:105-128); Forward maps (sensors x sources) by (sources x times)
    to (sensors x times)."""
    out, sizes, inp = built["files"]
    geom = _geometry(om, "files")
    both = om.GainEEGMEGadjoint(geom, inp["dip"], out["HeadMat"], out["Head2EEGMat"], out["Head2MEGMat"],
                                out["DipSource2MEGMat"])
    assert both.nlin() == sizes["E"] + sizes["M"]
    assert rel_fro(both.EEGleadfield.array(), out["GainEEGadjoint"].array()) < 1e-12
    assert rel_fro(both.MEGleadfield.array(), out["GainMEGadjoint"].array()) < 1e-12
    act = om.Matrix(os.path.join(H1, "Head1.srcdip"))
    for name in ("GainEEG(dipoles)", "GainMEGadjoint"):
        y = om.Forward(out[name], act, 0.0)
        assert (y.nlin(), y.ncol()) == (out[name].nlin(), act.ncol())
        np.testing.assert_allclose(y.array(), out[name].array() @ act.array(), rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_argument_contract(om, built, tmp_path):
    """Dipoles are an n x 6 array or Matrix (both give the same source
    matrix); another column count is a ValueError; HeadMat.invert() equals
    inverse(); the symmetric and full matrices survive save/load."""
    out, _, inp = built["files"]
    geom = _geometry(om, "files")
    as_array = om.DipSourceMat(geom, inp["dip"], "Brain").array()
    assert np.array_equal(as_array, out["DipSourceMat"].array())
    with pytest.raises(ValueError, match="6 columns"):
        om.DipSourceMat(geom, inp["dip"][:, :5], "Brain")
    with pytest.raises(ValueError, match="6 columns"):
        om.DipSource2MEGMat(np.asfortranarray(inp["dip"][:4].T), inp["meg"])
    inv2 = om.HeadMat(geom)
    inv2.invert()
    assert rel_fro(inv2.data(), out["HeadMat.inverse"].data()) < 1e-14
    for suffix in (".mat", ".bin"):
        hm_path, ssm_path = str(tmp_path / ("hm" + suffix)), str(tmp_path / ("ssm" + suffix))
        out["HeadMat"].save(hm_path)
        out["SurfSourceMat"].save(ssm_path)
        back_h, back_s = om.SymMatrix(), om.Matrix()
        back_h.load(hm_path)
        back_s.load(ssm_path)
        assert np.array_equal(back_h.data(), out["HeadMat"].data())
        assert np.array_equal(back_s.array(), out["SurfSourceMat"].array())
