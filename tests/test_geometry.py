"""Host geometry of the library (C++) against the oracle's restatement:
unknown numbering, barriers, parts, mesh pairs and coefficients."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import openmeeg_amd as om
from openmeeg_amd import spheres as S
from oracle import oracle as O


def _sizes(g):
    return (g.size, g.nb_parameters(), g.nb_current_barrier_triangles(), g.n_vertex_unknowns, g.is_nested())


def _osizes(g):
    return (g.size, g.nb_parameters, g.nb_current_barrier_triangles, g.nv_unknowns, g.nested)


@pytest.mark.parametrize("head", ["head1", "head3"])
def test_file_geometry_matches_oracle(head):
    k = head[-1]
    gf = os.path.join(GOLDEN, head, "Head%s.geom" % k)
    cf = os.path.join(GOLDEN, head, "Head%s.cond" % k)
    assert _sizes(om.read_geometry(gf, cf)) == _osizes(O.load_geometry(gf, cf))


@pytest.mark.parametrize("layers", [1, 3, 4])
def test_nested_geometry_matches_oracle(layers):
    radii = {1: (1.0,), 3: (0.87, 0.92, 1.0), 4: (0.8, 0.84, 0.9, 1.0)}[layers]
    cond = {1: (1.0,), 3: (1.0, 0.0125, 1.0), 4: (0.33, 1.79, 0.0042, 0.33)}[layers]
    meshes = S.nested_sphere_meshes(4, radii)
    g = om.make_nested_geometry(meshes, cond)
    assert _sizes(g) == _osizes(O.nested_geometry(meshes, cond))
    V, T = 10 * 16 + 2, 20 * 16
    assert g.size == layers * V + (layers - 1) * T


def test_make_geometry_equals_nested():
    meshes = S.nested_sphere_meshes(4, (0.87, 0.92, 1.0))
    g1 = om.make_nested_geometry(meshes, (1.0, 0.0125, 1.0))
    g2 = om.make_geometry({"Cortex": meshes[0], "Skull": meshes[1], "Head": meshes[2]},
                          {"Cortex": [("Cortex", 1)], "Skull": [("Skull", 1)], "Head": [("Head", 1)]},
                          {"Scalp": ([("Skull", "Outside"), ("Head", "Inside")], 1.0),
                           "Brain": ([("Cortex", "Inside")], 1.0),
                           "Air": ([("Head", "Outside")], 0.0),
                           "Skull": ([("Cortex", "Outside"), ("Skull", "Inside")], 0.0125)})
    assert _sizes(g1) == _sizes(g2)
    assert g1.domain_names() == g2.domain_names() == ["Scalp", "Brain", "Air", "Skull"]


def test_geometry_errors():
    meshes = S.nested_sphere_meshes(2, (0.9, 1.0))
    with pytest.raises(ValueError):
        om.make_nested_geometry(meshes, (1.0,))
    bad = [(meshes[0][0], meshes[0][1] + 1000)]
    with pytest.raises(ValueError, match="out of range"):
        om.make_nested_geometry(bad, (1.0,))
    with pytest.raises(ValueError, match="unknown interface"):
        om.make_geometry({"A": meshes[0]}, {"A": [("A", 1)]}, {"D": ([("B", "Inside")], 1.0)})
    with pytest.raises(ValueError):
        om.Geometry(os.path.join(GOLDEN, "head1", "Head1.geom"))


def test_sensors_load():
    s = om.Sensors()
    s.load(os.path.join(GOLDEN, "head1", "Head1.patches"))
    assert s.getNumberOfSensors() == 32
    assert s.labels[0] == "EEG001"
    np.testing.assert_allclose(s.positions[0], [-3.0882874957133399e-01, 9.5047715836211366e-01,
                                                -3.4899496702500955e-02])


def test_integrator_safe_order():
    assert om.Integrator(0).order == 1 and om.Integrator(7).order == 3
    i = om.Integrator(3, 0.001)
    assert (i.levels, i.tolerance) == (10, 0.001)
    i = om.Integrator(3, 0, 0.005)
    assert (i.order, i.levels, i.tolerance) == (3, 0, 0.005)
