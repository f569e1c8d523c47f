import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running test")


def rel_fro(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def head1_oracle():
    """Oracle geometry + matrices of the reference's Head1 test head."""
    from oracle import oracle as O
    g = O.load_geometry(os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"))
    return g


@pytest.fixture(scope="session")
def c0_oracle():
    """Oracle results on the 642-vertex nested-sphere head (n = 4486)."""
    from oracle import oracle as O
    from openmeeg_amd import spheres as S
    meshes, cond, dip, el = S.make_config("c0")
    g = O.nested_geometry(meshes, cond)
    Hp = O.head_mat(g)
    H = O.unpack(Hp, g.size)
    D = O.dip_source_mat(g, dip, "Brain")
    cols, vals = O.head2eeg(g, el)
    E = O.h2em_dense(g, cols, vals)
    G = O.gain_eeg(H, D, E)
    return dict(meshes=meshes, cond=cond, dip=dip, el=el, g=g, H=H, D=D, cols=cols, vals=vals, E=E, G=G)
