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


HEAD3 = (os.path.join(GOLDEN, "head3", "Head3.geom"), os.path.join(GOLDEN, "head3", "Head3.cond"))


def c0_inputs():
    """The 642-vertex three-sphere head of the reference's Head3 test
    (data/Head3: skull, cortex, scalp .tri files, radii 0.87/0.92/1.0,
    conductivities 1/0.0125/1) with 100 seeded dipoles uniform in r < 0.8
    (unit random moments) and 64 seeded electrodes on r = 1."""
    from openmeeg_amd import spheres as S
    rng = np.random.default_rng(0)
    dip = S.random_dipoles(rng, 100, 0.8)
    el = S.sphere_electrodes(rng, 64, 1.0)
    return dip, el


@pytest.fixture(scope="session")
def c0_oracle():
    """Oracle results on the Head3 head (n = 4486)."""
    from oracle import oracle as O
    dip, el = c0_inputs()
    g = O.load_geometry(*HEAD3)
    Hp = O.head_mat(g)
    H = O.unpack(Hp, g.size)
    D = O.dip_source_mat(g, dip, "Brain")
    cols, vals = O.head2eeg(g, el)
    E = O.h2em_dense(g, cols, vals)
    G = O.gain_eeg(H, D, E)
    return dict(geom_files=HEAD3, cond=(1.0, 0.0125, 1.0), radii=(0.87, 0.92, 1.0), dip=dip, el=el, g=g, H=H, D=D,
                cols=cols, vals=vals, E=E, G=G)
