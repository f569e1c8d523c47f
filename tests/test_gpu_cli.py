"""The command-line tools (om_assemble, om_minverser, om_gain) end to end on
Head1: every matrix they write, in MAT 7.3, raw binary or ASCII, against the
reference's goldens of the same command (tests/initialTest/Head1.*).  Needs
a HIP GPU.  Tolerances as the parity tests: assembled matrices 1e-13
(relative Frobenius), inverse and gains 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openmeeg_amd", "bin")
H1 = os.path.join(GOLDEN, "head1")


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def run(tool, *args):
    print("run", tool, " ".join(os.path.basename(a) for a in args), flush=True)
    r = subprocess.run([os.path.join(BIN, tool)] + list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def load(path):
    from openmeeg_amd import io as omio
    return omio.read(path)


def dense(sp):
    r, c, v, shape = sp
    E = np.zeros(shape)
    E[r, c] = v
    return E


def h1(name):
    return os.path.join(H1, name)


def test_cli_head1_eeg_meg_surface(gpu, tmp_path):
    from oracle import oracle as O
    t = lambda f: str(tmp_path / f)
    g, c = h1("Head1.geom"), h1("Head1.cond")
    run("om_assemble", "-HM", g, c, t("hm.mat"))
    run("om_assemble", "-DSM", g, c, h1("Head1.dip"), t("dsm.bin"), "Brain")
    run("om_assemble", "-H2EM", g, c, h1("Head1.patches"), t("h2em.mat"))
    run("om_minverser", t("hm.mat"), t("hminv.bin"))
    run("om_gain", "-EEG", t("hminv.bin"), t("dsm.bin"), t("h2em.mat"), t("gain.mat"))
    run("om_gain", "-EEGadjoint", g, c, h1("Head1.dip"), t("hm.mat"), t("h2em.mat"), t("gainadj.txt"))
    run("om_assemble", "-H2MM", g, c, h1("Head1.squids"), t("h2mm.mat"))
    run("om_assemble", "-DS2MM", h1("Head1.dip"), h1("Head1.squids"), t("ds2mm.mat"))
    run("om_gain", "-MEG", t("hminv.bin"), t("dsm.bin"), t("h2mm.mat"), t("ds2mm.mat"), t("gmeg.mat"))
    run("om_gain", "-EEGMEGadjoint", g, c, h1("Head1.dip"), t("hm.mat"), t("h2em.mat"), t("h2mm.mat"),
        t("ds2mm.mat"), t("ge.bin"), t("gm.bin"))
    run("om_assemble", "-SSM", g, c, h1("Head1.tri"), t("ssm.mat"))
    run("om_assemble", "-SS2MM", h1("Head1.tri"), h1("Head1.squids"), t("ss2mm.mat"))
    run("om_gain", "-EEG", t("hminv.bin"), t("ssm.mat"), t("h2em.mat"), t("sgem.mat"))
    gold = lambda n: np.load(os.path.join(H1, n + ".npy"))
    kind, hm = load(t("hm.mat"))
    assert kind == "symmetric"
    errs = {
        "hm": rel_fro(hm, gold("hm")),
        "dsm": rel_fro(load(t("dsm.bin"))[1], gold("dsm")),
        "hminv": rel_fro(load(t("hminv.bin"))[1], gold("hm_inv")),
        "gain": rel_fro(load(t("gain.mat"))[1], gold("dgem")),
        "gain_adjoint_txt": rel_fro(load(t("gainadj.txt"))[1], gold("dgem")),
        "h2mm": rel_fro(load(t("h2mm.mat"))[1], gold("h2mm")),
        "ds2mm": rel_fro(load(t("ds2mm.mat"))[1], gold("ds2mm")),
        "gain_meg": rel_fro(load(t("gmeg.mat"))[1], gold("dgmm")),
        "eegmeg_eeg": rel_fro(load(t("ge.bin"))[1], gold("dgem")),
        "eegmeg_meg": rel_fro(load(t("gm.bin"))[1], gold("dgmm")),
        "ssm": rel_fro(load(t("ssm.mat"))[1], gold("ssm")),
        "ss2mm": rel_fro(load(t("ss2mm.mat"))[1], gold("ss2mm")),
        "sgem": rel_fro(load(t("sgem.mat"))[1], gold("sgem")),
    }
    kind, sp = load(t("h2em.mat"))
    data, ir, jc = (gold("h2em_" + k) for k in ("data", "ir", "jc"))
    E = np.zeros((32, 286))
    for j in range(286):
        E[ir[jc[j]:jc[j + 1]].astype(int), j] = data[jc[j]:jc[j + 1]]
    errs["h2em"] = rel_fro(dense(sp), E)
    print("CLI vs goldens:", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    for k in ("hm", "dsm", "ssm"):
        assert errs[k] < 1e-13, k
    assert errs["h2em"] < 1e-14
    for k in ("ds2mm",):
        assert errs[k] < 1e-14, k
    for k in ("h2mm", "ss2mm"):
        assert errs[k] < 1e-10, k
    for k in ("hminv", "gain", "gain_meg", "eegmeg_eeg", "eegmeg_meg", "sgem"):
        assert errs[k] < 1e-10, k
    assert errs["gain_adjoint_txt"] < 1e-5          # ASCII: 6 significant digits


def test_cli_old_ordering(gpu, tmp_path):
    """om_assemble -old-ordering (apps/assemble.cpp:41, geometry.cpp:256-295):
    unknowns numbered mesh by mesh, (V1, p1, V2, p2, V3).  Head1 (cortex,
    skull, scalp; 42 vertices and 80 triangles each; scalp triangles are
    current-barrier unknowns outside the matrix): the HeadMat diagonal must
    read 42 negative (vertex) then 80 positive (triangle) entries per inner
    mesh, then the scalp's 42 vertices; the matrices equal the default ones
    under that renumbering (HeadMat both ways, DipSourceMat rows,
    Head2EEGMat columns), bit for bit."""
    t = lambda f: str(tmp_path / f)
    g, c = h1("Head1.geom"), h1("Head1.cond")
    for flag, tag in (([], "new"), (["-old-ordering"], "old")):
        run("om_assemble", *flag, "-HM", g, c, t("hm_%s.bin" % tag))
        run("om_assemble", *flag, "-DSM", g, c, h1("Head1.dip"), t("dsm_%s.bin" % tag), "Brain")
        run("om_assemble", *flag, "-H2EM", g, c, h1("Head1.patches"), t("h2em_%s.bin" % tag))
    from oracle import oracle as O
    H_new = O.unpack(load(t("hm_new.bin"))[1], 286)
    H_old = O.unpack(load(t("hm_old.bin"))[1], 286)
    # reference numbering: V_cortex V_skull V_scalp (3 x 42), T_cortex T_skull (2 x 80)
    perm = np.empty(286, dtype=np.int64)
    nv, nt = 42, 80
    for m in range(3):
        base = m * (nv + nt)
        perm[m * nv:(m + 1) * nv] = base + np.arange(nv)
        if m < 2:
            perm[3 * nv + m * nt:3 * nv + (m + 1) * nt] = base + nv + np.arange(nt)
    d = np.sign(np.diag(H_old))
    expect = np.concatenate([-np.ones(nv), np.ones(nt), -np.ones(nv), np.ones(nt), -np.ones(nv)])
    assert np.array_equal(d, expect)
    P = np.zeros((286, 286))
    P[perm, np.arange(286)] = 1.0
    assert np.array_equal(H_old, P @ H_new @ P.T)
    D_new, D_old = load(t("dsm_new.bin"))[1], load(t("dsm_old.bin"))[1]
    assert np.array_equal(D_old[perm], D_new)
    E_new, E_old = dense(load(t("h2em_new.bin"))[1]), dense(load(t("h2em_old.bin"))[1])
    assert np.array_equal(E_old[:, perm], E_new)
