"""The reference's matrix file formats (libopenmeeg_amd_io, host only):
reading the reference's own MAT-file 7.3 goldens, writing MAT / raw binary /
ASCII files that hold the same variables, byte layouts and text as the
reference's writers, and round trips of every storage type."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

from openmeeg_amd import io as omio

MAT = os.path.join(GOLDEN, "head1", "mat")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _npy(name):
    return np.load(os.path.join(GOLDEN, "head1", name + ".npy"))


def test_reads_reference_mat_files():
    """The MAT 7.3 goldens (symmetric struct, full matrices, sparse CSC) read
    back to the values extracted from them with h5dump."""
    kind, p = omio.read(os.path.join(MAT, "Head1.hm"))
    assert kind == "symmetric" and np.array_equal(p, _npy("hm"))
    for name in ("dsm", "ss2mm", "est_eeg"):
        kind, a = omio.read(os.path.join(MAT, "Head1." + name))
        assert kind == "full" and np.array_equal(a, _npy(name)), name
    kind, (r, c, v, shape) = omio.read(os.path.join(MAT, "Head1.h2em"))
    assert kind == "sparse" and shape == (32, 286) and len(v) == 96
    data, ir, jc = (_npy("h2em_" + k) for k in ("data", "ir", "jc"))
    E = np.zeros(shape)
    for j in range(shape[1]):
        for k in range(int(jc[j]), int(jc[j + 1])):
            E[int(ir[k]), j] = data[k]
    Er = np.zeros(shape)
    Er[r, c] = v
    assert np.array_equal(E, Er)
    assert np.all(np.diff(r * shape[1] + c) > 0)          # (row, col) order


@pytest.mark.parametrize("suffix", ["mat", "bin", "txt"])
def test_round_trips(tmp_path, suffix):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((7, 5))
    v = rng.standard_normal(9)
    P = rng.standard_normal(6 * 7 // 2)
    r = np.array([4, 0, 2, 0]), np.array([1, 3, 0, 0])
    sv = rng.standard_normal(4)
    tol = 0 if suffix != "txt" else 1e-5        # ASCII keeps 6 significant digits
    f = str(tmp_path / ("m." + suffix))
    omio.write_full(f, A)
    kind, B = omio.read(f)
    assert kind == "full" and B.shape == A.shape and np.allclose(B, A, rtol=tol, atol=0)
    omio.write_full(f, v)
    kind, w = omio.read(f)
    assert kind == "full" and np.allclose(np.ravel(w), v, rtol=tol, atol=0)
    omio.write_symmetric(f, P)
    kind, Q = omio.read(f)
    assert kind == "symmetric" and np.allclose(Q, P, rtol=tol, atol=0)
    omio.write_sparse(f, r[0], r[1], sv, (5, 4))
    kind, (rr, cc, vv, shape) = omio.read(f)
    order = np.lexsort((r[1], r[0]))
    assert kind == "sparse" and shape == (5, 4)
    assert np.array_equal(rr, r[0][order]) and np.array_equal(cc, r[1][order])
    assert np.allclose(vv, sv[order], rtol=tol, atol=0)


def test_bin_and_ascii_bytes(tmp_path):
    """Byte layouts of TrivialBinIO (u32 header, column-major doubles, packed
    upper, (u32 i, u32 j, f64 v) records in (i, j) order) and the text of
    AsciiIO (default stream precision, tab separated rows)."""
    A = np.array([[1.5, -2.0, 3.25], [4.0, 5.125, -6.0]])
    f = str(tmp_path / "a.bin")
    omio.write_full(f, A)
    assert open(f, "rb").read() == struct.pack("<II", 2, 3) + A.tobytes(order="F")
    P = np.array([1.0, 2.0, 3.0])                  # 2 x 2: (0,0) (0,1) (1,1)
    omio.write_symmetric(f, P)
    assert open(f, "rb").read() == struct.pack("<I", 2) + P.tobytes()
    omio.write_sparse(f, [1, 0], [0, 2], [7.0, -1.0], (2, 3))
    assert open(f, "rb").read() == struct.pack("<II", 2, 3) + struct.pack("<IId", 0, 2, -1.0) + \
        struct.pack("<IId", 1, 0, 7.0)
    t = str(tmp_path / "a.txt")
    B = np.array([[0.1234567891, 12345678.9], [1e-7, -3.0]])
    omio.write_full(t, B)
    assert open(t).read() == "0.123457\t1.23457e+07\n1e-07\t-3\n"
    omio.write_symmetric(t, np.array([1.0, 2.0, 3.0]))
    assert open(t).read() == "1\t2\n3\n"
    omio.write_sparse(t, [1, 0], [0, 2], [7.0, -1.5], (2, 3))
    assert open(t).read() == "2 3\n0 2 -1.5\n1 0 7\n"


def _structure(path):
    """Groups, datasets, types, shapes and attribute values from h5dump -H -A,
    without the storage layout (chunking / compression) or the header."""
    out = subprocess.run([H5DUMP, "-H", "-A", path], capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()[1:]
    return [ln.strip() for ln in lines]


def _layout(path):
    """Per dataset its storage layout (CHUNKED with the chunk shape, or
    CONTIGUOUS) and filters, from h5dump -p -H (compressed sizes left out)."""
    out = subprocess.run([H5DUMP, "-p", "-H", path], capture_output=True, text=True, check=True).stdout
    keep = ("DATASET", "CHUNKED", "CONTIGUOUS", "COMPACT", "COMPRESSION DEFLATE")
    return [ln.strip() for ln in out.splitlines() if ln.strip().startswith(keep)]


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
@pytest.mark.parametrize("name", ["hm", "dsm", "h2em", "ss2mm"])
def test_mat_files_hold_the_reference_variables(tmp_path, name):
    """A MAT file written here holds the same variables as the reference's
    (names, MATLAB classes and fields, HDF5 types and shapes) -- what matio
    reads -- stored the same way (chunk shapes, DEFLATE level, contiguous
    sparse components), and starts with a MATLAB 7.3 header in a 512-byte
    user block."""
    ref = os.path.join(MAT, "Head1." + name)
    kind, val = omio.read(ref)
    f = str(tmp_path / (name + ".mat"))
    if kind == "symmetric":
        omio.write_symmetric(f, val)
    elif kind == "sparse":
        omio.write_sparse(f, *val)
    else:
        omio.write_full(f, val)
    assert _structure(f) == _structure(ref)
    # the reference's storage: dense datasets in DEFLATE-9 chunks of the same
    # shape, sparse components contiguous
    assert _layout(f) == _layout(ref), (_layout(f), _layout(ref))
    head = open(f, "rb").read(128)
    assert head.startswith(b"MATLAB 7.3 MAT-file") and head[124:128] == b"\x00\x02IM"
    kind2, val2 = omio.read(f)
    assert kind2 == kind
    if kind == "sparse":
        assert all(np.array_equal(a, b) for a, b in zip(val[:3], val2[:3])) and val[3] == val2[3]
    else:
        assert np.array_equal(val, val2)


def test_unknown_suffix_writes_mat(tmp_path):
    """Matrix::save falls back to the first format that stores the matrix
    when the suffix names none (matrix.cpp:219-226): MAT, as the reference's
    Head1.hm, Head1.dsm, ... are."""
    for name in ("a.hm", "a"):
        f = str(tmp_path / name)
        omio.write_full(f, np.eye(2))
        assert open(f, "rb").read(10) == b"MATLAB 7.3"
        assert omio.info(f).format == 0 and np.array_equal(omio.read(f)[1], np.eye(2))


def test_tools_print_usage_and_refuse_unknown_options():
    """The command-line tools are built, print their usage without a GPU and
    refuse the options this build does not provide with a message."""
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openmeeg_amd", "bin")
    for tool in ("om_assemble", "om_minverser", "om_gain"):
        r = subprocess.run([os.path.join(root, tool)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout, tool
    r = subprocess.run([os.path.join(root, "om_assemble"), "-CorticalMat", "a", "b", "c", "d", "e"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not supported" in r.stderr


def test_io_library_exports_every_declared_symbol():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = open(os.path.join(root, "include", "openmeeg_amd_io.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(omh_\w+)\s*\(", txt)))
    assert names == ["omh_io_info", "omh_io_last_error", "omh_io_read", "omh_io_write"]
    L = omio.lib()
    for n in names:
        assert hasattr(L, n), n
