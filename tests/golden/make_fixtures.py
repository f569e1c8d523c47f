"""Build the golden fixtures used by the parity tests.

Run in a container that holds the reference checkout (default /root/reference)
and conda's HDF5 tools (``/opt/conda/bin/h5dump``).  The output is committed:
the GPU box never reads the reference.

What is extracted (all data files that the reference's own tests hold):

* ``head1/``: the Head1 regression goldens ``tests/initialTest/Head1.{hm,hm_inv,
  dsm,h2em,dgem}`` (MAT v7.3 = HDF5), converted to ``.npy``:
  - ``hm.npy`` / ``hm_inv.npy``: packed upper storage of the SymMatrix
    (index ``i + j*(j+1)/2`` for ``i <= j``), length n(n+1)/2;
  - ``dsm.npy`` / ``dgem.npy``: dense matrices (rows x cols);
  - ``h2em_{data,ir,jc}.npy``: compressed-sparse-column H2EM;
  - ``h2mm.npy`` / ``ds2mm.npy`` / ``dgmm.npy``: the MEG goldens (Head2MEGMat,
    DipSource2MEGMat, GainMEG with ``Head1.squids``), dense;
  - ``ssm.npy`` / ``sgem.npy`` / ``ss2mm.npy`` / ``sgmm.npy`` / ``est_eeg.npy``:
    the surface-source goldens (SurfSourceMat and SurfSource2MEGMat of the
    source mesh ``Head1.tri``, the surface EEG and MEG gains, and the EEG
    forward of the activation ``Head1.src`` with no noise), dense;
  plus the Head1 inputs (``data/Head1``: meshes, .geom, .cond, .dip, .patches,
  .squids, the source mesh .tri and its activation .src), and under
  ``head1/mat/`` the MAT-file goldens Head1.{hm,dsm,h2em,ss2mm,est_eeg}
  unconverted (a symmetric, full, sparse and full matrix as the reference
  writes them) for the file-format tests.
* ``head3/``: the Head3 inputs (642-vertex spheres, same geometry as the
  0.87/0.92/1.0 nested-sphere configuration), .dip and .patches.
* ``headnna1/``: the HeadNNa1 inputs (non-nested: the north and south meshes
  share the cut mesh's vertices), .dip, .patches and .squids.
* ``analytic/``: ``tests/analytic/eeg_head{1,2,3}_analytic.txt`` (multilayer
  sphere series, n_electrodes x 6 dipoles) and the internal-potential points
  and analytic values ``eeg_internal_{points,analytic}.txt``; ``head1/`` also
  holds ``Head1.srcdip`` (the dipole activations) and ``Head1-EIT.patches``
  (EIT electrodes with radii).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = os.environ.get("OPENMEEG_REFERENCE", "/root/reference")
H5DUMP = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")
HERE = os.path.dirname(os.path.abspath(__file__))


def h5_dataset(path, dset, dtype):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "raw.bin")
        subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path],
                       check=True, stdout=subprocess.DEVNULL)
        return np.fromfile(out, dtype=dtype)


def main():
    initial = os.path.join(REF, "tests", "initialTest")
    h1 = os.path.join(HERE, "head1")
    os.makedirs(h1, exist_ok=True)

    for name in ("hm", "hm_inv"):
        path = os.path.join(initial, "Head1." + name)
        size = h5_dataset(path, "/symmatrix/size", "<u4")
        data = h5_dataset(path, "/symmatrix/data", "<f8")
        n = int(size[0])
        assert data.size == n * (n + 1) // 2, (name, n, data.size)
        np.save(os.path.join(h1, name + ".npy"), data)

    for name in ("dsm", "dgem", "h2mm", "ds2mm", "dgmm", "ssm", "sgem", "ss2mm", "sgmm", "est_eeg"):
        path = os.path.join(initial, "Head1." + name)
        raw = h5_dataset(path, "/linop", "<f8")
        # The HDF5 dataset is (ncol, nrow) row-major, i.e. the column-major
        # (nrow, ncol) matrix of the reference; shapes from the header.
        hdr = subprocess.run([H5DUMP, "-H", "-d", "/linop", path],
                             check=True, capture_output=True, text=True).stdout
        dims = hdr.split("SIMPLE { (")[1].split(")")[0]
        ncol, nrow = (int(x) for x in dims.split(","))
        mat = raw.reshape(ncol, nrow).T.copy()
        np.save(os.path.join(h1, name + ".npy"), mat)

    # the MAT-file goldens themselves, for the file-format tests
    mat = os.path.join(h1, "mat")
    os.makedirs(mat, exist_ok=True)
    for name in ("hm", "dsm", "h2em", "ss2mm", "est_eeg"):
        shutil.copy(os.path.join(initial, "Head1." + name), os.path.join(mat, "Head1." + name))

    path = os.path.join(initial, "Head1.h2em")
    for part, dt in (("data", "<f8"), ("ir", "<u8"), ("jc", "<u8")):
        np.save(os.path.join(h1, "h2em_" + part + ".npy"),
                h5_dataset(path, "/matrix/" + part, dt))

    data1 = os.path.join(REF, "data", "Head1")
    for f in ("Head1.geom", "Head1.cond", "Head1.dip", "Head1.patches", "Head1.squids",
              "Head1.tri", "Head1.src", "Head1.srcdip", "Head1-EIT.patches", "cortex.1.tri", "skull.1.tri", "scalp.1.tri"):
        shutil.copy(os.path.join(data1, f), h1)

    h3 = os.path.join(HERE, "head3")
    os.makedirs(h3, exist_ok=True)
    data3 = os.path.join(REF, "data", "Head3")
    for f in ("Head3.geom", "Head3.cond", "Head3.dip", "Head3.patches",
              "cortex.3.tri", "skull.3.tri", "scalp.3.tri"):
        shutil.copy(os.path.join(data3, f), h3)

    hn = os.path.join(HERE, "headnna1")
    os.makedirs(hn, exist_ok=True)
    datan = os.path.join(REF, "data", "HeadNNa1")
    for f in ("HeadNNa1.geom", "HeadNNa1.cond", "HeadNNa1.dip", "HeadNNa1.patches", "HeadNNa1.squids",
              "cut.tri", "north.tri", "south.tri", "skull.tri", "scalp.tri"):
        shutil.copy(os.path.join(datan, f), hn)
    an = os.path.join(HERE, "analytic")
    os.makedirs(an, exist_ok=True)
    for k in (1, 2, 3):
        shutil.copy(os.path.join(REF, "tests", "analytic",
                                 "eeg_head%d_analytic.txt" % k), an)
    for f in ("eeg_internal_points.txt", "eeg_internal_analytic.txt"):
        shutil.copy(os.path.join(REF, "tests", "analytic", f), an)
    print("fixtures written under", HERE)


if __name__ == "__main__":
    sys.exit(main())
