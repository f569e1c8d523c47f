"""Builds the HIP library openmeeg_amd/libopenmeeg_amd.so for gfx950, the
host library of the matrix file formats openmeeg_amd/libopenmeeg_amd_io.so
and the command-line tools openmeeg_amd/bin/{om_assemble,om_minverser,om_gain}.

hipcc compiles every csrc/*.hip (device + host) and csrc/*.cpp (host only,
except maths_io.cpp) into openmeeg_amd/_build/, then links the shared
library in-tree so that it travels with the source tree.  maths_io.cpp is
built with g++ against the HDF5 C library (OMH_HDF5_ROOT, default
/opt/conda); the tools link both libraries.  Usage: ``python -m
openmeeg_amd.build``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopenmeeg_amd.so")
IOLIB = os.path.join(HERE, "libopenmeeg_amd_io.so")
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")
HDF5 = os.environ.get("OMH_HDF5_ROOT", "/opt/conda")
IO_SRC = "maths_io.cpp"
TOOLS = ("om_assemble", "om_minverser", "om_gain")
ARCH = os.environ.get("OMH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
# packed FP32 instructions are not emitted (they have given silent wrong
# results on this hardware pool); hipcc's host pass warns about the flag.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + DEVICE + COMMON + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed: %s\n%s%s" % (what, " ".join(cmd), r.stdout, r.stderr))


def build_io(verbose=True):
    """The file-format library (host only, HDF5 for MAT 7.3)."""
    tmp = IOLIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(HDF5, "include"),
          os.path.join(CSRC, IO_SRC), "-o", tmp, "-L", os.path.join(HDF5, "lib"), "-lhdf5",
          "-Wl,-rpath," + os.path.join(HDF5, "lib")], "io library")
    os.replace(tmp, IOLIB)
    if verbose:
        print("built", IOLIB)


def build_tools(verbose=True):
    """om_assemble, om_minverser, om_gain (host code over both
    libraries)."""
    os.makedirs(BIN, exist_ok=True)

    def one(name):
        out = os.path.join(BIN, name)
        _run([HIPCC, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", os.path.join(APPS, name + ".cpp"), "-o",
              out + ".tmp", "-L", HERE, "-lopenmeeg_amd", "-lopenmeeg_amd_io", "-Wl,-rpath,$ORIGIN/.."],
             "tool " + name)
        os.replace(out + ".tmp", out)
        return out

    with cf.ThreadPoolExecutor(max_workers=len(TOOLS)) as ex:
        outs = list(ex.map(one, TOOLS))
    if verbose:
        print("built", " ".join(outs))


def build(verbose=True):
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")) and f != IO_SRC)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_io(verbose)
    build_tools(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
