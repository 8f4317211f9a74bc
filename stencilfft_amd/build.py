"""Build libstencilfft.so (gfx950) in-tree with hipcc.

Usage: ``python -m stencilfft_amd.build`` (or ``build()`` from Python).  Object
files go to ``stencilfft_amd/_build/``; the shared library lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(HERE, "libstencilfft.so")
SOURCES = ["plan.hip", "fft.hip", "line_fast.hip", "stencil.hip", "spectral.hip"]
ARCH = "gfx950"

# No packed-FP32 instructions: hipcc would otherwise form v_pk_* FP32 ops from
# float2 arithmetic.
COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


# The stencils reproduce the reference's rounding term by term: no FMA
# contraction anywhere in that translation unit (helpers included).
PER_FILE_FLAGS = {"stencil.hip": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, *PER_FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        # keep real warnings visible, drop the host pass's note about the feature flag
        lines = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    """Compile every HIP source for gfx950 and link libstencilfft.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = list(SOURCES)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    build_example()
    build_hdf5()
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


EXAMPLE_SRC = os.path.join(HERE, "examples", "example.cpp")
EXAMPLE_BIN = os.path.join(HERE, "examples", "zmodel_example")


def build_example() -> str:
    """The C++ drop-in example, linked against libstencilfft.so (rpath $ORIGIN/..)."""
    include = os.path.join(os.path.dirname(HERE), "include")
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", EXAMPLE_SRC, f"-I{include}", f"-L{HERE}", "-lstencilfft",
           "-Wl,-rpath,$ORIGIN/..", "-o", EXAMPLE_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"example build failed:\n{res.stdout}\n{res.stderr}")
    return EXAMPLE_BIN


HDF5_SRC = os.path.join(CSRC, "hdf5_io.cpp")
HDF5_LIB = os.path.join(HERE, "libstencilfft_hdf5.so")


def hdf5_prefix() -> str | None:
    """Install prefix of the serial HDF5 C library: $HDF5_HOME, else the first
    of /opt/conda, /usr/local, /usr holding include/hdf5.h."""
    cands = [os.environ.get("HDF5_HOME", ""), "/opt/conda", "/usr/local", "/usr"]
    for c in cands:
        if c and os.path.exists(os.path.join(c, "include", "hdf5.h")):
            return c
    return None


def build_hdf5() -> str | None:
    """libstencilfft_hdf5.so: the optional HDF5 field I/O (host code only, g++).
    Skipped, with a message, when no HDF5 installation is found."""
    prefix = hdf5_prefix()
    if prefix is None:
        sys.stderr.write("HDF5 not found (set HDF5_HOME): libstencilfft_hdf5.so not built\n")
        return None
    include = os.path.join(os.path.dirname(HERE), "include")
    libdir = os.path.join(prefix, "lib")
    tmp = HDF5_LIB + ".tmp"
    # static libstdc++: the HDF5 prefix may ship an older libstdc++.so that the
    # rpath below would otherwise put ahead of the one the HIP runtime needs
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-static-libstdc++", "-static-libgcc",
           HDF5_SRC, f"-I{include}", f"-I{os.path.join(prefix, 'include')}", f"-L{HERE}", "-lstencilfft", f"-L{libdir}", "-lhdf5",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{libdir}", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"HDF5 I/O build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, HDF5_LIB)
    build_hdf5_example()
    return HDF5_LIB


HDF5_EXAMPLE_SRC = os.path.join(HERE, "examples", "hdf5_example.cpp")
HDF5_EXAMPLE_BIN = os.path.join(HERE, "examples", "zmodel_hdf5_example")


def build_hdf5_example() -> str:
    """A field through the HDF5 file format and the GPU, over the zmodel.hpp
    facade (linked against both libraries, rpath $ORIGIN/..)."""
    include = os.path.join(os.path.dirname(HERE), "include")
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", HDF5_EXAMPLE_SRC, f"-I{include}", f"-L{HERE}",
           "-lstencilfft_hdf5", "-lstencilfft", "-Wl,-rpath,$ORIGIN/..", "-o", HDF5_EXAMPLE_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"HDF5 example build failed:\n{res.stdout}\n{res.stderr}")
    return HDF5_EXAMPLE_BIN


if __name__ == "__main__":
    build()
