"""Build libfopis_hip.so for gfx950 in-tree with hipcc.

Every HIP unit is compiled with packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops). Objects go to
fopis_amd/_build/, the library to fopis_amd/libfopis_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfopis_hip.so")
ARCH = os.environ.get("FOPIS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["capi.cpp", "particles.hip", "fields.hip", "poisson.hip", "objects.hip", "radix.hip", "guess.hip", "host.cpp"]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON]
    if src.endswith(".cpp") and src != "host.cpp":
        cmd += ["-x", "hip"]
    if src == "host.cpp":
        # the host draws keep the reference's arithmetic: every product and sum rounded separately
        cmd += ["-ffp-contract=off"]
    cmd += ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_io(verbose)
    return LIB


IO_LIB = os.path.join(HERE, "libfopis_io.so")
HDF5_PREFIX = os.environ.get("FOPIS_HDF5_PREFIX", "/opt/conda")


def build_io(verbose: bool = True) -> str | None:
    """libfopis_io.so (resume and capacity-matrix files in HDF5) when the HDF5 C
    library is present; the step itself does not need it."""
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"HDF5 not found under {HDF5_PREFIX}: libfopis_io.so not built")
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           "-I", inc, os.path.join(SRC, "h5io.cpp"), "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile h5io.cpp failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    if verbose:
        print(f"built {IO_LIB}")
    return IO_LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
