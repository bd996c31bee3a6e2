"""Build the native parts in-tree.

    python -m score_amd.build            # libscore_gm.so (gfx950), libsynth.so, libscore_io.so, libscore_setup.so
    python -m score_amd.build --oracle   # also oracle/liboracle.so (test infrastructure)

hipcc compiles device and host code of the growth machine for gfx950 only,
with -ffp-contract=off so no multiply-add is fused (the results are compared
bit for bit with a serial FP64 run).
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", INCLUDE,
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gm(out: str | None = None) -> str:
    out = out or os.path.join(HERE, "libscore_gm.so")
    srcs = [os.path.join(CSRC, f) for f in ("gm_kernels.hip", "gm_regions.hip", "gm_noalloc.hip", "gm_host.cpp")]
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_synth(out: str | None = None) -> str:
    out = out or os.path.join(HERE, "libsynth.so")
    cc = shutil.which("gcc") or "cc"
    tmp = out + ".tmp"
    _run([cc, "-O2", "-std=c99", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          os.path.join(CSRC, "synth.c"), "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


HDF5_PREFIX = os.environ.get("SCORE_HDF5_PREFIX", "/opt/conda")


def build_io(out: str | None = None) -> str | None:
    """libscore_io.so: grain-map files in SCORE's formats (raw, HDF5, XDMF); needs the
    HDF5 C library (headers and libhdf5 under HDF5_PREFIX)."""
    out = out or os.path.join(HERE, "libscore_io.so")
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"score_io: no HDF5 headers under {HDF5_PREFIX}; libscore_io.so not built", flush=True)
        return None
    cc = shutil.which("gcc") or "cc"
    tmp = out + ".tmp"
    _run([cc, "-O2", "-std=c99", "-fPIC", "-shared", "-Wall", "-I", inc, "-I", INCLUDE,
          os.path.join(CSRC, "score_io.c"), "-o", tmp, "-L", lib, "-lhdf5", "-Wl,-rpath," + lib])
    os.replace(tmp, out)
    return out


def build_setup(out: str | None = None) -> str:
    """The host setup library (UDS files, SCORE's random stream, cuboid structure,
    CSR nucleation): g++ without -march and with -ffp-contract=off, as the
    reference's GCC build rounds."""
    out = out or os.path.join(HERE, "libscore_setup.so")
    tmp = out + ".tmp"
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wextra", "-I", INCLUDE,
          os.path.join(CSRC, "score_setup.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_caller(out: str | None = None) -> str | None:
    """tests/abi_caller/score_gm_caller: a C++ program compiled against the C
    headers with g++ -std=c++11 (SCORE's language level) and linked to
    libscore_gm.so and libscore_io.so, as SCORE's adapter would be (the ABI
    tests run it)."""
    out = out or os.path.join(ROOT, "tests", "abi_caller", "score_gm_caller")
    lib_io = os.path.join(HERE, "libscore_io.so")
    if not os.path.exists(lib_io):
        print("score_gm_caller: libscore_io.so not built; caller not built", flush=True)
        return None
    cxx = os.environ.get("CXX", "g++")
    tmp = out + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    _run([cxx, "-std=c++11", "-O2", "-Wall", "-Wextra", "-I", INCLUDE,
          os.path.join(ROOT, "tests", "abi_caller", "caller.cpp"), "-o", tmp,
          "-L", HERE, "-lscore_gm", "-lscore_io",
          "-Wl,-rpath,$ORIGIN/../../score_amd",
          "-Wl,-rpath-link," + rocm_lib, "-Wl,-rpath-link," + os.path.join(HDF5_PREFIX, "lib")])
    os.replace(tmp, out)
    return out


def build_radd_device(out: str | None = None) -> str:
    """tests/radd_device/radd_device: the device build of gm_radd.h against
    plain device additions (a GPU test runs it)."""
    out = out or os.path.join(ROOT, "tests", "radd_device", "radd_device")
    flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")]
    tmp = out + ".tmp"
    _run([HIPCC, *flags, "-I", CSRC, os.path.join(ROOT, "tests", "radd_device", "radd_device.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


REFERENCE = "/root/reference"


def build_oracle() -> str:
    """The oracle (test infrastructure); where the reference tree is present,
    also its random-number and orientation code into oracle/_ref (machines
    without the reference tree use the file built here)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "-B", "liboracle.so"])
    if os.path.isdir(os.path.join(REFERENCE, "src")):
        _run(["make", "-s", "-C", odir, "-B", "ref", "REF=" + REFERENCE])
    return os.path.join(odir, "liboracle.so")


def build_all(oracle: bool = True) -> None:
    build_synth()
    build_gm()
    build_io()
    build_setup()
    build_caller()
    build_radd_device()
    if oracle:
        build_oracle()


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true", help="also build the oracle (test infrastructure)")
    ap.add_argument("--only", choices=["gm", "synth", "io", "setup", "caller", "radd", "oracle"], default=None)
    a = ap.parse_args(argv)
    if a.only == "gm":
        build_gm()
    elif a.only == "synth":
        build_synth()
    elif a.only == "io":
        build_io()
    elif a.only == "setup":
        build_setup()
    elif a.only == "caller":
        build_caller()
    elif a.only == "radd":
        build_radd_device()
    elif a.only == "oracle":
        build_oracle()
    else:
        build_all(oracle=a.oracle)
    return 0


if __name__ == "__main__":
    sys.exit(main())
