"""Builds ``libcadycore.so`` (the C-ABI library with the gfx950 kernels)
in-tree with hipcc.  Packed-FP32 instruction selection is disabled for all
device code (``-target-feature -packed-fp32-ops``)."""
from __future__ import annotations

import os
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(_HERE)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
LIB = os.path.join(_HERE, "libcadycore.so")
HOST_LIB = os.path.join(_HERE, "libcadycore_host.so")
HOST_SOURCES = ["ca_dss_table.cpp", "ca_host_tables.cpp"]
IO_LIB = os.path.join(_HERE, "libcaio.so")
IO_SOURCES = ["ca_io.cpp"]
HDF5_ROOT = os.environ.get("CA_HDF5_ROOT", "/opt/conda")
SOURCES = ["ca_unity.hip"]
ARCH = os.environ.get("CA_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def command(out=LIB, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    return [
        hipcc(), "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-I", INCLUDE, "-I", CSRC, "-o", out, *extra, *srcs,
    ]


def host_command(out=HOST_LIB):
    """The host-only table library the CPU tests load (g++, no HIP)."""
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    return ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", CSRC, "-o", out, *srcs]


def io_command(out=IO_LIB):
    """The checkpoint I/O library (host C++ over libhdf5)."""
    srcs = [os.path.join(CSRC, s) for s in IO_SOURCES]
    lib = os.path.join(HDF5_ROOT, "lib")
    return ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
            "-I", os.path.join(HDF5_ROOT, "include"), "-o", out, *srcs,
            f"-L{lib}", f"-Wl,-rpath,{lib}", "-lhdf5", "-static-libstdc++", "-static-libgcc"]


def have_hdf5() -> bool:
    return os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")) and any(
        os.path.exists(os.path.join(HDF5_ROOT, "lib", n)) for n in ("libhdf5.so", "libhdf5.so.103"))


def _run(cmd, out, verbose):
    tmp = out + ".tmp"
    cmd = [tmp if c == out else c for c in cmd]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    _run(command(LIB), LIB, verbose)
    _run(host_command(HOST_LIB), HOST_LIB, verbose)
    if have_hdf5():
        _run(io_command(IO_LIB), IO_LIB, verbose)
    else:
        print(f"libcaio.so not built: no libhdf5 under {HDF5_ROOT} (set CA_HDF5_ROOT)", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
