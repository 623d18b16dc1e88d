"""Build the gfx950 shared library libgdremc.so in-tree with hipcc.

The library holds the HIP kernels (csrc/gd_kernels.hip), the host image of a
complex (csrc/gd_image.cpp), the engine and the C ABI entry points
(csrc/gd_engine.cpp), the serving loop (csrc/gd_server.cpp), shared host utilities
(csrc/gd_host.cpp), the host preparation / parameter parsing
(csrc/gd_prep.cpp) and the input-file loaders (csrc/gd_load.cpp); then the
command-line front end gd_dock (csrc/gd_dock.cpp) beside it.  All device code is compiled with contraction off
(the canonical arithmetic relies on it) and without packed FP32 ops.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgdremc.so")
DOCK = os.path.join(HERE, "gd_dock")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    """Compile every source and link libgdremc.so; returns the library path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, dev in (("gd_kernels.hip", True), ("gd_image.cpp", False), ("gd_engine.cpp", False),
                     ("gd_server.cpp", False), ("gd_host.cpp", False), ("gd_prep.cpp", False),
                     ("gd_load.cpp", False)):
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE if dev else [f"--offload-arch={ARCH}"]) + [
            "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the command-line front end, linked against the library beside it
    tmp = DOCK + ".tmp"
    _run([HIPCC] + COMMON + ["-I", CSRC, os.path.join(CSRC, "gd_dock.cpp"), "-o", tmp,
                             "-L", HERE, "-lgdremc", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DOCK)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
