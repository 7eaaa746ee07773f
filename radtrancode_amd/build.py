"""Build the gfx950 shared library ``radtrancode_amd/lib/libradg.so``.

``python -m radtrancode_amd.build`` compiles every HIP source with hipcc for
gfx950 only.  Packed-FP32 instructions are switched off for device code
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libradg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["radg_kernels.hip", "radg_api.hip", "radg_server.hip", "radg_oe.hip", "radg_conv.hip"]

DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *DEVICE_FLAGS, "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_HOST = os.path.join(LIBDIR, "radg_host")


def build_fortran(verbose: bool = False) -> None:
    """The ISO_C_BINDING module and a Fortran host linked to libradg.so."""
    if not os.path.exists(FLANG):
        return
    fdir = os.path.join(HERE, "fortran")
    cmds = [
        [FLANG, "-O2", "-fPIC", "-J", LIBDIR, "-c", os.path.join(fdir, "radg_mod.f90"),
         "-o", os.path.join(LIBDIR, "radg_mod.o")],
        [FLANG, "-O2", "-J", LIBDIR, os.path.join(fdir, "radg_host.f90"),
         os.path.join(LIBDIR, "radg_mod.o"), "-L", LIBDIR, "-lradg",
         "-Wl,-rpath,$ORIGIN", "-o", FORTRAN_HOST],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
