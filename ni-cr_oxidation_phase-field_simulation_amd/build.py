"""Build libnicr.so (HIP, gfx950) and the drop-in host programs in-tree.

    python ni-cr_oxidation_phase-field_simulation_amd/build.py

The device code is compiled for gfx950 only. -ffp-contract=off keeps the binary64
step bitwise equal to the reference (the mixed-precision kernels enable
contraction locally with `#pragma clang fp contract(fast)`). Packed FP32 ops are
disabled for every HIP translation unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function", "-Wno-unused-variable",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fno-semantic-interposition", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread"]


def _run(cmd, quiet):
    if not quiet:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(quiet: bool = True) -> None:
    lib = os.path.join(PKG, "libnicr.so")
    tmp = lib + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "nicr_api.hip", "-o", tmp], quiet)
    os.replace(tmp, lib)
    # host side of the drop-in programs (g++): initial state, writers, CLI
    hlib = os.path.join(PKG, "libnicr_host.so")
    _run(["g++", *HOST_FLAGS, "-shared", "host/nicr_host.cpp", "-o", hlib + ".tmp",
          f"-L{PKG}", "-lnicr", "-Wl,-rpath,$ORIGIN", "-lz"], quiet)
    os.replace(hlib + ".tmp", hlib)
    for dim in (2, 3):
        exe = os.path.join(PKG, f"nicr{dim}d")
        _run(["g++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-DNICR_CLI_DIM={dim}",
              "host/nicr_cli.cpp", "-o", exe + ".tmp",
              f"-L{PKG}", "-lnicr_host", "-lnicr", "-Wl,-rpath,$ORIGIN",
              "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-lz"], quiet)
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    build(quiet=False)
