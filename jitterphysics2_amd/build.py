"""Build of libjitterhip.so (gfx950) from csrc/ with hipcc, in-tree."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libjitterhip.so")
CLI = os.path.join(PKG, "jh_cli")

# -ffp-contract=off: no FMA contraction, so FP32 results round like the reference's scalar code.
# The packed-FP32 feature is disabled for all device code.
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-o", OUT + ".tmp", os.path.join(CSRC, "jh_lib.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(" ".join(cmd))
    # native driver (host C++ only, links the library next to it)
    cli = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-o", CLI + ".tmp",
           os.path.join(CSRC, "jh_cli.cpp"), "-L" + PKG, "-ljitterhip", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath,$ORIGIN:/opt/rocm/lib"]
    r = subprocess.run(cli, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ (jh_cli) failed ({r.returncode}):\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
