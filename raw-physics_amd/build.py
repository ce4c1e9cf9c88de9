"""Builds librawphys.so (HIP for gfx950) in-tree with hipcc.

Flags that matter for results:
  -ffp-contract=off   no FMA contraction anywhere: the step is bit-exact with
                      the reference only if every multiply and add rounds
                      separately, as in the reference's CPU build.
  no -ffast-math      IEEE division and square root (correctly rounded on the
                      device; checked by the GPU tests).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "librawphys.so"
SOURCES = [HERE / "csrc" / "rp_kernels.hip", HERE / "csrc" / "rp_large.hip", HERE / "csrc" / "rp_adjoint.hip",
           HERE / "csrc" / "rp_world.cpp", HERE / "csrc" / "rp_refsyms.cpp", HERE / "csrc" / "rp_obj.cpp"]
HEADERS = [HERE / "csrc" / "rp_kernels.h", HERE / "csrc" / "rp_device.h", HERE / "csrc" / "rp_dual.h",
           HERE / "csrc" / "rp_step.h",
           HERE / "csrc" / "rp_math.h",
           HERE.parent / "include" / "rawphys.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RP_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         # packed FP32 instructions off (this build has no FP32 math on the hot path anyway)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> Path:
    cmd = [HIPCC, f"--offload-arch={ARCH}", *FLAGS, "-o", str(LIB), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with exit code {r.returncode}")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
