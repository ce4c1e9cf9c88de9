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
import re
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


# The runtime reserves a kernel's scratch (private) memory for every wave the
# device can hold at the kernel's occupancy: bytes per lane x 64 lanes x
# waves per SIMD x 4 SIMDs x 256 CUs, and keeps it per hardware queue.
# Reservations of 2 to 4 GiB have failed with HSA_STATUS_ERROR_OUT_OF_RESOURCES
# in a process that had already run other large ones, so the build refuses a
# kernel above 2 GiB.
SCRATCH_LIMIT = 2 << 30
DEVICE_SIMDS = 4 * 256


def kernel_resources(remarks: str) -> dict:
    """{kernel: (scratch bytes per lane, waves per SIMD)} from hipcc's
    kernel-resource-usage remarks."""
    out, name, scratch = {}, None, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name, scratch = m.group(1), None
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            scratch = int(m.group(1))
            continue
        m = re.search(r"Occupancy \[waves/SIMD\]: (\d+)", line)
        if m and name and scratch is not None:
            out[name] = (scratch, int(m.group(1)))
            name = None
    return out


def build(verbose: bool = False) -> Path:
    cmd = [HIPCC, f"--offload-arch={ARCH}", *FLAGS, "-Rpass-analysis=kernel-resource-usage",
           "-o", str(LIB), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with exit code {r.returncode}")
    res = kernel_resources(r.stderr)
    if not res:
        raise RuntimeError("hipcc printed no kernel resource usage")
    over = {k: v for k, v in res.items() if v[0] * 64 * v[1] * DEVICE_SIMDS > SCRATCH_LIMIT}
    if over:
        LIB.unlink(missing_ok=True)
        raise RuntimeError(f"kernels whose scratch exceeds the runtime's limit at their occupancy: {over}")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
