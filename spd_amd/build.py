"""Build ``libspd_hip.so`` (the C-ABI library of the HIP kernels) in-tree.

``python -m spd_amd.build`` compiles ``spd_amd/csrc/*.hip`` for gfx950 with
hipcc into ``spd_amd/libspd_hip.so``.  Packed-FP32 instructions are disabled
for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libspd_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "spd_step.hip")]
ARCH = os.environ.get("SPD_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out=LIB):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"),
            "-o", out] + SOURCES


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libspd_hip.so")
    os.replace(tmp, LIB)
    if verbose:
        warn = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if warn:
            print("\n".join(warn[:40]))
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
