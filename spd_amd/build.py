"""Build ``libspd_hip.so`` (the C-ABI library of the HIP kernels) in-tree.

``python -m spd_amd.build`` compiles each ``spd_amd/csrc/*.hip`` translation
unit for gfx950 with hipcc (in parallel) and links them into
``spd_amd/libspd_hip.so``.  Packed-FP32 instructions are disabled for all
device code.
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
SOURCES = [os.path.join(HERE, "csrc", "spd_step.hip"),
           os.path.join(HERE, "csrc", "spd_adjoint.hip")]
ARCH = os.environ.get("SPD_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def object_path(src):
    return os.path.join(HERE, "csrc", os.path.basename(src).replace(".hip", ".o"))


def compile_command(src):
    return [hipcc()] + FLAGS + ["-I", os.path.join(REPO, "include"), "-c", "-o",
                                object_path(src) + ".tmp", src]


def link_command(out=LIB):
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + \
        [object_path(s) for s in SOURCES]


def build(verbose=False):
    t0 = time.time()
    procs = [(src, subprocess.Popen(compile_command(src), stdout=subprocess.PIPE,
                                    stderr=subprocess.PIPE, text=True)) for src in SOURCES]
    errs, warn = [], []
    for src, pr in procs:
        out, err = pr.communicate()
        if pr.returncode != 0:
            errs.append(out + err)
        else:
            os.replace(object_path(src) + ".tmp", object_path(src))
            warn += [l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed for libspd_hip.so")
    tmp = LIB + ".tmp"
    r = subprocess.run(link_command(tmp), capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed for libspd_hip.so")
    os.replace(tmp, LIB)
    if verbose:
        if warn:
            print("\n".join(warn[:40]))
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
