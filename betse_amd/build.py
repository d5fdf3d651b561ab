"""Build ``libbetse_amd.so`` in-tree with hipcc for gfx950."""

from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "betse_step.hip"), os.path.join(HERE, "csrc", "gnn.hip")]
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "libbetse_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the reference's rounding: no a*b+c contraction into FMA
    "-ffp-contract=off",
    # no packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source into ``betse_amd/libbetse_amd.so`` (always rebuilds)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE] + SOURCES + ["-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    if verbose and (res.stdout or res.stderr):
        print(res.stdout + res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
