"""Build recipe of the native library ``libssplat_hip.so`` (hand-written HIP for gfx950).

Run ``python -m structsplat_amd._build`` (or ``__graft_entry__.build()``) to compile the library
in-tree; the Python package loads it from next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
SOURCES = [os.path.join(PKG_DIR, "csrc", name)
           for name in ("ssplat.hip", "loss.hip", "optim.hip", "field.hip", "codec.hip")]
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
LIB_PATH = os.path.join(PKG_DIR, "libssplat_hip.so")
ARCH = "gfx950"

# -packed-fp32-ops keeps the compiler from emitting packed FP32 (v_pk_*_f32) instructions.
HIPCC_FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def build(verbose: bool = False) -> str:
    """Compile every HIP source into LIB_PATH (always recompiles) and return the path."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE_DIR}", "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "not a recognized feature for this target"
    diag = "\n".join(line for line in (proc.stderr or "").splitlines() if noise not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{diag}")
    if verbose and diag.strip():
        print(diag, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
