"""Build libadock_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension machinery).

python -m adock_amd.build  [--verbose]
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libadock_hip.so"
SOURCES = ["grid.hip", "dock.hip", "stagea.hip", "mlp.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 (v_pk_*_f32) code generation is switched off for every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build(verbose=False, extra=()):
    cmd = [HIPCC, *FLAGS, *extra, *[str(CSRC / s) for s in SOURCES], "-o", str(LIB) + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose and r.stderr.strip():
        print(r.stderr[-3000:])
    return LIB


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv)
    print(LIB)
