"""Build libadock_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension machinery).

Every source compiles to its own object in parallel (the chain kernels are split into one translation unit per torsion
bucket for that), then one hipcc link makes the shared library.

python -m adock_amd.build  [--verbose] [-j N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libadock_hip.so"
KM_BUCKETS = (8, 16, 24, 32, 40, 48, 56)
SOURCES = (["grid.hip", "dock.hip"] + [f"dock_km{k}.hip" for k in KM_BUCKETS] +
           ["stagea.hip", "mlp.hip", "capi.hip"])
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = ["--offload-arch=gfx950"]
COMPILE = ARCH + ["-O3", "-std=c++17", "-fPIC",
                  # packed FP32 (v_pk_*_f32) code generation is switched off for every kernel
                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                  "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}): {' '.join(cmd)}\n{r.stderr[-6000:]}")
    return r.stderr


def build(verbose=False, extra=(), jobs=None):
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or max(1, min(16, len(os.sched_getaffinity(0))))
    objs = [OBJ / (Path(s).stem + ".o") for s in SOURCES]
    cmds = [[HIPCC, *COMPILE, *extra, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
    if verbose:
        print(f"compiling {len(cmds)} sources with {jobs} jobs:", " ".join(cmds[0]), flush=True)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        logs = list(pool.map(_run, cmds))
    link = [HIPCC, *ARCH, "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(LIB) + ".tmp"]
    logs.append(_run(link))
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose:
        text = "\n".join(l for l in logs if l.strip() and "packed-fp32-ops" not in l)
        if text.strip():
            print(text[-3000:])
    return LIB


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else None
    build(verbose="--verbose" in sys.argv, jobs=j)
    print(LIB)
