"""TEST INFRASTRUCTURE ONLY: builds the reference's own host (C++/OpenMP) kernels into oracle/_ref/.

The reference keeps its search and grid kernels as C++ source text inside Python modules and compiles them for the
host with g++ and a small compatibility header (src/adock/cpukernel.py:37-58,80-142). This recipe does the same from
the files where they lie under the reference checkout, without importing any of its Python:
  * reads src/adock/mckernel.py (`_SRC`), src/adock/grid.py (`_KERNEL_SRC`) and src/adock/cpukernel.py (`_COMPAT` and
    the raw-string tiled grid launcher) as text;
  * applies the reference's documented substitutions (cpukernel.py:82-88): __AM__ / __KM__ -> the bucket, __NW__ -> 1,
    and the 17 lane-stride rewrites `x += 32)` -> `x += 1)`;
  * appends one OpenMP launcher per kernel, `ref_<kernel>(long long n_blocks, <kernel params>)`, a loop over blocks
    with blockIdx.x set, as the reference's generated launchers do (cpukernel.py:97-101);
  * pipes the result to g++ on stdin (no source file is written) with the reference's flag set (cpukernel.py:33), except
    -march=x86-64-v3 instead of -march=native so the library also runs on another host.
Output: oracle/_ref/libadock_ref_am<AM>_km<KM>.so (git-ignored). Only tests/ and bench.py's CPU-baseline leg load it.

python oracle/build_ref.py [--ref /root/reference] [AM:KM ...]   (default: BUCKETS below)
"""
import argparse
import os
import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_ref"
DEFAULT_REF = Path(os.environ.get("ADOCK_REFERENCE", "/root/reference"))
# the buckets the checks use: one atom per lane at every torsion bucket to 32, two atoms per lane, and past 32 torsions
BUCKETS = [(64, 8), (64, 16), (64, 24), (64, 32), (64, 40), (96, 16), (128, 24), (128, 48), (128, 56)]
CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "-march=x86-64-v3", "-ffast-math", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-w"]
LANE_REWRITES = 17   # the count the reference asserts (cpukernel.py:87-88)


def _raw_string(text, name):
    """Body of the module-level raw string `name = r\"\"\"...\"\"\"` in a Python source text."""
    m = re.search(name + r'\s*=\s*r"""(.*?)"""', text, re.S)
    if not m:
        raise RuntimeError(f"raw string {name} not found")
    return m.group(1)


def _tiled_launcher(text):
    """The raw-string block of cpukernel.py that defines launch_build_grid_tiled / set_threads / get_threads."""
    for m in re.finditer(r'r"""(.*?)"""', text, re.S):
        if "launch_build_grid_tiled" in m.group(1):
            return m.group(1)
    raise RuntimeError("tiled grid launcher not found")


_SIG = re.compile(r'extern "C" __global__ void (\w+)\((.*?)\)\s*\{', re.S)


def _params(sig):
    out = []
    for p in sig.split(","):
        p = " ".join(p.replace("__restrict__", "").split())
        name = re.findall(r"(\w+)$", p)[0]
        out.append((p[: -len(name)].strip(), name))
    return out


def source(ref, am, km):
    """C++ text of the host build for bucket (am, km)."""
    src_dir = Path(ref) / "src" / "adock"
    mk = (src_dir / "mckernel.py").read_text()
    gr = (src_dir / "grid.py").read_text()
    ck = (src_dir / "cpukernel.py").read_text()
    body = _raw_string(mk, "_SRC").replace("__AM__", str(am)).replace("__KM__", str(km)).replace("__NW__", "1")
    body, n = re.subn(r"(\b\w+) \+= 32\)", r"\1 += 1)", body)
    if n != LANE_REWRITES:
        raise RuntimeError(f"lane-stride rewrite count {n} != {LANE_REWRITES}: the reference text changed")
    body += "\n" + _raw_string(gr, "_KERNEL_SRC")
    launchers = []
    for name, sig in _SIG.findall(body):
        if name == "build_grid":
            continue
        ps = _params(sig)
        decl = ", ".join(f"{t} {v}" for t, v in ps)
        call = ", ".join(v for _, v in ps)
        launchers.append(f'extern "C" void ref_{name}(long long n_blocks, {decl}) {{\n'
                         f"#pragma omp parallel for schedule(dynamic, 1)\n"
                         f"  for (long long b = 0; b < n_blocks; ++b) {{ blockIdx.x = b; {name}({call}); }}\n}}\n")
    return _raw_string(ck, "_COMPAT") + body + "\n" + "".join(launchers) + _tiled_launcher(ck)


def lib_path(am, km):
    return OUT / f"libadock_ref_am{am}_km{km}.so"


def build(ref=DEFAULT_REF, buckets=BUCKETS, force=False, verbose=False):
    """Compile every bucket whose library is missing (all with force). Returns the built paths; raises if the
    reference checkout is absent and a library is missing."""
    OUT.mkdir(exist_ok=True)
    done = []
    for am, km in buckets:
        so = lib_path(am, km)
        if so.exists() and not force:
            done.append(so)
            continue
        if not (Path(ref) / "src" / "adock" / "mckernel.py").exists():
            raise FileNotFoundError(f"reference checkout not found at {ref}")
        tmp = str(so) + ".tmp"
        cmd = [CXX, *FLAGS, "-x", "c++", "-", "-o", tmp]
        if verbose:
            print(" ".join(cmd), f"< reference kernels (AM={am}, KM={km})", flush=True)
        r = subprocess.run(cmd, input=source(ref, am, km), capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on the reference kernels ({am}, {km}):\n{r.stderr[-4000:]}")
        os.replace(tmp, so)
        done.append(so)
    return done


def available(am=64, km=8):
    return lib_path(am, km).exists()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default=str(DEFAULT_REF))
    ap.add_argument("buckets", nargs="*")
    a = ap.parse_args()
    bk = [tuple(map(int, b.split(":"))) for b in a.buckets] or BUCKETS
    for p in build(a.ref, bk, force=True, verbose=True):
        print(p)
    sys.exit(0)
