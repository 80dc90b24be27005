"""Build libfyfft.so (HIP kernels + C ABI) for gfx950 with hipcc.

Every translation unit under fyfft_amd/csrc is compiled to an object in
fyfft_amd/_build (in parallel), then linked into fyfft_amd/lib/libfyfft.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfyfft.so")
ARCH = os.environ.get("FYFFT_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 ops are switched off for all device code: some of their forms
# have produced silently wrong results on the target machines.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _headers(src: str) -> list[str]:
    """Headers a source may include: every header under csrc/, and for
    sources that include the public C ABI header, include/fyfft.h."""
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    with open(src) as f:
        if "fyfft.h" in f.read():
            hdrs.append(os.path.join(ROOT, "include", "fyfft.h"))
    return hdrs


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return t < os.path.getmtime(src) or any(t < os.path.getmtime(h) for h in _headers(src))


def _compile(src: str, extra: tuple = ()) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, *extra, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every source (incremental: only those older than their
    object's inputs -- the source itself or any header) and link."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]
    todo = srcs
    if incremental:
        todo = [s for s, o in zip(srcs, objs) if _stale(s, o)]
    jobs = jobs or min(max(1, len(todo)), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, todo))
    if incremental and not todo and os.path.exists(LIB):
        return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_dropin_test() -> str:
    """Compile tests/cpp/dropin_main.cpp with plain g++ against the C ABI and
    include/foye_FFT.hpp (no HIP headers needed on the caller's side)."""
    exe = os.path.join(LIBDIR, "dropin_main")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-o", exe,
           f"-L{LIBDIR}", "-lfyfft", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in test build failed:\n{r.stderr}")
    return exe


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv))
    print(build_dropin_test())
