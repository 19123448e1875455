"""Build librocbta.so (all HIP sources of the package) for gfx950 with hipcc.

Usage: python -m serinv_amd.csrc.build   (or serinv_amd.csrc.build.build())
The library is written in-tree to serinv_amd/_lib/librocbta.so.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_DIR = os.path.join(PKG, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "librocbta.so")
SOURCES = ["rocbta.hip", "noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the compiler from emitting packed FP32 math (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build librocbta.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(
        ln for ln in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in ln
    )
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, flush=True)


def build(verbose=False):
    """Compile every source to an object (in parallel), then link the library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    hipcc = hipcc_path()
    objs = [os.path.join(LIB_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = [[hipcc, *HIPCC_FLAGS, "-c", "-o", o, os.path.join(HERE, s)]
            for s, o in zip(SOURCES, objs)]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(_run, cmd, verbose) for cmd in jobs]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
