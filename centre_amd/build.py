"""Build the HIP library ``libcentre_hip.so`` in place (gfx950 only).

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled with hipcc for
``--offload-arch=gfx950`` and linked into ``centre_amd/libcentre_hip.so``.
The C++ host (``host/*.cpp``, the drop-in for the centre_OMP executable) is
compiled with g++ against that library's C ABI into ``centre_amd/centre_hip``
(it loads libhdf5 at run time for NetCDF-4 input).
Objects go to ``centre_amd/_build/``.  ``build()`` recompiles every source;
``build(force=False)`` recompiles only sources newer than their object (or
than any header under ``csrc/`` and ``include/``).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcentre_hip.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST_BIN = os.path.join(HERE, "centre_hip")
ARCH = "gfx950"
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def _headers_mtime() -> float:
    hs = glob.glob(os.path.join(CSRC, "*.h")) + \
        glob.glob(os.path.join(os.path.dirname(HERE), "include", "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src: str, force: bool = True) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not force and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(src) and t > _headers_mtime():
            return obj
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, force: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                  glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_host()
    if verbose:
        print(f"built {OUT} and {HOST_BIN}")
    return OUT


def build_host() -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(os.path.dirname(HERE), "include")]
    libs = ["-L" + HERE, "-lcentre_hip", "-Wl,-rpath,$ORIGIN", "-ldl"]
    tmp = HOST_BIN + ".tmp"
    r = subprocess.run(cmd + srcs + ["-o", tmp] + libs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, HOST_BIN)
    return HOST_BIN


if __name__ == "__main__":
    build(verbose=True, force="--changed" not in sys.argv)
    sys.exit(0)
