"""Build recipe for libsegtaco (gfx950) and its tools.

Compiles every HIP/C++ source under segtaco_amd/csrc with hipcc for gfx950
into segtaco_amd/libsegtaco.so (in-tree, so it travels with the repository
snapshot), and the driver executable segtaco_amd/matrix-eval-gpu.

    python -m segtaco_amd.build        # or: python segtaco_amd/build.py
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
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsegtaco.so")
DRIVER = os.path.join(HERE, "matrix-eval-gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 (unreliable on this pool).
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-function",
]

LIB_SOURCES = [
    "k_rb.hip", "k_eb.hip", "k_batch.hip", "engine.cpp", "legacy.cpp", "legacy_cxx.cpp",
    "gen.cpp", "autotune.cpp", "batch.cpp", "serve.cpp", "io.cpp", "vendor.cpp",
    "k_sddmm.hip", "sddmm.cpp", "k_cs.hip", "k_spmv.hip", "k_bd.hip", "staging.cpp",
]
# rocSPARSE backs the vendor comparison row only (vendor.cpp).
LIB_LINK = ["-L/opt/rocm/lib", "-lrocsparse", "-Wl,-rpath,/opt/rocm/lib"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32" not in l and l.strip())
    if warn:
        print(f"[{src}]\n{warn}", file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + LIB_LINK
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    drv_src = os.path.join(HERE, "driver", "matrix_main.cpp")
    if os.path.exists(drv_src):
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), drv_src,
               "-o", DRIVER, f"-L{HERE}", "-lsegtaco", "-Wl,-rpath,$ORIGIN"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        if verbose:
            print(f"built {DRIVER}")
    return LIB


def build_oracle(verbose: bool = True) -> None:
    """Build oracle/liboracle.so and, when the reference tree is present,
    oracle/_ref/libsegtaco_ref.so (test infrastructure; see oracle/Makefile)."""
    res = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], capture_output=True,
                         text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(res.stdout.strip())


if __name__ == "__main__":
    build()
    build_oracle()
