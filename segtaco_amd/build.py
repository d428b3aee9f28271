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

# (source, object name, extra flags): the batched and element-balanced
# kernels are compiled once per storage type (and, for the batch, per N = 1
# form), so those objects build in parallel instead of in one long pass.
LIB_SOURCES = (
    [(s, os.path.splitext(s)[0], []) for s in (
        "k_rb.hip", "k_eb.hip", "k_batch.hip", "engine.cpp", "legacy.cpp", "legacy_cxx.cpp",
        "gen.cpp", "autotune.cpp", "batch.cpp", "serve.cpp", "io.cpp", "vendor.cpp",
        "k_sddmm.hip", "sddmm.cpp", "k_cs.hip", "k_spmv.hip", "k_bd.hip", "staging.cpp",
        "noalloc.hip")]
    + [("k_batch_inst.hip", f"k_batch_{dt}_{sv}",
        [f"-DSEGTACO_BATCH_DT={dt}", f"-DSEGTACO_BATCH_SV={sv}"])
       for dt in range(3) for sv in range(2)]
    + [("k_eb_inst.hip", f"k_eb_{dt}", [f"-DSEGTACO_EB_DT={dt}"]) for dt in range(3)]
)
# rocSPARSE backs the vendor comparison row only (vendor.cpp).
LIB_LINK = ["-L/opt/rocm/lib", "-lrocsparse", "-Wl,-rpath,/opt/rocm/lib"]


def _stale(obj: str) -> bool:
    """Whether obj is older than any source or header under csrc/ or include/
    (used only by incremental developer builds)."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    for d in (CSRC, os.path.join(ROOT, "include", "segtaco")):
        for f in os.listdir(d):
            if os.path.getmtime(os.path.join(d, f)) > t:
                return True
    return False


def _compile(entry, incremental: bool = False) -> str:
    src, name, extra = entry
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, name + ".o")
    if incremental and not _stale(obj):
        return obj
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + COMMON + extra + lang + ["-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32" not in l and l.strip())
    if warn:
        print(f"[{src}]\n{warn}", file=sys.stderr)
    return obj


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every source and link libsegtaco.so and the driver.
    incremental=True (developer use) recompiles only objects older than a
    source or header; the default compiles everything."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(8, os.cpu_count() or 4)
    # longest compiles first, so the pool is not left waiting on one of them
    order = sorted(LIB_SOURCES, key=lambda e: not e[0].endswith("_inst.hip"))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda e: _compile(e, incremental), order))
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
    build(incremental="--incremental" in sys.argv)
    build_oracle()
