"""Memory-safety fuzz of the MatrixMarket reader and COO -> CSR (host code).

Builds segtaco_amd/csrc/io.cpp with AddressSanitizer and
UndefinedBehaviorSanitizer (host only, no HIP runtime: tests/asan/error_stub.cpp
stands in for the library's error recorder) together with the driver
tests/asan/mtx_fuzz.cpp, which writes random valid and mutated MatrixMarket
files and random COO triplets. Every input must be either accepted with a
structurally valid result or rejected with an error status; any sanitizer
report aborts the driver.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def fuzz_binary(tmp_path_factory):
    if not os.path.exists(CLANG) or shutil.which("ld") is None:
        pytest.skip("no host clang++ for the sanitizer build")
    out = str(tmp_path_factory.mktemp("asan") / "mtx_fuzz")
    cmd = [CLANG, "-x", "c++", "-std=c++17", "-O1", "-g", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "segtaco_amd", "csrc"),
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ROOT, "segtaco_amd", "csrc", "io.cpp"),
           os.path.join(HERE, "asan", "error_stub.cpp"),
           os.path.join(HERE, "asan", "mtx_fuzz.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-4000:])
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mtx_reader_survives_mutated_files(fuzz_binary, tmp_path, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([fuzz_binary, "4000", str(seed), str(tmp_path)], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    accepted, rejected = (int(x) for x in r.stdout.split()[1::2])
    assert accepted > 0 and rejected > 0
