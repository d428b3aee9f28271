"""The drop-in boundary, proved by compiling and linking against the
reference's own header.

A C++ translation unit includes the reference's gpu_kernels.cuh (its entry
points have C++ linkage, gpu_kernels.cuh:4-93) and taco_tensor_t.h, takes the
address of every function the header declares and calls spmm_rb_sr and
spmm_csr_gpu_taco as the reference's driver does (spmm_csr_gpu.h:283-343).
It is compiled with g++ and linked against libsegtaco.so the way the
reference's Makefile links its driver (Makefile:11-15): the link must succeed
with every reference the object leaves undefined resolved by libsegtaco's
exports. A second translation unit includes the reference's operand struct
and this library's in separate namespaces and static_asserts that size,
alignment and every field offset agree.

Skipped when the reference tree is absent (only the compile and link are
checked here; calls through the C++-linkage names run in the GPU tests).
"""
import os
import re
import subprocess

import pytest

import segtaco_amd as S

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REF, "gpu_kernels.cuh")

pytestmark = pytest.mark.skipif(not os.path.exists(HEADER), reason="reference tree absent")


def _declared_in_reference_header():
    """Names of the functions gpu_kernels.cuh declares outside comments."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return re.findall(r"^\s*(?:void|float)\s+(\w+)\s*\(", text, flags=re.M)


CALLER = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "gpu_kernels.cuh"

typedef void (*op_t)(taco_tensor_t *, taco_tensor_t *, taco_tensor_t *);

// every entry point the header declares (address taken: one undefined
// reference per name in this object)
static const op_t k_ops[] = { OPS };

static taco_tensor_t dense(int32_t *dims, float *v, int32_t rows, int32_t cols) {
  dims[0] = rows; dims[1] = cols;
  taco_tensor_t t{};
  t.order = 2; t.dimensions = dims; t.vals = v; t.vals_size = rows * cols;
  return t;
}

int main(int argc, char **argv) {
  // 3 x 4 CSR times 4 x 2 (as spmm_csr_gpu builds its operands)
  int32_t pos[] = {0, 2, 2, 5}, crd[] = {0, 3, 1, 2, 3};
  float av[] = {1, 2, 3, 4, 5};
  int32_t *idx1[] = {pos, crd};
  int32_t **idx[] = {nullptr, idx1};
  int32_t adims[] = {3, 4};
  taco_tensor_t A{};
  A.order = 2; A.dimensions = adims; A.indices = idx; A.vals = av; A.vals_size = 5;
  std::vector<float> b(8, 1.0f), c(6, -1.0f);
  int32_t bd[2], cd[2];
  taco_tensor_t B = dense(bd, b.data(), 4, 2), C = dense(cd, c.data(), 3, 2);
  const int n = (int)(sizeof(k_ops) / sizeof(k_ops[0]));
  if (argc > 2) k_ops[std::atoi(argv[2]) % n](&C, &A, &B);  // any entry point by index
  if (argc > 1) {  // run: the two entry points the driver uses most
    spmm_rb_sr(&C, &A, &B);
    float ms = get_gpu_timer_result();
    spmm_csr_gpu_taco(&C, &A, &B);
    std::printf("%g %g %g %g %g %g %g\n", c[0], c[1], c[2], c[3], c[4], c[5], ms);
  }
  return 0;
}
"""

LAYOUT = r"""
#include <cstddef>
#include <cstdint>
namespace ref {
#include "taco_tensor_t.h"
}
namespace mine {
#include "segtaco/taco_tensor_t.h"
}
#define SAME(f) static_assert(offsetof(ref::taco_tensor_t, f) == offsetof(mine::taco_tensor_t, f), #f)
static_assert(sizeof(ref::taco_tensor_t) == sizeof(mine::taco_tensor_t), "size");
static_assert(alignof(ref::taco_tensor_t) == alignof(mine::taco_tensor_t), "alignment");
static_assert(sizeof(ref::taco_mode_t) == sizeof(mine::taco_mode_t), "mode enum");
static_assert((int)ref::taco_mode_sparse == (int)mine::taco_mode_sparse, "mode values");
SAME(order); SAME(dimensions); SAME(csize); SAME(mode_ordering); SAME(mode_types);
SAME(indices); SAME(vals); SAME(vals_size);
int layout_ok() { return 1; }
"""


def _run(cmd, cwd):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert res.returncode == 0, f"{' '.join(cmd)}\n{res.stdout}\n{res.stderr}"
    return res.stdout


def _undefined(obj, cwd):
    out = _run(["nm", "-u", obj], cwd)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_reference_header_links_against_libsegtaco(tmp_path):
    names = _declared_in_reference_header()
    ops = [n for n in names if n != "get_gpu_timer_result"]
    assert len(ops) == 79 and "get_gpu_timer_result" in names
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER.replace("OPS", ", ".join(ops)))
    d = str(tmp_path)
    _run(["g++", "-std=c++14", "-O2", "-Wall", f"-I{REF}", "-c", "caller.cpp", "-o",
          "caller.o"], d)
    # the object refers to every entry point by its C++ (mangled) name
    undef = _undefined("caller.o", d)
    mangled = {f"_Z{len(n)}{n}P13taco_tensor_tS0_S0_" for n in ops} | {
        "_Z20get_gpu_timer_resultv"}
    assert mangled <= undef, sorted(mangled - undef)[:5]
    assert mangled <= _exported()
    # the link of the driver (Makefile:11-15) with libsegtaco in place of
    # gpu_kernels.cu: every reference resolves, none is left undefined
    libdir = os.path.dirname(S.LIB_PATH)
    _run(["g++", "caller.o", "-o", "caller", f"-L{libdir}", "-lsegtaco",
          f"-Wl,-rpath,{libdir}", "-Wl,--no-allow-shlib-undefined"], d)
    # load the program with every symbol bound at start-up (no call is made
    # without arguments): each of its references binds to libsegtaco.so
    env = dict(os.environ, LD_BIND_NOW="1", LD_DEBUG="bindings")
    res = subprocess.run(["./caller"], cwd=d, capture_output=True, text=True, env=env)
    assert res.returncode == 0, res.stderr[-2000:]
    bound = set()
    for line in res.stderr.splitlines():
        m = re.search(r"binding file \./caller .* to (\S+) .*symbol `([^']+)'", line)
        if m and m.group(1).endswith("libsegtaco.so"):
            bound.add(m.group(2))
    assert mangled <= bound, sorted(mangled - bound)[:5]


def test_operand_struct_layout_identical(tmp_path):
    src = tmp_path / "layout.cpp"
    src.write_text(LAYOUT)
    _run(["g++", "-std=c++14", "-c", f"-I{REF}", f"-I{os.path.join(ROOT, 'include')}",
          "layout.cpp", "-o", "layout.o"], str(tmp_path))
