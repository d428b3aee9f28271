"""segtaco_spmm_noalloc: segtaco_spmm with the library's fixed device
workspace instead of a caller buffer (no allocation in the call).

Through the C-ABI, each case runs two calls in a row on one GPU, the second
with a smaller product, and compares every result bit for bit with
segtaco_spmm's on the same plan (the same kernels, deterministic: integer
partitions and a fixed summation order), and the smaller one also with the
FP64 oracle (scaled error <= 1e-5). A product whose workspace exceeds the
fixed array is refused on the host before any device work (CPU test).
"""
import numpy as np
import pytest

import oracle as O
import segtaco_amd as S


def test_oversized_workspace_refused_on_host():
    """eb_sr at tile 8 over 40 000 entries and N = 4096 needs 2 x 5000 x 4096
    floats of carries (164 MB) > the 128 MiB array: refused with
    SEGTACO_ERR_CAPACITY before any device call (no GPU is touched)."""
    A = S.csr_desc(1000, 1000, 40000, 0x10000, 0x20000, 0x30000)
    plan = S.make_plan("eb_sr", tile=8)
    need = S.workspace_bytes(A, 4096, S.plan_resolve(A, 4096, S.DT_F32, plan))
    assert need > S.NOALLOC_WORKSPACE_BYTES
    with pytest.raises(S.SegtacoError) as ei:
        S.spmm_noalloc(A, 0x40000, S.DT_F32, 4096, 4096, 0x50000, 4096, plan)
    assert ei.value.status == S.ERR_CAPACITY
    assert "workspace" in str(ei.value)


def _operands(kind, M, N, seed):
    from gpu_util import DeviceCSR, dense_dev
    if kind == "powerlaw":
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, seed, M, M)
    else:
        pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, seed, M, M, mean=12)
    B = S.gen_dense(seed + 100, M, N)
    return (pos, crd, vals, B), DeviceCSR(pos, crd, vals, M), dense_dev(B)


CASES = [
    # (schedule, plan fields, first product (kind, M, N), second, smaller)
    ("eb_sr", {}, ("powerlaw", 262144, 128), ("powerlaw", 20000, 64)),
    ("eb_sr", dict(tile=16), ("powerlaw", 60000, 40), ("uniform", 3000, 8)),
    ("eb_pr", {}, ("powerlaw", 60000, 128), ("powerlaw", 5000, 32)),
    ("eb_pr", {}, ("powerlaw", 60000, 1), ("uniform", 4000, 1)),
    ("eb_cs", {}, ("powerlaw", 60000, 64), ("powerlaw", 8000, 16)),
    ("rb_sr", {}, ("powerlaw", 60000, 64), ("uniform", 3000, 32)),
    ("rb_pr", {}, ("uniform", 40000, 32), ("uniform", 2000, 4)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("sched,kw,first,second", CASES,
                         ids=[f"{c[0]}-{c[2][0]}{c[2][1]}n{c[2][2]}" for c in CASES])
def test_noalloc_matches_spmm_two_calls(sched, kw, first, second):
    from gpu_util import run_spmm, stream_ptr, torch
    t = torch()
    for call, (kind, M, N) in enumerate((first, second)):
        print(f"[noalloc] {sched} {kw} call {call}: {kind} {M} N={N}", flush=True)
        (pos, crd, vals, B), A, Bd = _operands(kind, M, N, seed=3 + call)
        plan = S.make_plan(sched, **kw)
        ref = run_spmm(A, Bd, N, plan)[: A.M].cpu().numpy()
        C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
        S.spmm_noalloc(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, plan,
                       stream_ptr())
        t.cuda.synchronize()
        got = C.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"call {call}"
        if call == 1:
            err = O.scaled_err(got, O.spmm_f64(pos, crd, vals, B),
                               O.spmm_absbound(pos, crd, vals, B))
            assert err <= 1e-5, err
