"""Plan resolution and workspace sizing on the host (no GPU needed): the
bytes reported for a plan with open fields cover the resolved plan's."""
import numpy as np
import pytest

import segtaco_amd as S


def _desc(M, nnz, K=None, dtype=S.DT_F32, vals=True):
    # pointers are never dereferenced by plan resolution or sizing
    return S.csr_desc(M, K or M, nnz, 0x1000, 0x2000, 0x3000 if vals else None, dtype)


@pytest.mark.parametrize("sched", ["eb_sr", "eb_pr", "eb_cs"])
@pytest.mark.parametrize("nnz", [100, 5000, 70000, 4_000_000])
@pytest.mark.parametrize("N", [1, 3, 16, 128])
@pytest.mark.parametrize("dtype,vals", [(S.DT_F32, True), (S.DT_BF16, True), (S.DT_F32, False)])
def test_open_plan_workspace_covers_resolved(sched, nnz, N, dtype, vals):
    A = _desc(max(nnz // 10, 1), nnz, dtype=dtype, vals=vals)
    open_plan = S.make_plan(sched)
    need_open = S.workspace_bytes(A, N, open_plan)
    b_dtypes = [dtype] if vals else [S.DT_F32, S.DT_BF16, S.DT_F16]
    for bd in b_dtypes:
        for ld in (N, N + 1):
            try:
                p = S.plan_resolve(A, N, bd, open_plan, ld, ld)
            except S.SegtacoError:
                continue
            need = S.workspace_bytes(A, N, p)
            assert need > 0
            assert need_open >= need, (bd, ld, p.as_dict(), need_open, need)


def test_row_balanced_needs_no_workspace():
    A = _desc(1000, 20000)
    for sched in ("rb_sr", "rb_pr"):
        assert S.workspace_bytes(A, 32, S.make_plan(sched)) == 0
