"""eb_cs (column-swept element-balanced tiles) on the GPU.

eb_cs changes only the order in which a tile's entries are visited (ascending
column instead of CSR order) and where interior rows accumulate (LDS instead
of registers); every row still sees its entries in CSR order for
column-sorted CSR. So besides the oracle bound, its C must be bit-identical
to eb_sr with the same tile size — including units whose interior rows
overflow the LDS accumulators (their waves walk in CSR order), units where
some waves sweep and others walk, columns too wide for the packed records,
launches with several units per workgroup, 16-bit storage and the full
power-law workload.
"""
import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _pair(pos, crd, vals, B, K=None, dtype="f32", **kw):
    """(eb_cs C, eb_sr C with the same tile, resolved eb_cs plan)."""
    K = B.shape[0] if K is None else K
    A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
    N = B.shape[1]
    Bd = dense_dev(B, dtype)
    bd = S.DTYPES[dtype]
    p = S.plan_resolve(A.desc, N, bd, S.make_plan("eb_cs", **kw))
    C_cs = run_spmm(A, Bd, N, p, b_dtype=dtype)[: A.M].cpu().numpy()
    p_sr = S.make_plan("eb_sr", tile=p.tile, vec=p.vec, comp_threshold=0)
    C_sr = run_spmm(A, Bd, N, p_sr, b_dtype=dtype)[: A.M].cpu().numpy()
    return C_cs, C_sr, p


def _bound(pos, crd, vals, B, C):
    ref = O.spmm_f64(pos, crd, vals, B)
    scale = O.spmm_absbound(pos, crd, vals, B)
    return O.scaled_err(C, ref, scale)


def _mk(rng, lens, K, N):
    lens = np.minimum(lens, K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    return pos, crd, vals, B


@pytest.mark.parametrize("kw", [{}, dict(tile=64), dict(tile=512, group=16), dict(vec=2),
                                dict(tile=96, group=4)])
def test_bit_identical_to_eb_sr_powerlaw(kw):
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 5, 30000, 30000)
    B = S.gen_dense(6, 30000, 96)
    C_cs, C_sr, p = _pair(pos, crd, vals, B, **kw)
    print(f"[cs] plan {p}", flush=True)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_long_rows_empty_runs_and_partial_panel():
    rng = np.random.default_rng(8)
    lens = rng.integers(0, 20, 6000)
    lens[100:900] = 0
    lens[2000] = 15000
    lens[-50:] = 0
    pos, crd, vals, B = _mk(rng, lens, 20000, 44)  # 44 columns: last panel partial
    C_cs, C_sr, _ = _pair(pos, crd, vals, B, tile=128)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_lds_overflow_units_walk():
    """One entry per row: a unit of 64 tiles x 512 entries holds 32768 interior
    rows, far beyond the LDS accumulators (1216 slots at vec 4), so every wave
    walks its tiles in CSR order; results are unchanged."""
    rng = np.random.default_rng(4)
    M = 70000
    pos = np.arange(M + 1, dtype=np.int32)
    crd = rng.integers(0, 5000, M).astype(np.int32)
    vals = rng.uniform(-1, 1, M).astype(np.float32)
    B = rng.uniform(-1, 1, (5000, 32)).astype(np.float32)
    C_cs, C_sr, _ = _pair(pos, crd, vals, B, tile=512)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_units_with_sweeping_and_walking_waves():
    """Short rows first, long rows after: with 128-entry tiles a wave of 8
    tiles over one-entry rows holds 1024 interior rows, which fit the 1088
    interior slots once, so in those units wave 0 sweeps and the rest walk;
    the units over long rows all sweep, and the unit at the boundary mixes
    both kinds of wave."""
    rng = np.random.default_rng(12)
    lens = np.concatenate([np.ones(60000, np.int64), rng.integers(20, 60, 6000)])
    pos, crd, vals, B = _mk(rng, lens, 30000, 64)
    C_cs, C_sr, p = _pair(pos, crd, vals, B, tile=128)
    assert p.vec == 4
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_wide_columns_unpacked_records():
    """K above 2^20: columns do not fit the packed record, so records take
    three words; rows of one entry make some waves walk there too."""
    rng = np.random.default_rng(13)
    K = (1 << 20) + 4099
    lens = np.concatenate([rng.integers(0, 30, 8000), np.ones(20000, np.int64)])
    pos, crd, vals, B = _mk(rng, lens, K, 8)
    crd[-5:] = K - 1 - np.arange(5)[::-1]  # the last columns of B in use
    C_cs, C_sr, p = _pair(pos, crd, vals, B, tile=256)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_several_units_per_workgroup():
    """Small tiles on a 3.2M-entry matrix: ~1560 units for 256 workgroups."""
    pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, 7, 200000, 200000, mean=16)
    B = S.gen_dense(8, 200000, 32)
    C_cs, C_sr, p = _pair(pos, crd, vals, B, tile=32)
    assert p.tile == 32
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_sixteen_bit_storage(dtype):
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 9, 20000, 20000)
    B = S.gen_dense(10, 20000, 128)
    C_cs, C_sr, p = _pair(pos, crd, vals, B, dtype=dtype)
    assert p.vec == 4
    vr, Br = O.round_to(dtype, vals)[1], O.round_to(dtype, B)[1]
    assert _bound(pos, crd, vr, Br, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_full_size_powerlaw():
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    B = S.gen_dense(3, 262144, 128)
    C_cs, C_sr, p = _pair(pos, crd, vals, B)
    print(f"[cs full] plan {p}", flush=True)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


@pytest.mark.parametrize("sched", ["eb_cs", "eb_sr"])
def test_prepared_calls_match_full_calls(sched):
    """One preparation, then run-only calls with two different B: each C is
    bit-identical to a full call. Empty rows are zeroed by every run (C starts
    as NaN). A full call with another matrix on the same workspace replaces
    the preparation, after which a run-only call for the first matrix is
    refused (eb_cs), and a fresh preparation makes it valid again."""
    t = torch()
    rng = np.random.default_rng(21)
    lens = rng.integers(0, 40, 30000)
    lens[::7] = 0
    pos, crd, vals, B1 = _mk(rng, lens, 30000, 64)
    B2 = rng.uniform(-1, 1, B1.shape).astype(np.float32)
    A = DeviceCSR(pos, crd, vals, 30000)
    N = 64
    p = S.plan_resolve(A.desc, N, S.DT_F32, S.make_plan(sched), N, N)
    need = S.workspace_bytes(A.desc, N, p)
    ws = t.empty(need, dtype=t.uint8, device="cuda")
    S.spmm_prepare(A.desc, S.DT_F32, N, N, N, p, ws.data_ptr(), need, stream_ptr())
    for B in (B1, B2):
        Bd = dense_dev(B)
        C_full = run_spmm(A, Bd, N, p)[: A.M].cpu().numpy()
        C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
        S.spmm_prepared(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, p,
                        ws.data_ptr(), need, stream_ptr())
        t.cuda.synchronize()
        C = C.cpu().numpy()
        assert _bound(pos, crd, vals, B, C) <= TOL
        assert np.array_equal(C.view(np.uint32), C_full.view(np.uint32))
    # another product through the same workspace
    pos2, crd2, vals2, B3 = _mk(rng, rng.integers(0, 40, 30000), 30000, 64)
    A2 = DeviceCSR(pos2, crd2, vals2, 30000)
    p2 = S.plan_resolve(A2.desc, N, S.DT_F32, S.make_plan(sched), N, N)
    need2 = S.workspace_bytes(A2.desc, N, p2)
    ws2 = ws if need2 <= need else t.empty(need2, dtype=t.uint8, device="cuda")
    C2 = t.empty((A2.M, N), dtype=t.float32, device="cuda")
    Bd = dense_dev(B3)
    S.spmm(A2.desc, Bd.data_ptr(), S.DT_F32, N, N, C2.data_ptr(), N, p2, ws2.data_ptr(),
           max(need, need2), stream_ptr())
    C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
    if sched == "eb_cs" and ws2 is ws:
        with pytest.raises(S.SegtacoError, match="not prepared"):
            S.spmm_prepared(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, p,
                            ws.data_ptr(), need, stream_ptr())
    S.spmm_prepare(A.desc, S.DT_F32, N, N, N, p, ws.data_ptr(), need, stream_ptr())
    S.spmm_prepared(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, p,
                    ws.data_ptr(), need, stream_ptr())
    t.cuda.synchronize()
    C_full = run_spmm(A, Bd, N, p)[: A.M].cpu().numpy()
    assert np.array_equal(C.cpu().numpy().view(np.uint32), C_full.view(np.uint32))


def test_batch_refuses_eb_cs():
    t = torch()
    pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, 1, 500, 500, mean=8)
    A = DeviceCSR(pos, crd, vals, 500)
    B = dense_dev(S.gen_dense(2, 500, 16))
    C = t.empty((500, 16), dtype=t.float32, device="cuda")
    item = S.batch_item(A.desc, B.data_ptr(), 16, C.data_ptr(), 16, 16, S.make_plan("eb_cs"))
    with pytest.raises(S.SegtacoError, match="standalone"):
        S.Batch([item])
