"""GPU parity of the five schedules against the CPU oracle, through the C-ABI.

Tolerances (stated on the metric the oracle defines, oracle.scaled_err:
max |C - ref| / sum_p |v_p| |B[crd_p, k]|):
  * FP32 storage: <= 1e-5 against the FP64 closed form and against the
    Eigen-order FP32 product (the reference's CPU path).
  * Partition arrays: bit-exact.
  * Every schedule is atomic-free: two runs are bit-identical.
"""
import glob
import os
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu

TOL_F32 = 1e-5
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))

PLANS = [
    ("rb_sr", {}), ("rb_sr", dict(group=4)), ("rb_sr", dict(vec=1)),
    ("rb_pr", {}), ("rb_pr", dict(group=8)), ("rb_pr", dict(col_lanes=4, group=4)),
    ("rb_pr", dict(col_lanes=1, group=32)),
    ("eb_sr", {}), ("eb_sr", dict(tile=8)), ("eb_sr", dict(tile=33)), ("eb_sr", dict(tile=1000)),
    ("eb_pr", {}), ("eb_pr", dict(group=32, tile=64)), ("eb_pr", dict(group=4, tile=512)),
    ("eb_pr", dict(group=64, tile=8)),
    ("eb_cs", {}), ("eb_cs", dict(tile=32)), ("eb_cs", dict(tile=512, group=8)),
    ("eb_cs", dict(vec=1, tile=40)),
]
PLAN_IDS = [f"{s}-" + "-".join(f"{k}{v}" for k, v in kw.items()) for s, kw in PLANS]


def _check(pos, crd, vals, B, C, tol=TOL_F32, what=""):
    ref = O.spmm_f64(pos, crd, vals, B)
    scale = O.spmm_absbound(pos, crd, vals, B)
    err = O.scaled_err(C, ref, scale)
    assert err <= tol, f"{what}: scaled error {err:.3e} > {tol:.1e}"
    eig = O.spmm_f32_seq(pos, crd, vals, B)
    err2 = O.scaled_err(C, eig.astype(np.float64), scale)
    assert err2 <= tol, f"{what}: vs Eigen order {err2:.3e}"
    return err


def _device_run(pos, crd, vals, B, sched, kw, K=None, runs=1):
    K = B.shape[0] if K is None else K
    A = DeviceCSR(pos, crd, vals, K)
    N = B.shape[1]
    Bd = dense_dev(B)
    outs = []
    for _ in range(runs):
        C = run_spmm(A, Bd, N, S.make_plan(sched, **kw))
        outs.append(C[: A.M].cpu().numpy())
    return outs


# ---------------------------------------------------------------------------
# Operand families
# ---------------------------------------------------------------------------

def _case(name):
    if name.startswith("golden:"):
        d = np.load(name.split(":", 1)[1], allow_pickle=False)
        vals = d["vals"] if int(d["has_vals"]) else None
        return d["pos"], d["crd"], vals, d["B"]
    if name == "uniform_4096_n32":
        pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, 1, 4096, 4096, mean=16)
        return pos, crd, vals, S.gen_dense(11, 4096, 32)
    if name == "powerlaw_20000_n128":
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 20000, 20000)
        return pos, crd, vals, S.gen_dense(12, 20000, 128)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "empty_runs_n20":
        M, K = 3000, 500
        lens = rng.integers(1, 12, M)
        for s in rng.integers(0, M - 60, 25):
            lens[s: s + rng.integers(1, 60)] = 0
        lens[:17] = 0
        lens[-23:] = 0
        return _mk(rng, lens, K, 20)
    if name == "one_long_row_n130":
        lens = np.array([0, 3, 20000, 0, 5, 9000, 1], np.int64)
        return _mk(rng, lens, 25000, 130)
    if name == "spmv_n1":
        lens = rng.integers(0, 40, 5000)
        return _mk(rng, lens, 4000, 1)
    if name == "n3_odd":
        lens = rng.integers(0, 30, 2000)
        return _mk(rng, lens, 2000, 3)
    if name == "n512_panels":
        lens = rng.integers(0, 24, 1500)
        return _mk(rng, lens, 1500, 512)
    if name == "single_row":
        return _mk(rng, np.array([777]), 1000, 64)
    if name == "all_empty_n8":
        return _mk(rng, np.zeros(100, np.int64), 50, 8)
    if name == "sparse_rows_n8":
        # one row in ~12 holds entries: every eb tile spans more rows than the
        # staged row-pointer window (the kernels' out-of-window lookups)
        lens = np.where(rng.random(20000) < 0.08, rng.integers(1, 4, 20000), 0)
        return _mk(rng, lens, 3000, 8)
    if name == "novals_n16":
        pos, crd, _, B = _mk(rng, rng.integers(0, 50, 800), 900, 16)
        return pos, crd, None, B
    raise KeyError(name)


def _mk(rng, lens, K, N):
    lens = np.minimum(lens, K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    return pos, crd, vals, B


CASES = ([f"golden:{p}" for p in GOLDEN] +
         ["uniform_4096_n32", "powerlaw_20000_n128", "empty_runs_n20", "one_long_row_n130",
          "spmv_n1", "n3_odd", "n512_panels", "single_row", "all_empty_n8", "novals_n16",
          "sparse_rows_n8"])
CASE_IDS = [c.split("/")[-1] for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=CASE_IDS)
@pytest.mark.parametrize("sched,kw", PLANS, ids=PLAN_IDS)
def test_schedule_parity(case, sched, kw):
    pos, crd, vals, B = _case(case)
    print(f"[parity] {case} {sched} {kw} M={len(pos)-1} N={B.shape[1]}", flush=True)
    C, C2 = _device_run(pos, crd, vals, B, sched, kw, runs=2)
    _check(pos, crd, vals, B, C, what=f"{case}/{sched}{kw}")
    assert np.array_equal(C, C2), "not deterministic"


def _spmv_case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "powerlaw_20000_n1":
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 20000, 20000)
        return pos, crd, vals, S.gen_dense(13, 20000, 1)
    if name == "sparse_rows_n1":  # tiles spanning more rows than the window
        lens = np.where(rng.random(30000) < 0.03, rng.integers(1, 4, 30000), 0)
        return _mk(rng, lens, 3000, 1)
    if name == "long_rows_n1":  # rows spanning many lanes and tiles
        lens = np.array([0, 7, 20000, 0, 0, 3, 1, 9000, 2, 0, 65, 64, 63, 1], np.int64)
        return _mk(rng, lens, 25000, 1)
    if name == "short_rows_n1":  # many rows inside one lane
        lens = rng.integers(0, 4, 20000)
        return _mk(rng, lens, 5000, 1)
    raise KeyError(name)


@pytest.mark.parametrize("tile", [64, 192, 256, 512, 1024])
@pytest.mark.parametrize("case", ["powerlaw_20000_n1", "sparse_rows_n1", "long_rows_n1",
                                  "short_rows_n1"])
def test_spmv_kernel_parity(case, tile):
    """eb_pr at N = 1 with tiles of 64 * E entries runs the SpMV kernel (E
    consecutive entries per lane, one segmented scan over the wave; full
    tiles read each lane's run with 64/128-bit loads)."""
    pos, crd, vals, B = _spmv_case(case)
    print(f"[spmv] {case} tile {tile} M={len(pos)-1} nnz={len(crd)}", flush=True)
    C, C2 = _device_run(pos, crd, vals, B, "eb_pr", dict(tile=tile), runs=2)
    _check(pos, crd, vals, B, C, what=f"spmv {case}/tile {tile}")
    assert np.array_equal(C, C2), "not deterministic"


@pytest.mark.parametrize("has_vals", [True, False])
@pytest.mark.parametrize("shift", [0, 1, 2])
@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp16"])
def test_spmv_kernel_alignment_and_storage(dtype, shift, has_vals):
    """crd and vals starting `shift` elements past an aligned address: shift 0
    takes the vector loads on full tiles, 1 and 2 the per-entry loads (16-bit
    vals at shift 2 are 4-byte but not 8-byte aligned); every storage type,
    with and without values, at the default tile and at 1024."""
    t = torch()
    pos, crd, vals, B = _spmv_case("powerlaw_20000_n1")
    K, M, nnz = B.shape[0], len(pos) - 1, len(crd)
    v = vals if has_vals else None
    if dtype == "f32":
        vr, Br = v, B
        vt = None if v is None else t.from_numpy(v)
        xt = t.from_numpy(B)
    else:
        vr = None if v is None else O.round_to(dtype, v)[1]
        Br = O.round_to(dtype, B)[1]
        vt = None if v is None else t.from_numpy(S.to_half(dtype, v).view(np.int16))
        xt = t.from_numpy(S.to_half(dtype, B).view(np.int16))
    dpos = dev(np.asarray(pos, np.int32))
    dcrd = t.zeros(nnz + shift, dtype=t.int32, device="cuda")
    dcrd[shift:] = t.from_numpy(np.asarray(crd, np.int32)).cuda()
    dval = None
    if vt is not None:
        dval = t.zeros(nnz + shift, dtype=vt.dtype, device="cuda")
        dval[shift:] = vt.cuda()
    desc = S.csr_desc(M, K, nnz, dpos.data_ptr(), dcrd[shift:].data_ptr(),
                      None if dval is None else dval[shift:].data_ptr(), S.DTYPES[dtype])
    x = xt.cuda()
    for tile in (0, 1024):
        plan = S.make_plan("eb_pr", **({"tile": tile} if tile else {}))
        print(f"[spmv] {dtype} shift {shift} vals {has_vals} tile {tile}", flush=True)
        bd = S.DTYPES[dtype]
        rp = S.plan_resolve(desc, 1, bd, plan, 1, 1)
        wsb = S.workspace_bytes(desc, 1, rp)
        ws = t.empty(max(wsb, 1), dtype=t.uint8, device="cuda")
        y = t.full((M, 1), float("nan"), dtype=t.float32, device="cuda")
        S.spmm(desc, x.data_ptr(), bd, 1, 1, y.data_ptr(), 1, plan, ws.data_ptr(), wsb,
               stream_ptr())
        t.cuda.synchronize()
        _check(pos, crd, vr, Br, y.cpu().numpy(), what=f"spmv {dtype}/shift {shift}/tile {tile}")


def test_spmv_kernel_strided_vectors():
    """x and y with leading dimensions 3 and 2 (every third / second element)."""
    pos, crd, vals, B = _spmv_case("short_rows_n1")
    K = B.shape[0]
    A = DeviceCSR(pos, crd, vals, K)
    Bs = np.zeros((K, 3), np.float32)
    Bs[:, 0] = B[:, 0]
    C = run_spmm(A, dense_dev(Bs), 1, S.make_plan("eb_pr"), ldb=3, ldc=2)
    _check(pos, crd, vals, B, C[: A.M, :1].cpu().numpy(), what="spmv strided")


@pytest.mark.parametrize("sched", ["rb_sr", "rb_pr", "eb_sr", "eb_pr", "eb_cs"])
def test_b_beyond_32bit_offsets(sched):
    """K > 2^24 rows of B: gathers take the 64-bit offset path; entries reach
    the last rows of B."""
    rng = np.random.default_rng(7)
    K, N, M = (1 << 24) + 1000, 4, 1500
    lens = rng.integers(0, 12, M)
    lens[-1] = 5
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    rows = [np.sort(rng.choice(K, int(l), replace=False)) for l in lens[:-1]]
    rows.append(np.arange(K - 5, K))  # the last row holds B's last five rows
    crd = np.concatenate(rows).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    print(f"[parity] K={K} N={N} nnz={len(crd)} {sched}", flush=True)
    (C,) = _device_run(pos, crd, vals, B, sched, {})
    _check(pos, crd, vals, B, C, what=f"K>2^24/{sched}")


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_against_eigen_output(path):
    d = np.load(path, allow_pickle=False)
    vals = d["vals"] if int(d["has_vals"]) else None
    scale = O.spmm_absbound(d["pos"], d["crd"], vals, d["B"])
    for sched in ("rb_sr", "rb_pr", "eb_sr", "eb_pr"):
        (C,) = _device_run(d["pos"], d["crd"], vals, d["B"], sched, {})
        err = O.scaled_err(C, d["C_eigen"].astype(np.float64), scale)
        assert err <= TOL_F32, f"{sched}: {err}"


def test_strided_operands():
    """ldb > N and ldc > N (row-major with padding)."""
    t = torch()
    pos, crd, vals, B = _case("powerlaw_20000_n128")
    N = 40
    Bp = np.zeros((B.shape[0], 48), np.float32)
    Bp[:, :N] = B[:, :N]
    A = DeviceCSR(pos, crd, vals, B.shape[0])
    for sched in ("rb_sr", "rb_pr", "eb_sr", "eb_pr"):
        C = t.full((A.M, 44), 7.0, dtype=t.float32, device="cuda")
        run_spmm(A, dev(Bp), N, S.make_plan(sched), ldb=48, ldc=44, C=C)
        Ch = C.cpu().numpy()
        _check(pos, crd, vals, B[:, :N].copy(), Ch[:, :N].copy(), what=sched)
        assert np.all(Ch[:, N:] == 7.0), "wrote outside [0, N)"


@pytest.mark.parametrize("T", [1, 7, 32, 256, 1024])
def test_partition_bitexact(T):
    rng = np.random.default_rng(T)
    lens = rng.integers(0, 9, 50000)
    lens[rng.random(50000) < 0.3] = 0
    lens[1000:1800] = 0
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    M, nnz = len(pos) - 1, int(pos[-1])
    nb = (nnz + T - 1) // T
    out = torch().full((nb + 1,), -1, dtype=torch().int32, device="cuda")
    S.block_starts(dev(pos).data_ptr(), M, T, nb, out.data_ptr(), stream_ptr())
    torch().cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), O.block_starts(pos, M, T, nb))


def test_full_size_powerlaw_all_schedules():
    """The power-law workload at full size (262144^2, N = 128)."""
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    B = S.gen_dense(3, 262144, 128)
    ref = O.spmm_f64(pos, crd, vals, B)
    scale = O.spmm_absbound(pos, crd, vals, B)
    A = DeviceCSR(pos, crd, vals, 262144)
    Bd = dense_dev(B)
    for sched in ("rb_sr", "rb_pr", "eb_sr", "eb_pr"):
        print(f"[full] {sched}", flush=True)
        C = run_spmm(A, Bd, 128, S.make_plan(sched)).cpu().numpy()
        err = O.scaled_err(C, ref, scale)
        assert err <= TOL_F32, f"{sched}: {err}"


@pytest.mark.parametrize("case", ["powerlaw_20000_n128", "empty_runs_n20", "n3_odd"])
def test_vendor_row_matches_oracle(case):
    """rocSPARSE (the comparison row) computes the same product."""
    t = torch()
    pos, crd, vals, B = _case(case)
    A = DeviceCSR(pos, crd, vals, B.shape[0])
    N = B.shape[1]
    Bd = dense_dev(B)
    C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
    ms = S.vendor_spmm(A.desc, Bd.data_ptr(), N, N, C.data_ptr(), N, 0, stream_ptr(), timed=True)
    t.cuda.synchronize()
    assert ms > 0
    _check(pos, crd, vals, B, C.cpu().numpy(), what=f"rocsparse {case}")


@pytest.mark.parametrize("alg", [0, 1, 2, 3])
def test_vendor_spmv_matches_oracle(alg):
    """rocSPARSE's SpMV (the N = 1 comparison row) computes the same product."""
    t = torch()
    pos, crd, vals, B = _spmv_case("powerlaw_20000_n1")
    A = DeviceCSR(pos, crd, vals, B.shape[0])
    xd = dense_dev(B)
    y = t.full((A.M, 1), float("nan"), dtype=t.float32, device="cuda")
    ms = S.vendor_spmv(A.desc, xd.data_ptr(), y.data_ptr(), alg, stream_ptr(), timed=True)
    t.cuda.synchronize()
    assert ms > 0
    _check(pos, crd, vals, B, y.cpu().numpy(), what=f"rocsparse spmv alg {alg}")


@pytest.mark.parametrize("sched", ["rb_sr", "eb_sr", "eb_pr"])
def test_stage_times_run_the_same_product(sched):
    """segtaco_spmm_stage_times computes the same C as segtaco_spmm and reports
    a positive main-kernel time (partition / fix-up only for element-balanced)."""
    t = torch()
    pos, crd, vals, B = _case("powerlaw_20000_n128")
    A = DeviceCSR(pos, crd, vals, B.shape[0])
    N = B.shape[1]
    Bd = dense_dev(B)
    plan = S.plan_resolve(A.desc, N, S.DT_F32, S.make_plan(sched))
    ref = run_spmm(A, Bd, N, plan).cpu().numpy()
    ws_bytes = S.workspace_bytes(A.desc, N, plan)
    ws = t.empty(max(ws_bytes, 1), dtype=t.uint8, device="cuda")
    C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
    st = S.spmm_stage_times(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, plan,
                            ws.data_ptr(), ws_bytes, stream_ptr(), reps=3)
    t.cuda.synchronize()
    assert np.array_equal(C.cpu().numpy(), ref)
    assert st[1] > 0
    if sched.startswith("eb"):
        assert st[0] > 0 and st[2] > 0


@pytest.mark.parametrize("name", [
    "spmm_rb_sr", "spmm_rb_pr", "spmm_eb_sr", "spmm_eb_pr", "spmm_csr_gpu_taco"])
def test_primary_entry_points(name):
    pos, crd, vals, B = _case("empty_runs_n20")
    C, ms = S.legacy(name, pos, crd, vals, B)
    _check(pos, crd, vals, B, C, what=name)
    assert ms > 0


@pytest.mark.parametrize("name", ["spmm_rb_sr", "spmm_csr_gpu_taco", "spmm_eb_pr",
                                  "spmm_rb_pr_256_2_16_8", "spmm_csr_gpu_cusparse"])
def test_cxx_linkage_entry_points(name):
    """The C++-linkage symbols — what an object compiled against the
    reference's gpu_kernels.cuh binds to (the link itself is checked on the
    CPU) — compute the product and set the timer."""
    import ctypes
    L = S.lib()
    fn = getattr(L, f"_Z{len(name)}{name}P13taco_tensor_tS0_S0_")
    fn.restype = None
    fn.argtypes = [ctypes.POINTER(S.TacoTensor)] * 3
    timer = L._Z20get_gpu_timer_resultv
    timer.restype = ctypes.c_float
    pos, crd, vals, B = _case("empty_runs_n20")
    M, K = len(pos) - 1, B.shape[0]
    C = np.full((M, B.shape[1]), np.nan, np.float32)
    keep = []
    tA = S._taco_csr(M, K, np.asarray(pos, np.int32), np.asarray(crd, np.int32),
                     np.asarray(vals, np.float32), keep)
    tB, tC = S._taco_dense(np.ascontiguousarray(B), keep), S._taco_dense(C, keep)
    prev = L.segtaco_set_error_mode(S.ERRMODE_RECORD)
    try:
        L.segtaco_clear_error()
        fn(ctypes.byref(tC), ctypes.byref(tA), ctypes.byref(tB))
        S.check(L.segtaco_last_error())
    finally:
        L.segtaco_set_error_mode(prev)
    _check(pos, crd, vals, B, C, what=f"C++ linkage {name}")
    assert timer() > 0


@pytest.mark.parametrize("N", [64, 37])
def test_operator_surface_large_operands(N):
    """Operands above the staging threshold (crd, vals, B and C of several
    8 MiB chunks, sizes not a multiple of a chunk) go through the pinned
    ring: the result matches the oracle on every call, and a second call that
    reuses the ring gives the same bits."""
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    B = S.gen_dense(12, 262144, N)
    C1, _ = S.legacy("spmm_eb_sr", pos, crd, vals, B)
    _check(pos, crd, vals, B, C1, what=f"staged N={N}")
    C2, _ = S.legacy("spmm_rb_pr", pos, crd, vals, B)
    _check(pos, crd, vals, B, C2, what=f"staged rb_pr N={N}")
    C3, _ = S.legacy("spmm_eb_sr", pos, crd, vals, B)
    assert np.array_equal(C1.view(np.uint32), C3.view(np.uint32))


def test_all_operator_surface_names():
    pos, crd, vals, B = _case("n3_odd")
    pos2, crd2, vals2, B2 = _case("uniform_4096_n32")
    for name in S.legacy_names():
        print(f"[legacy] {name}", flush=True)
        C, _ = S.legacy(name, pos, crd, vals, B)
        _check(pos, crd, vals, B, C, what=name)
        C2, _ = S.legacy(name, pos2, crd2, vals2, B2)
        _check(pos2, crd2, vals2, B2, C2, what=name)
