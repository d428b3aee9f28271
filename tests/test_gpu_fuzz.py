"""Seeded random sweep through the C-ABI: random CSR structures (empty rows,
short rows, log-normal rows, sparse long rows, K down to 1), every schedule,
FP32 / BF16 / FP16 storage, column counts from 1 to 300, padded leading
dimensions, with and without values, and tiles / groups drawn from the
values plan resolution accepts (a combination it refuses is skipped).

Tolerances as everywhere: FP32 storage within 1e-5 of the FP64 oracle on the
oracle's scaled metric; 16-bit storage within 1e-5 of FP64 on the rounded
inputs. Each case prints its seed and plan first, so a failure names it.
"""
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm

pytestmark = pytest.mark.gpu

SEEDS = list(range(96))


def _structure(rng):
    M = int(rng.integers(1, 1500))
    K = int(rng.integers(1, 1500))
    kind = int(rng.integers(0, 4))
    if kind == 0:
        lens = rng.integers(0, 8, M)
    elif kind == 1:
        lens = rng.poisson(20, M)
    elif kind == 2:
        lens = rng.lognormal(2.0, 1.5, M).astype(np.int64)
    else:
        lens = np.where(rng.random(M) < 0.9, 0, rng.integers(1, 1400, M))
    lens = np.minimum(lens, K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    return pos, crd, vals, K


def _plan(rng, dtype):
    scheds = ["rb_sr", "rb_pr", "eb_sr", "eb_pr", "eb_cs"] + (["bd"] if dtype != "f32" else [])
    name = str(rng.choice(scheds))
    kw = {}
    if name == "rb_sr":
        kw["group"] = int(rng.choice([0, 1, 2, 4, 16]))
    elif name == "rb_pr":
        kw["group"] = int(rng.choice([0, 2, 8, 32]))
    elif name == "eb_sr":
        kw["tile"] = int(rng.choice([0, 8, 16, 40, 128, 200, 512]))
    elif name == "eb_pr":
        kw["tile"] = int(rng.choice([0, 64, 128, 256, 1024, 96]))
    elif name == "eb_cs":
        kw["tile"] = int(rng.choice([0, 32, 64, 256]))
    return name, {k: v for k, v in kw.items() if v}


@pytest.mark.parametrize("seed", SEEDS)
def test_random_structures_plans_and_shapes(seed):
    rng = np.random.default_rng(zlib.crc32(f"fuzz{seed}".encode()))
    pos, crd, vals, K = _structure(rng)
    dtype = str(rng.choice(["f32", "bf16", "fp16"]))
    N = int(rng.choice([1, 2, 3, 7, 16, 31, 64, 100, 128, 256, 300]))
    ldb = N + int(rng.choice([0, 0, 1, 3, 8]))
    ldc = N + int(rng.choice([0, 0, 2, 5, 8]))
    has_vals = bool(rng.random() < 0.85)
    name, kw = _plan(rng, dtype)
    M = len(pos) - 1
    print(f"[fuzz] seed {seed}: M={M} K={K} nnz={len(crd)} {dtype} N={N} ldb={ldb} ldc={ldc} "
          f"vals={has_vals} {name} {kw}", flush=True)
    v = vals if has_vals else None
    A = DeviceCSR(pos, crd, v, K, dtype=dtype)
    plan = S.make_plan(name, **kw)
    try:
        S.plan_resolve(A.desc, N, S.DTYPES[dtype], plan, ldb, ldc)
    except S.SegtacoError as e:
        pytest.skip(f"plan refused: {e}")
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Bp = np.zeros((K, ldb), np.float32)
    Bp[:, :N] = B
    C = run_spmm(A, dense_dev(Bp, dtype), N, plan, ldb=ldb, ldc=ldc, b_dtype=dtype)
    Cg = C[:M, :N].cpu().numpy()
    vr = v if (dtype == "f32" or v is None) else O.round_to(dtype, v)[1]
    Br = B if dtype == "f32" else O.round_to(dtype, B)[1]
    err = O.scaled_err(Cg, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
    assert err <= 1e-5, f"seed {seed}: scaled error {err:.3e}"


@pytest.mark.parametrize("seed", list(range(24)))
def test_random_batches(seed):
    """Random batches of 4-16 products (random structures, widths, padded
    leading dimensions; the autotuner's plan or a random rb/eb plan per
    product) in one persistent launch: every product within tolerance of the
    FP64 oracle and bit-identical to the same plan run alone."""
    from gpu_util import stream_ptr, torch
    t = torch()
    rng = np.random.default_rng(zlib.crc32(f"batchfuzz{seed}".encode()))
    dtype = str(rng.choice(["f32", "bf16", "fp16"]))
    n = int(rng.integers(4, 17))
    prods, items, keep = [], [], []
    for i in range(n):
        pos, crd, vals, K = _structure(rng)
        N = int(rng.choice([1, 3, 16, 32, 64, 100, 128]))
        ldb = N + int(rng.choice([0, 0, 4]))
        ldc = N + int(rng.choice([0, 0, 3]))
        if rng.random() < 0.5:
            plan = None
        else:
            name, kw = _plan(rng, "f32")
            if name == "eb_cs":
                name, kw = "eb_sr", {}
            plan = S.make_plan(name, **kw)
        A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        Bp = np.zeros((K, ldb), np.float32)
        Bp[:, :N] = B
        Bd = dense_dev(Bp, dtype)
        C = t.full((max(A.M, 1), ldc), float("nan"), dtype=t.float32, device="cuda")
        prods.append((pos, crd, vals, B, N, ldb, ldc))
        keep.append((A, Bd, C))
        items.append(S.batch_item(A.desc, Bd.data_ptr(), ldb, C.data_ptr(), ldc, N, plan,
                                  np.ascontiguousarray(pos, np.int32)))
    print(f"[batch fuzz] seed {seed}: {n} products {dtype}", flush=True)
    try:
        batch = S.Batch(items, dtype, stream_ptr())
    except S.SegtacoError as e:
        pytest.skip(f"batch refused: {e}")
    batch.run(stream_ptr())
    t.cuda.synchronize()
    for i, ((pos, crd, vals, B, N, ldb, ldc), (A, Bd, C)) in enumerate(zip(prods, keep)):
        M = len(pos) - 1
        Cg = C[:M, :N].cpu().numpy()
        vr = vals if dtype == "f32" else O.round_to(dtype, vals)[1]
        Br = B if dtype == "f32" else O.round_to(dtype, B)[1]
        err = O.scaled_err(Cg, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
        plan = batch.plan(i)
        assert err <= 1e-5, f"seed {seed} product {i} plan {plan.as_dict()}: {err:.3e}"
        if M:
            alone = run_spmm(A, Bd, N, plan, ldb=ldb, ldc=ldc, b_dtype=dtype)
            assert np.array_equal(alone[:M, :N].cpu().numpy(), Cg), \
                f"seed {seed} product {i}: batch != single call"


@pytest.mark.parametrize("seed", list(range(24)))
def test_random_sddmm(seed):
    """SDDMM over random structures, storage types, widths J and padded
    leading dimensions of X and Y, with and without values: within 1e-5 of
    the FP64 oracle (16-bit storage: on the rounded inputs)."""
    from test_sddmm import TOL, _err, _run
    rng = np.random.default_rng(zlib.crc32(f"sddmmfuzz{seed}".encode()))
    pos, crd, vals, K = _structure(rng)
    M = len(pos) - 1
    dtype = str(rng.choice(["f32", "bf16", "fp16"]))
    J = int(rng.choice([1, 2, 5, 8, 16, 37, 64, 128, 200, 300]))
    ldx = J + int(rng.choice([0, 0, 1, 8]))
    ldy = J + int(rng.choice([0, 0, 3, 8]))
    v = vals if rng.random() < 0.85 else None
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    if dtype != "f32":
        X, Y = O.round_to(dtype, X)[1], O.round_to(dtype, Y)[1]
        v = None if v is None else O.round_to(dtype, v)[1]
    print(f"[sddmm fuzz] seed {seed}: M={M} K={K} nnz={len(crd)} {dtype} J={J} "
          f"ldx={ldx} ldy={ldy} vals={v is not None}", flush=True)
    out = _run(pos, crd, v, X, Y, dtype, ldx=ldx, ldy=ldy)[0]
    ref = O.sddmm_f64(pos, crd, v, X, Y)
    scale = O.sddmm_f64(pos, crd, v, X, Y, absval=True)
    assert _err(out, ref, scale) <= TOL, f"seed {seed}"


def test_random_operator_surface():
    """Every one of the 79 operator-surface names (host operands in, C out,
    the reference's calling convention) on its own random structure and a
    random width from 1 to 300: within 1e-5 of the FP64 oracle."""
    for name in S.legacy_names():
        rng = np.random.default_rng(zlib.crc32(f"legacy{name}".encode()))
        pos, crd, vals, K = _structure(rng)
        N = int(rng.choice([1, 2, 4, 17, 32, 128, 130, 256, 300]))
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        print(f"[legacy fuzz] {name}: M={len(pos) - 1} K={K} nnz={len(crd)} N={N}", flush=True)
        C, _ = S.legacy(name, pos, crd, vals, B)
        err = O.scaled_err(C, O.spmm_f64(pos, crd, vals, B), O.spmm_absbound(pos, crd, vals, B))
        assert err <= 1e-5, f"{name}: scaled error {err:.3e}"


@pytest.mark.parametrize("seed", list(range(32)))
def test_random_banded_block_dense(seed):
    """bd (32-row blocks on the matrix cores) on random banded structures —
    band half-width 0-100 around a shifted diagonal, random density, M not a
    multiple of 32, K from 1 — so that most blocks' column windows fit the
    256-column limit and take the MFMA path; random widths and padded
    leading dimensions, BF16 and FP16, with and without values."""
    rng = np.random.default_rng(zlib.crc32(f"bdfuzz{seed}".encode()))
    M = int(rng.integers(1, 3000))
    K = int(rng.integers(1, 3000))
    hw = int(rng.integers(0, 101))
    shift = int(rng.integers(-50, 51))
    dens = float(rng.uniform(0.05, 1.0))
    rows = []
    for i in range(M):
        c = int(round(i * K / max(M, 1))) + shift
        lo, hi = max(0, c - hw), min(K, c + hw + 1)
        if hi <= lo:
            rows.append(np.zeros(0, np.int64))
            continue
        cols = np.arange(lo, hi)
        rows.append(cols[rng.random(len(cols)) < dens])
    pos = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    crd = np.concatenate(rows + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    dtype = str(rng.choice(["bf16", "fp16"]))
    N = int(rng.choice([1, 7, 32, 64, 100, 128, 200, 256, 300]))
    ldb = N + int(rng.choice([0, 0, 5, 8]))
    ldc = N + int(rng.choice([0, 0, 3]))
    v = vals if rng.random() < 0.85 else None
    print(f"[bd fuzz] seed {seed}: M={M} K={K} nnz={len(crd)} hw={hw} {dtype} N={N} "
          f"ldb={ldb} ldc={ldc} vals={v is not None}", flush=True)
    A = DeviceCSR(pos, crd, v, K, dtype=dtype)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Bp = np.zeros((K, ldb), np.float32)
    Bp[:, :N] = B
    C = run_spmm(A, dense_dev(Bp, dtype), N, S.make_plan("bd"), ldb=ldb, ldc=ldc, b_dtype=dtype)
    Cg = C[:M, :N].cpu().numpy()
    vr = None if v is None else O.round_to(dtype, v)[1]
    Br = O.round_to(dtype, B)[1]
    err = O.scaled_err(Cg, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
    assert err <= 1e-5, f"seed {seed}: scaled error {err:.3e}"
