"""GPU parity of 16-bit storage (BF16 / FP16 vals and B, FP32 accumulation).

Tolerances:
  * vs FP64 of the *rounded* inputs (what the kernel computes, up to FP32
    accumulation), on the oracle's per-element scaled metric
    (max |C - ref| / sum |v||b|): <= 1e-5;
  * vs FP64 of the original FP32 inputs (the rounding of both operands
    included), norm-wise: ||C - ref||_F / ||ref||_F <= 1e-2 and
    max |C - ref| / max |ref| <= 1e-2 (per element, a product whose inputs
    fall below FP16's subnormal range may round to zero).
Compensated (Neumaier) summation is pinned by a row whose uncompensated FP32
sum provably loses 2^-25 * 4000 relative to 1.
"""
import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm

pytestmark = pytest.mark.gpu

SCHEDS = ["rb_sr", "rb_pr", "eb_sr", "eb_pr"]


def _rounded(dtype, x):
    return O.round_to(dtype, x)[1]


def _run(pos, crd, vals, B, dtype, sched, **kw):
    A = DeviceCSR(pos, crd, vals, B.shape[0], dtype=dtype)
    C = run_spmm(A, dense_dev(B, dtype), B.shape[1], S.make_plan(sched, **kw), b_dtype=dtype)
    return C[: A.M].cpu().numpy()


def _mk(rng, lens, K, N):
    lens = np.minimum(lens, K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    return pos, crd, vals, B


def _cases():
    rng = np.random.default_rng(31)
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 5, 20000, 20000)
    yield "powerlaw_20000_n256", (pos, crd, vals, S.gen_dense(6, 20000, 256))
    pos, crd, vals = S.gen_csr(S.GEN_BANDED, 7, 30000, 30000, bw=32)
    yield "banded_30000_n256", (pos, crd, vals, S.gen_dense(8, 30000, 256))
    lens = rng.integers(0, 30, 600)
    lens[[5, 300]] = [3000, 2500]  # rows above the 1024 compensation threshold
    yield "longrows_n40", _mk(rng, lens, 4000, 40)
    yield "n130", _mk(rng, rng.integers(0, 20, 700), 900, 130)
    yield "n7", _mk(rng, rng.integers(0, 20, 700), 900, 7)
    lens = rng.integers(0, 30, 3000)
    lens[[7, 2000]] = [4000, 1500]
    yield "spmv_n1", _mk(rng, lens, 5000, 1)


CASES = list(_cases())


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
@pytest.mark.parametrize("sched", SCHEDS)
@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_half_storage_parity(case, sched, dtype):
    pos, crd, vals, B = dict(CASES)[case]
    print(f"[precision] {case} {sched} {dtype}", flush=True)
    C = _run(pos, crd, vals, B, dtype, sched)
    vr, Br = _rounded(dtype, vals), _rounded(dtype, B)
    err_r = O.scaled_err(C, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
    assert err_r <= 1e-5, f"vs FP64 of rounded inputs: {err_r:.3e}"
    fro, mx = O.rel_norm_errs(C, O.spmm_f64(pos, crd, vals, B))
    assert fro <= 1e-2 and mx <= 1e-2, f"vs FP64 of FP32 inputs: fro {fro:.3e} max {mx:.3e}"


def _absorbing_row(dtype):
    # One row: 1.0 then 4000 terms of 2^-25 (all exact in BF16 and FP16 via B).
    # Uncompensated FP32 gives 1.0 (each 2^-25 is below half an ulp of 1);
    # the exact sum is 1 + 4000 * 2^-25.
    L = 4001
    pos = np.array([0, L], np.int32)
    crd = np.arange(L, dtype=np.int32)
    vals = np.ones(L, np.float32)
    B = np.full((L, 4), 2.0 ** -25, np.float32)
    B[0, :] = 1.0
    if dtype == "fp16":  # 2^-25 is below FP16's smallest subnormal: scale up
        vals[:] = 2.0 ** -10
        vals[0] = 1.0
        B[1:, :] = 2.0 ** -15
    return pos, crd, vals, B


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
@pytest.mark.parametrize("sched", SCHEDS)
def test_compensated_long_row(sched, dtype):
    pos, crd, vals, B = _absorbing_row(dtype)
    vr, Br = _rounded(dtype, vals), _rounded(dtype, B)
    ref = O.spmm_f64(pos, crd, vr, Br)
    scale = O.spmm_absbound(pos, crd, vr, Br)
    assert ref[0, 0] - 1.0 == pytest.approx(4000 * 2.0 ** -25)
    C = _run(pos, crd, vals, B, dtype, sched, comp_threshold=1024)
    assert O.scaled_err(C, ref, scale) <= 1e-6


def test_uncompensated_sequential_loses_the_tail():
    # The same row through rb_sr without compensation shows the loss the
    # compensated test guards against (so that test has teeth).
    pos, crd, vals, B = _absorbing_row("bf16")
    vr, Br = _rounded("bf16", vals), _rounded("bf16", B)
    C = _run(pos, crd, vals, B, "bf16", "rb_sr", comp_threshold=0)
    assert C[0, 0] == 1.0
    C2 = _run(pos, crd, vals, B, "bf16", "rb_sr", comp_threshold=1024)
    assert C2[0, 0] > 1.0


def test_half_storage_full_powerlaw_n256():
    """The precision workload's power-law matrix at full size (262144^2, N=256)."""
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    B = S.gen_dense(4, 262144, 256)
    ref = O.spmm_f64(pos, crd, vals, B)
    for dtype in ("bf16", "fp16"):
        C = _run(pos, crd, vals, B, dtype, "eb_sr")
        fro, mx = O.rel_norm_errs(C, ref)
        print(f"[precision full] {dtype}: fro {fro:.3e} max {mx:.3e}", flush=True)
        assert fro <= 1e-2 and mx <= 1e-2
