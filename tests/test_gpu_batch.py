"""GPU checks of the batched persistent launch and of graph capture/replay.

Every product of a heterogeneous batch (different shapes, N, row-length
distributions and schedules, auto-tuned or explicit) must match the oracle
(scaled error <= 1e-5 against FP64 of the stored inputs), be bit-identical to
the single-product call with the same plan, and stay bit-identical across
runs and across graph replays interleaved with other device work.
"""
import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu


def _mk(rng, lens, K, N):
    lens = np.minimum(np.asarray(lens), K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    return pos, crd, vals, B


def _products(seed=17):
    rng = np.random.default_rng(seed)
    out = []
    for i, N in enumerate([16, 32, 64, 128, 16, 32, 64, 128]):
        M = int(rng.integers(500, 6000))
        mean = float(rng.uniform(4, 64))
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 100 + i, M, M, mean=mean)
        out.append((pos, crd, vals, rng.uniform(-1, 1, (M, N)).astype(np.float32), None))
    pos, crd, vals = S.gen_csr(S.GEN_BANDED, 3, 4000, 4000, bw=32)
    out.append((pos, crd, vals, rng.uniform(-1, 1, (4000, 64)).astype(np.float32), None))
    pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, 4, 3000, 5000, mean=12)
    out.append((pos, crd, vals, rng.uniform(-1, 1, (5000, 32)).astype(np.float32), None))
    explicit = [
        ("rb_sr", {}), ("rb_sr", dict(group=4)), ("rb_pr", {}), ("rb_pr", dict(group=8)),
        ("eb_sr", {}), ("eb_sr", dict(tile=64)), ("eb_sr", dict(tile=512)),
        ("eb_pr", {}), ("eb_pr", dict(group=4, tile=128)),
    ]
    for j, (name, kw) in enumerate(explicit):
        N = [1, 3, 7, 16, 40, 130, 64, 128, 24][j]
        lens = rng.integers(0, 40, int(rng.integers(200, 2500)))
        lens[rng.integers(0, len(lens), 5)] = rng.integers(500, 3000, 5)
        pos, crd, vals, B = _mk(rng, lens, 3000, N)
        out.append((pos, crd, vals, B, S.make_plan(name, **kw)))
    # edge products: no stored entries, empty leading/trailing rows, one row
    out.append((np.zeros(51, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32),
                rng.uniform(-1, 1, (20, 16)).astype(np.float32), None))
    lens = np.zeros(300, np.int64)
    lens[100:120] = rng.integers(1, 50, 20)
    out.append(_mk(rng, lens, 700, 32) + (S.make_plan("eb_sr", tile=16),))
    out.append(_mk(rng, [5000], 6000, 16) + (S.make_plan("eb_pr"),))
    return out


def _build(products, dtype):
    dev = []
    items = []
    for pos, crd, vals, B, plan in products:
        A = DeviceCSR(pos, crd, vals, B.shape[0], dtype=dtype)
        Bd = dense_dev(B, dtype)
        N = B.shape[1]
        C = torch().full((max(A.M, 1), N), float("nan"), dtype=torch().float32, device="cuda")
        dev.append((A, Bd, C))
        items.append(S.batch_item(A.desc, Bd.data_ptr(), N, C.data_ptr(), N, N, plan,
                                  np.ascontiguousarray(pos, np.int32)))
    return dev, items


def _refs(products, dtype):
    refs = []
    for pos, crd, vals, B, _ in products:
        if dtype != "f32":
            vals = O.round_to(dtype, vals)[1]
            B = O.round_to(dtype, B)[1]
        refs.append((O.spmm_f64(pos, crd, vals, B), O.spmm_absbound(pos, crd, vals, B)))
    return refs


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_batch_matches_oracle_and_single_calls(dtype):
    products = _products()
    dev, items = _build(products, dtype)
    batch = S.Batch(items, dtype, stream_ptr())
    print(f"[batch] {len(items)} products, {batch.units()} units, "
          f"{batch.workspace_bytes()} B workspace", flush=True)
    batch.run(stream_ptr())
    torch().cuda.synchronize()
    outs = [C[: A.M].cpu().numpy() for A, _, C in dev]
    refs = _refs(products, dtype)
    for i, ((pos, crd, vals, B, _), (ref, scale), C) in enumerate(zip(products, refs, outs)):
        err = O.scaled_err(C, ref, scale)
        plan = batch.plan(i)
        assert err <= 1e-5, f"product {i} plan {plan}: {err:.3e}"
        A, Bd, _ = dev[i]
        if A.M and B.shape[1]:
            single = run_spmm(A, Bd, B.shape[1], plan, b_dtype=dtype)[: A.M].cpu().numpy()
            assert np.array_equal(single, C), f"product {i}: batch != single call"
    # second run: bit-identical
    for _, _, C in dev:
        C.fill_(float("nan"))
    batch.run(stream_ptr())
    torch().cuda.synchronize()
    for (A, _, C), prev in zip(dev, outs):
        assert np.array_equal(C[: A.M].cpu().numpy(), prev)
    batch.destroy()


def test_batch_graph_replay():
    products = _products(seed=23)
    dev, items = _build(products, "f32")
    batch = S.Batch(items, "f32", stream_ptr())
    batch.run(stream_ptr())
    torch().cuda.synchronize()
    eager = [C[: A.M].cpu().numpy() for A, _, C in dev]
    batch.capture()
    scratch = torch().randn(1 << 20, device="cuda")
    for rep in range(4):
        for _, _, C in dev:
            C.fill_(float("nan"))
        scratch.mul_(1.0001).add_(0.5)  # unrelated eager work between replays
        torch().cuda.synchronize()
        batch.replay(stream_ptr())
        torch().cuda.synchronize()
        for i, ((A, _, C), ref) in enumerate(zip(dev, eager)):
            assert np.array_equal(C[: A.M].cpu().numpy(), ref), f"replay {rep} product {i}"
    batch.destroy()


def test_autotune_measure_picks_a_correct_plan():
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 41, 30000, 30000)
    B = S.gen_dense(42, 30000, 64)
    A = DeviceCSR(pos, crd, vals, 30000)
    Bd = dense_dev(B)
    C = torch().full((30000, 64), float("nan"), dtype=torch().float32, device="cuda")
    st = S.row_stats(pos, 30000)
    S.tune_cache_clear()
    plan, ms = S.tune_measure(A.desc, st, Bd.data_ptr(), S.DT_F32, 64, 64, C.data_ptr(), 64,
                              reps=2, stream=stream_ptr())
    torch().cuda.synchronize()
    print(f"[tune] chose {plan} at {ms:.4f} ms", flush=True)
    assert ms > 0
    err = O.scaled_err(C.cpu().numpy(), O.spmm_f64(pos, crd, vals, B),
                       O.spmm_absbound(pos, crd, vals, B))
    assert err <= 1e-5
    cached = S.tune(st, 64)  # cached under the signature
    if plan.schedule == S.EB_CS:  # standalone only: the cache keeps a shared schedule
        assert S.RB_SR <= cached.schedule <= S.EB_PR
    else:
        assert cached.as_dict() == plan.as_dict()
    S.tune_cache_clear()


def test_autotune_measure_offers_eb_cs_to_standalone_calls():
    """On the power-law 262144^2 product (4.2 M entries) at N = 128, FP32,
    eb_cs's step is about a fifth faster than every other candidate's; the
    measured choice returns it, leaves C correct, and caches a plan that
    batches can run."""
    M = 262144
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 43, M, M)
    B = S.gen_dense(44, M, 128)
    A = DeviceCSR(pos, crd, vals, M)
    Bd = dense_dev(B)
    C = torch().full((M, 128), float("nan"), dtype=torch().float32, device="cuda")
    st = S.row_stats(pos, M)
    S.tune_cache_clear()
    plan, ms = S.tune_measure(A.desc, st, Bd.data_ptr(), S.DT_F32, 128, 128, C.data_ptr(), 128,
                              reps=3, stream=stream_ptr())
    torch().cuda.synchronize()
    print(f"[tune] nnz {int(pos[-1])} chose {plan} at {ms:.4f} ms", flush=True)
    assert plan.schedule == S.EB_CS
    err = O.scaled_err(C.cpu().numpy(), O.spmm_f64(pos, crd, vals, B),
                       O.spmm_absbound(pos, crd, vals, B))
    assert err <= 1e-5
    cached = S.tune(st, 128)
    assert S.RB_SR <= cached.schedule <= S.EB_PR
    S.tune_cache_clear()
