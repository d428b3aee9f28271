"""The three large workloads checked at their stated size (GPU).

* Banded 1048576 x 1048576 (|i - j| <= 32, 65 entries per row), values and B
  stored BF16 and FP16, N = 256, FP32 accumulation: the block-dense schedule
  (matrix cores) and rb_sr against the FP64 oracle of the rounded inputs on
  a sample of 163 840 rows — every block's first two and last two rows (the
  block edges, where windows and the tile scatter end) and 32 768 random
  others — scaled error <= 1e-5; and norm-wise <= 1e-2 against FP64 of the
  FP32 inputs.
* The batching workload: 512 power-law products (M ~ U[1000, 200000], mean
  row length ~ U[8, 64], sigma 1.5, cap 8192, N in {16, 32, 64, 128}, FP32,
  about 1.9 G stored entries) in one persistent launch with autotuned plans:
  every product bit-equal to its single call with the same plan (compared on
  the device), and 32 sampled products against the FP64 oracle (<= 1e-5).
* Serving over a pool of the batching distribution larger than the CSR
  cache (BF16 storage): a Poisson trace (lambda = 2000 / s, Zipf(1)
  popularity) with cache misses and evictions; sampled requests bit-equal
  to the same product run alone, and against the oracle.
"""
import time

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _sub_rows(pos, crd, vals, rows):
    """The CSR of the chosen rows (columns unchanged)."""
    lens = (pos[rows + 1] - pos[rows]).astype(np.int64)
    spos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(pos[r], pos[r + 1]) for r in rows]) if len(rows) else \
        np.zeros(0, np.int64)
    return spos, crd[idx], (None if vals is None else vals[idx])


@pytest.fixture(scope="module")
def banded():
    t0 = time.time()
    M = 1048576
    pos, crd, vals = S.gen_csr(S.GEN_BANDED, 9, M, M, bw=32)
    assert int(pos[-1]) == 65 * M - 1056
    rng = np.random.default_rng(3)
    blk = np.arange(0, M, 32)
    edges = np.concatenate([blk, blk + 1, blk + 30, blk + 31])
    rows = np.unique(np.concatenate([edges, rng.choice(M, 32768, replace=False)]))
    B = S.gen_dense(10, M, 256)
    print(f"[configs] banded 1M^2 built, {len(rows)} sampled rows ({time.time()-t0:.1f}s)",
          flush=True)
    return pos, crd, vals, B, rows


@pytest.mark.parametrize("sched", ["bd", "rb_sr"])
@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_banded_1m_n256_full_size(banded, dtype, sched):
    pos, crd, vals, B, rows = banded
    M, N = len(pos) - 1, 256
    t0 = time.time()
    A = DeviceCSR(pos, crd, vals, M, dtype=dtype)
    Bd = dense_dev(B, dtype)
    C = run_spmm(A, Bd, N, S.make_plan(sched), b_dtype=dtype)
    Cs = C[torch().from_numpy(rows).cuda()].cpu().numpy()
    del C, Bd, A
    torch().cuda.empty_cache()
    print(f"[configs] banded {dtype} {sched}: device run + copy {time.time()-t0:.1f}s",
          flush=True)
    spos, scrd, svals = _sub_rows(pos, crd, vals, rows)
    vr = O.round_to(dtype, svals)[1]
    Br = O.round_to(dtype, B)[1]
    err = O.scaled_err(Cs, O.spmm_f64(spos, scrd, vr, Br), O.spmm_absbound(spos, scrd, vr, Br))
    assert err <= 1e-5, f"{dtype}/{sched}: scaled error vs FP64 of rounded inputs {err:.3e}"
    fro, _ = O.rel_norm_errs(Cs, O.spmm_f64(spos, scrd, svals, B))
    assert fro <= 1e-2, f"{dtype}/{sched}: vs FP64 of FP32 inputs {fro:.3e}"
    print(f"[configs] banded {dtype} {sched}: err {err:.2e}, fro {fro:.2e} "
          f"({time.time()-t0:.1f}s)", flush=True)


def _batch_set(n, seed=41):
    """The batching workload's matrices (the bench's generator and seeds)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        M = int(rng.integers(1000, 200001))
        mean = float(rng.uniform(8, 64))
        N = int(rng.choice([16, 32, 64, 128]))
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 5000 + i, M, M, mean=mean)
        out.append((pos, crd, vals, M, N))
    return out


def test_batch_512_products_full_size():
    t = torch()
    t0 = time.time()
    mats = _batch_set(512)
    total = sum(int(m[0][-1]) for m in mats)
    print(f"[configs] batch set: 512 products, {total} entries ({time.time()-t0:.1f}s)",
          flush=True)
    assert total > 1_000_000_000
    keep, items = [], []
    for i, (pos, crd, vals, M, N) in enumerate(mats):
        A = DeviceCSR(pos, crd, vals, M)
        g = t.Generator(device="cuda")
        g.manual_seed(900 + i)
        Bd = t.rand((M, N), device="cuda", generator=g) * 2 - 1
        C = t.full((M, N), float("nan"), dtype=t.float32, device="cuda")
        keep.append((A, Bd, C))
        items.append(S.batch_item(A.desc, Bd.data_ptr(), N, C.data_ptr(), N, N, None, pos))
    b = S.Batch(items, "f32", stream_ptr())
    b.capture()
    b.replay(stream_ptr())
    t.cuda.synchronize()
    print(f"[configs] batch ran: {b.units()} units ({time.time()-t0:.1f}s)", flush=True)
    rng = np.random.default_rng(5)
    sample = set(rng.choice(512, 32, replace=False).tolist())
    for i, ((pos, crd, vals, M, N), (A, Bd, C)) in enumerate(zip(mats, keep)):
        alone = run_spmm(A, Bd, N, b.plan(i))[:M]
        assert t.equal(alone.view(t.int32), C.view(t.int32)), f"product {i}: batch != alone"
        if i in sample:
            Bh = Bd.cpu().numpy()
            err = O.scaled_err(C.cpu().numpy(), O.spmm_f64(pos, crd, vals, Bh),
                               O.spmm_absbound(pos, crd, vals, Bh))
            assert err <= 1e-5, f"product {i} (M {M}, N {N}): {err:.3e}"
        del alone
    print(f"[configs] batch checked ({time.time()-t0:.1f}s)", flush=True)
    b.destroy()


def test_serving_pool_larger_than_cache():
    t = torch()
    t0 = time.time()
    mats = _batch_set(96, seed=43)
    csr_bytes = sum(4 * (m[3] + 1) + 6 * int(m[0][-1]) for m in mats)
    cache = csr_bytes // 4  # the pool is four times the cache
    server = S.Server(cache_bytes=cache, inflight_bytes=4096 * MiB, max_batch=32, streams=2,
                      dtype="bf16")
    for pos, crd, vals, M, N in mats:
        server.add_matrix(pos, crd, vals, M)
    rng = np.random.default_rng(11)
    n = 400
    arrivals = np.cumsum(rng.exponential(1e6 / 2000.0, n))
    pop = 1.0 / np.arange(1, len(mats) + 1)
    mids = rng.choice(len(mats), n, p=pop / pop.sum())
    Ns = np.array([mats[m][4] for m in mids], np.int32)
    Bs = {}
    Cs = []
    for m, N in zip(mids, Ns):
        M = mats[m][3]
        if m not in Bs:
            Bs[m] = dense_dev(S.gen_dense(300 + int(m), M, int(N)), "bf16")
        Cs.append(t.full((M, int(N)), float("nan"), dtype=t.float32, device="cuda"))
    t.cuda.synchronize()
    print(f"[configs] serving pool: {len(mats)} matrices, {csr_bytes >> 20} MiB of CSR, cache "
          f"{cache >> 20} MiB ({time.time()-t0:.1f}s)", flush=True)
    server.replay(arrivals, mids, [Bs[m].data_ptr() for m in mids], Ns, Ns,
                  [C.data_ptr() for C in Cs], Ns, timeout=300.0)
    st = server.stats()
    print(f"[configs] serving: {st}", flush=True)
    assert st["completed"] == n and st["rejected"] == 0
    assert st["cache_misses"] > len(set(mids.tolist())) and st["evictions"] > 0
    for k in rng.choice(n, 24, replace=False):
        m, N = int(mids[k]), int(Ns[k])
        pos, crd, vals, M, _ = mats[m]
        A = DeviceCSR(pos, crd, vals, M, dtype="bf16")
        alone = run_spmm(A, Bs[m], N, server.plan(m, N), b_dtype="bf16")[:M]
        assert t.equal(alone.view(t.int32), Cs[k].view(t.int32)), f"request {k}: mixed != alone"
    for k in rng.choice(n, 4, replace=False):
        m, N = int(mids[k]), int(Ns[k])
        pos, crd, vals, M, _ = mats[m]
        vr = O.round_to("bf16", vals)[1]
        Br = O.round_to("bf16", S.gen_dense(300 + m, M, N))[1]
        err = O.scaled_err(Cs[k].cpu().numpy(), O.spmm_f64(pos, crd, vr, Br),
                           O.spmm_absbound(pos, crd, vr, Br))
        assert err <= 1e-5, f"request {k}: {err:.3e}"
    server.destroy()
    print(f"[configs] serving checked ({time.time()-t0:.1f}s)", flush=True)
