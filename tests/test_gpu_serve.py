"""GPU checks of the continuous-batching server.

* Every request's result matches the oracle and is bit-identical to the same
  product run alone (single call with the plan the server uses).
* A cache smaller than the working set evicts and still computes correctly.
* Ticks overlap on the device (two streams busy at once under a burst).
* A request that cannot fit the configured capacities is rejected, not hung.
* A Poisson trace replays to completion with latency statistics.
"""
import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _pool(seed, n, dtype="f32"):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(n):
        M = int(rng.integers(300, 8000))
        kind = [S.GEN_POWERLAW, S.GEN_UNIFORM, S.GEN_BANDED][i % 3]
        pos, crd, vals = S.gen_csr(kind, 1000 + i, M, M, mean=float(rng.uniform(4, 48)),
                                   bw=int(rng.integers(2, 20)))
        mats.append((pos, crd, vals, M))
    return mats


def _requests(rng, mats, n, Ns=(16, 32, 64)):
    reqs = []
    for _ in range(n):
        mi = int(rng.integers(0, len(mats)))
        N = int(rng.choice(Ns))
        K = mats[mi][3]
        reqs.append((mi, N, rng.uniform(-1, 1, (K, N)).astype(np.float32)))
    return reqs


def _check_results(server, mats, reqs, dev, dtype):
    for (mi, N, B), (Bd, C) in zip(reqs, dev):
        pos, crd, vals, K = mats[mi]
        Cg = C.cpu().numpy()
        vr = vals if dtype == "f32" else O.round_to(dtype, vals)[1]
        Br = B if dtype == "f32" else O.round_to(dtype, B)[1]
        err = O.scaled_err(Cg, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
        assert err <= 1e-5, f"matrix {mi} N {N}: {err:.3e}"
        A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
        alone = run_spmm(A, Bd, N, server.plan(mi, N), b_dtype=dtype)[: A.M].cpu().numpy()
        assert np.array_equal(alone, Cg), f"matrix {mi} N {N}: mixed != alone"


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_mixed_requests_equal_alone(dtype):
    rng = np.random.default_rng(5)
    mats = _pool(7, 12, dtype)
    server = S.Server(cache_bytes=256 * MiB, inflight_bytes=256 * MiB, max_batch=8, streams=2,
                      dtype=dtype)
    ids = [server.add_matrix(p, c, v, K) for p, c, v, K in mats]
    assert ids == list(range(len(mats)))
    reqs = _requests(rng, mats, 48)
    dev = []
    for mi, N, B in reqs:
        M = mats[mi][3]
        Bd = dense_dev(B, dtype)
        C = torch().full((M, N), float("nan"), dtype=torch().float32, device="cuda")
        dev.append((Bd, C))
    torch().cuda.synchronize()
    rids = [server.submit(mi, Bd.data_ptr(), N, N, C.data_ptr(), N)
            for (mi, N, _), (Bd, C) in zip(reqs, dev)]
    for rid in rids:
        assert server.wait(rid, 120.0) == S.OK
    st = server.stats()
    print(f"[serve] {st}", flush=True)
    assert st["completed"] == len(reqs) and st["ticks"] >= 1
    _check_results(server, mats, reqs, dev, dtype)
    server.destroy()


def test_eviction_under_small_cache():
    rng = np.random.default_rng(9)
    mats = _pool(11, 10)
    # CSR bytes (+ tile rows) of the largest matrix; the cache holds ~2 of 10
    biggest = max(4 * (K + 1) + 8 * int(p[-1]) + 8 * (int(p[-1]) // 64 + 2) + 4096
                  for p, _, _, K in mats)
    cache = int(2.2 * biggest)
    server = S.Server(cache_bytes=cache, inflight_bytes=128 * MiB, max_batch=4, streams=2)
    for p, c, v, K in mats:
        server.add_matrix(p, c, v, K)
    reqs = [(i % len(mats), 32, rng.uniform(-1, 1, (mats[i % len(mats)][3], 32)).astype(np.float32))
            for i in range(30)]
    dev = []
    for mi, N, B in reqs:
        dev.append((dense_dev(B), torch().full((mats[mi][3], N), float("nan"),
                                               dtype=torch().float32, device="cuda")))
    torch().cuda.synchronize()
    for (mi, N, _), (Bd, C) in zip(reqs, dev):
        server.submit(mi, Bd.data_ptr(), N, N, C.data_ptr(), N)
    server.drain(120.0)
    st = server.stats()
    print(f"[serve evict] {st}", flush=True)
    assert st["completed"] == 30 and st["evictions"] > 0 and st["cache_misses"] > len(mats)
    _check_results(server, mats, reqs, dev, "f32")
    server.destroy()


def test_ticks_overlap_and_capacity_rejection():
    rng = np.random.default_rng(13)
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 3, 60000, 60000, mean=24)
    server = S.Server(cache_bytes=64 * MiB, inflight_bytes=512 * MiB, max_batch=2, streams=2)
    mid = server.add_matrix(pos, crd, vals, 60000)
    B = dense_dev(rng.uniform(-1, 1, (60000, 64)).astype(np.float32))
    Cs = [torch().empty((60000, 64), dtype=torch().float32, device="cuda") for _ in range(16)]
    torch().cuda.synchronize()
    for C in Cs:
        server.submit(mid, B.data_ptr(), 64, 64, C.data_ptr(), 64)
    server.drain(120.0)
    st = server.stats()
    print(f"[serve overlap] {st}", flush=True)
    assert st["completed"] == 16 and st["max_concurrent_ticks"] == 2
    ref = Cs[0].cpu().numpy()
    for C in Cs[1:]:
        assert np.array_equal(C.cpu().numpy(), ref)
    # a matrix larger than the whole cache can never be admitted
    big_pos, big_crd, big_vals = S.gen_csr(S.GEN_UNIFORM, 4, 200000, 200000, mean=64)
    big = server.add_matrix(big_pos, big_crd, big_vals, 200000)
    Bb = dense_dev(rng.uniform(-1, 1, (200000, 16)).astype(np.float32))
    Cb = torch().empty((200000, 16), dtype=torch().float32, device="cuda")
    rid = server.submit(big, Bb.data_ptr(), 16, 16, Cb.data_ptr(), 16)
    assert server.wait(rid, 10.0) == 5  # SEGTACO_ERR_CAPACITY
    server.destroy()


def test_poisson_replay_latency_stats():
    rng = np.random.default_rng(17)
    mats = _pool(19, 16)
    server = S.Server(cache_bytes=512 * MiB, inflight_bytes=512 * MiB, max_batch=32, streams=2)
    for p, c, v, K in mats:
        server.add_matrix(p, c, v, K)
    n = 300
    arrivals = np.cumsum(rng.exponential(1e6 / 2000.0, n))  # lambda = 2000 / s
    pop = 1.0 / np.arange(1, len(mats) + 1)
    mids = rng.choice(len(mats), n, p=pop / pop.sum())
    Ns = rng.choice([16, 32, 64, 128], n)
    Bs = {}
    Cs = []
    for mi, N in zip(mids, Ns):
        K = mats[mi][3]
        if (mi, N) not in Bs:
            Bs[(mi, N)] = dense_dev(rng.uniform(-1, 1, (K, N)).astype(np.float32))
        Cs.append(torch().empty((K, N), dtype=torch().float32, device="cuda"))
    torch().cuda.synchronize()
    ids = server.replay(arrivals, mids, [Bs[(m, N)].data_ptr() for m, N in zip(mids, Ns)],
                        Ns, Ns, [C.data_ptr() for C in Cs], Ns, timeout=120.0)
    assert len(ids) == n
    st = server.stats()
    print(f"[serve poisson] {st}", flush=True)
    assert st["completed"] == n and st["rejected"] == 0
    assert 0 < st["p50_us"] <= st["p99_us"] <= st["max_us"]
    # spot-check a few results
    for k in (0, n // 2, n - 1):
        mi, N = int(mids[k]), int(Ns[k])
        pos, crd, vals, K = mats[mi]
        B = Bs[(mi, N)].cpu().numpy()
        err = O.scaled_err(Cs[k].cpu().numpy(), O.spmm_f64(pos, crd, vals, B),
                           O.spmm_absbound(pos, crd, vals, B))
        assert err <= 1e-5
    server.destroy()


def test_one_matrix_many_widths_across_streams():
    """One matrix served at column counts whose plans use different tile sizes,
    one request per tick on two streams: the second stream builds its tile
    rows while the first stream's upload tick may still run, and one tick may
    add several tile sizes to the matrix. Every result matches the oracle and
    the same product run alone."""
    rng = np.random.default_rng(23)
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 29, 6000, 6000, mean=40)
    mats = [(pos, crd, vals, 6000)]
    Ns = [16, 64, 128, 32, 16, 128, 64, 32] * 3
    for max_batch in (1, 8):
        server = S.Server(cache_bytes=256 * MiB, inflight_bytes=256 * MiB,
                          max_batch=max_batch, streams=2)
        server.add_matrix(pos, crd, vals, 6000)
        tiles = {server.plan(0, N).tile for N in set(Ns)}
        print(f"[serve widths] max_batch {max_batch} tiles {sorted(tiles)}", flush=True)
        assert len(tiles) > 1, "widths should map to different tile sizes"
        reqs = [(0, N, rng.uniform(-1, 1, (6000, N)).astype(np.float32)) for N in Ns]
        dev = [(dense_dev(B), torch().full((6000, N), float("nan"), dtype=torch().float32,
                                           device="cuda")) for _, N, B in reqs]
        torch().cuda.synchronize()
        for (mi, N, _), (Bd, C) in zip(reqs, dev):
            server.submit(mi, Bd.data_ptr(), N, N, C.data_ptr(), N)
        server.drain(120.0)
        assert server.stats()["completed"] == len(reqs)
        _check_results(server, mats, reqs, dev, "f32")
        server.destroy()


def test_failed_tick_reports_errors_and_stops():
    """A tick whose launch fails (forced from the host: the launch reports a
    HIP error and none of its device work runs) retires its requests with
    SEGTACO_ERR_HIP instead of OK; requests still queued fail the same way,
    later submissions are refused and drain reports the error. Requests of
    ticks before the failure complete normally. The scheduler runs
    concurrently with the submitting thread, so of the five requests sent
    after arming the failure the first is always accepted (no tick has
    failed yet) and the others are accepted or, once the failing tick has
    been launched, refused."""
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 31, 5000, 5000, mean=20)
    server = S.Server(cache_bytes=64 * MiB, inflight_bytes=64 * MiB, max_batch=1, streams=1)
    mid = server.add_matrix(pos, crd, vals, 5000)
    B = dense_dev(S.gen_dense(32, 5000, 32))
    Cs = [torch().full((5000, 32), float("nan"), dtype=torch().float32, device="cuda")
          for _ in range(6)]
    torch().cuda.synchronize()
    ok_id = server.submit(mid, B.data_ptr(), 32, 32, Cs[0].data_ptr(), 32)
    assert server.wait(ok_id, 30.0) == 0
    server.debug_fail_tick(1)
    ids = []
    for k, C in enumerate(Cs[1:]):
        try:
            ids.append(server.submit(mid, B.data_ptr(), 32, 32, C.data_ptr(), 32))
        except S.SegtacoError as e:
            assert k > 0 and e.status == S.ERR_HIP
    statuses = [server.wait(i, 30.0) for i in ids]
    print(f"[serve fail] statuses {statuses} ({5 - len(ids)} refused)", flush=True)
    assert len(ids) >= 1
    assert all(s == S.ERR_HIP for s in statuses)
    with pytest.raises(S.SegtacoError) as ei:
        server.submit(mid, B.data_ptr(), 32, 32, Cs[0].data_ptr(), 32)
    assert ei.value.status == S.ERR_HIP
    with pytest.raises(S.SegtacoError):
        server.drain(10.0)
    st = server.stats()
    assert st["completed"] == 1 and st["rejected"] == len(ids)
    ref = O.spmm_f64(pos, crd, vals, S.gen_dense(32, 5000, 32))
    err = O.scaled_err(Cs[0].cpu().numpy(), ref,
                       O.spmm_absbound(pos, crd, vals, S.gen_dense(32, 5000, 32)))
    assert err <= 1e-5
    server.destroy()


def test_small_cache_one_matrix_many_widths():
    """A cache that holds the matrix and one tile-row array but not two: a
    request at another width (another tile size) releases the matrix's idle
    array instead of stalling, and every result is right."""
    rng = np.random.default_rng(37)
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 41, 20000, 20000, mean=24)
    M = nnz = None
    M, nnz = 20000, int(pos[-1])
    probe = S.Server(cache_bytes=256 * MiB, inflight_bytes=64 * MiB, max_batch=1, streams=1)
    probe.add_matrix(pos, crd, vals, M)
    Ns = [16, 128, 32, 64, 16, 128]
    tiles = {N: probe.plan(0, N).tile for N in Ns}
    probe.destroy()
    assert len(set(tiles.values())) > 1, tiles
    csr = -(-4 * (M + 1) // 256) * 256 + 2 * (-(-4 * nnz // 256) * 256)

    def rows_bytes(T):
        return 2 * (-(-4 * (-(-nnz // T) + 1) // 256) * 256)
    cache = csr + max(rows_bytes(T) for T in tiles.values()) + 256
    server = S.Server(cache_bytes=cache, inflight_bytes=256 * MiB, max_batch=1, streams=1)
    server.add_matrix(pos, crd, vals, M)
    mats = [(pos, crd, vals, M)]
    reqs = [(0, N, rng.uniform(-1, 1, (M, N)).astype(np.float32)) for N in Ns]
    dev = [(dense_dev(B), torch().full((M, N), float("nan"), dtype=torch().float32,
                                       device="cuda")) for _, N, B in reqs]
    torch().cuda.synchronize()
    ids = []
    for (mi, N, _), (Bd, C) in zip(reqs, dev):
        ids.append(server.submit(mi, Bd.data_ptr(), N, N, C.data_ptr(), N))
    statuses = [server.wait(i, 60.0) for i in ids]
    assert statuses == [0] * len(reqs), statuses
    _check_results(server, mats, reqs, dev, "f32")
    server.destroy()
