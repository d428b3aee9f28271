"""Bit-exact GPU checks of the row partitions every element-balanced schedule
runs on, against the oracle:

* the first row of each tile (the reference's K21 search for b*T,
  gpu_library.h:185-215, followed by the repair loop of spmm_eb_sr.h:39-42)
  from both device routines that compute it (one scatter pass over pos for
  standalone calls and SDDMM, the search-and-repair per tile for batches and
  the server), and the unrepaired K21 array beside it;
* what a standalone eb_sr / eb_pr / eb_cs call leaves in its workspace: the
  tile rows and the fix-up's per-tile records;
* the row of every stored entry as eb_pr derives it (the segment ids of its
  segmented reduction; the reference's binary_search_segment_number,
  gpu_library.h:138-152);
* the K21 array and tile rows a batch holds for each product.

Integer work: every comparison is array_equal.
"""
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, dev, run_spmm, stream_ptr, torch

pytestmark = pytest.mark.gpu

TILES = [8, 32, 128, 256, 1024]


def _structure(name):
    """Row lengths of a named structure (seeded)."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "empty_runs":
        # runs of empty rows, including at both ends and across tile starts
        lens = rng.integers(1, 12, 6000)
        for s in rng.integers(0, 6000 - 80, 40):
            lens[s: s + rng.integers(1, 80)] = 0
        lens[:31] = 0
        lens[-29:] = 0
        return lens, 700
    if name == "sparse_rows":
        # one row in ~12 holds entries: tiles span more rows than the LDS
        # window of the row pointer
        return np.where(rng.random(40000) < 0.08, rng.integers(1, 4, 40000), 0), 3000
    if name == "powerlaw":
        pos, _, _ = S.gen_csr(S.GEN_POWERLAW, 5, 30000, 30000)
        return np.diff(pos), 30000
    if name == "long_rows":
        return np.array([0, 3, 9000, 0, 0, 1, 4096, 1024, 0, 1023, 1025, 2, 0], np.int64), 12000
    if name == "single_entry_rows":
        return np.ones(5000, np.int64), 5000
    raise KeyError(name)


STRUCTS = ["empty_runs", "sparse_rows", "powerlaw", "long_rows", "single_entry_rows"]


def _csr(name):
    lens, K = _structure(name)
    rng = np.random.default_rng(zlib.crc32(name.encode()) + 1)
    lens = np.minimum(lens, K)
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         + [np.zeros(0, np.int64)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    return pos, crd, vals, K


@pytest.mark.parametrize("T", TILES)
@pytest.mark.parametrize("name", STRUCTS)
def test_tile_rows_both_routines_bitexact(name, T):
    t = torch()
    pos, _, _, _ = _csr(name)
    M, nnz = len(pos) - 1, int(pos[-1])
    nt = (nnz + T - 1) // T
    print(f"[partition] tile rows {name} T={T} M={M} nnz={nnz}", flush=True)
    want = O.tile_rows(pos, M, T)
    dpos = dev(pos)
    rows = t.full((nt + 1,), -7, dtype=t.int32, device="cuda")
    S.tile_rows(dpos.data_ptr(), M, T, nt, rows.data_ptr(), S.ROWS_SCATTER, None, stream_ptr())
    t.cuda.synchronize()
    assert np.array_equal(rows.cpu().numpy(), want), "scatter routine"
    rows2 = t.full((nt + 1,), -7, dtype=t.int32, device="cuda")
    part = t.full((nt + 1,), -7, dtype=t.int32, device="cuda")
    S.tile_rows(dpos.data_ptr(), M, T, nt, rows2.data_ptr(), S.ROWS_SEARCH, part.data_ptr(),
                stream_ptr())
    t.cuda.synchronize()
    assert np.array_equal(rows2.cpu().numpy(), want), "search + repair routine"
    assert np.array_equal(part.cpu().numpy(), O.block_starts(pos, M, T, nt)), "K21 array"


def _workspace_partition(pos, crd, vals, K, sched, kw, N=8):
    """Runs one standalone call and returns (plan, tile rows, fix records, C)."""
    t = torch()
    A = DeviceCSR(pos, crd, vals, K)
    B = S.gen_dense(41, K, N)
    Bd = dense_dev(B)
    plan = S.plan_resolve(A.desc, N, S.DT_F32, S.make_plan(sched, **kw))
    ws_bytes = S.workspace_bytes(A.desc, N, plan)
    ws = t.full((max(ws_bytes, 1),), 0xAB, dtype=t.uint8, device="cuda")
    C = t.full((A.M, N), float("nan"), dtype=t.float32, device="cuda")
    S.spmm(A.desc, Bd.data_ptr(), S.DT_F32, N, N, C.data_ptr(), N, plan, ws.data_ptr(),
           ws_bytes, stream_ptr())
    t.cuda.synchronize()
    lay = S.workspace_layout(A.desc, N, plan)
    w = ws.cpu().numpy()
    nt = lay["ntiles"]
    rows = w[lay["rows_off"]: lay["rows_off"] + 4 * (nt + 1)].view(np.int32)
    fix = w[lay["fix_off"]: lay["fix_off"] + 4 * nt].view(np.int32)
    return plan, rows, fix, C.cpu().numpy(), B


@pytest.mark.parametrize("sched,kw", [
    ("eb_sr", {}), ("eb_sr", dict(tile=8)), ("eb_sr", dict(tile=256)), ("eb_sr", dict(tile=1024)),
    ("eb_pr", {}), ("eb_pr", dict(tile=32)), ("eb_pr", dict(tile=1024)),
    ("eb_cs", {}), ("eb_cs", dict(tile=32)),
], ids=lambda v: str(v) if not isinstance(v, dict) else "-".join(f"{k}{x}" for k, x in v.items()))
@pytest.mark.parametrize("name", ["empty_runs", "sparse_rows", "powerlaw", "long_rows"])
def test_standalone_call_partition_bitexact(name, sched, kw):
    pos, crd, vals, K = _csr(name)
    M = len(pos) - 1
    print(f"[partition] workspace {name} {sched} {kw}", flush=True)
    plan, rows, fix, C, B = _workspace_partition(pos, crd, vals, K, sched, kw)
    T = plan.tile
    assert np.array_equal(rows, O.tile_rows(pos, M, T)), f"tile rows, T={T}"
    assert np.array_equal(fix, O.fix_t0(pos, M, T)), f"fix-up records, T={T}"
    ref = O.spmm_f64(pos, crd, vals, B)
    scale = O.spmm_absbound(pos, crd, vals, B)
    assert O.scaled_err(C, ref, scale) <= 1e-5


@pytest.mark.parametrize("T", [8, 64, 128, 1000, 4096])
@pytest.mark.parametrize("name", STRUCTS)
def test_eb_pr_entry_rows_bitexact(name, T):
    t = torch()
    pos, _, _, _ = _csr(name)
    M, nnz = len(pos) - 1, int(pos[-1])
    nt = (nnz + T - 1) // T
    print(f"[partition] entry rows {name} T={T} nnz={nnz}", flush=True)
    dpos = dev(pos)
    trows = t.empty((nt + 1,), dtype=t.int32, device="cuda")
    S.tile_rows(dpos.data_ptr(), M, T, nt, trows.data_ptr(), S.ROWS_SCATTER, None, stream_ptr())
    out = t.full((max(nnz, 1),), -7, dtype=t.int32, device="cuda")
    S.entry_rows(dpos.data_ptr(), M, nnz, T, trows.data_ptr(), out.data_ptr(), stream_ptr())
    t.cuda.synchronize()
    got = out.cpu().numpy()[:nnz]
    assert np.array_equal(got, O.row_of_nnz(pos, M))
    # the reference's segment search at a sample of entries (it equals the
    # row of the entry, pinned exhaustively on the CPU)
    for e in np.linspace(0, nnz - 1, 64).astype(np.int64):
        assert got[e] == O.segment_number(pos, M, int(e))


def test_batch_partition_bitexact():
    """A batch computes each product's K21 array and tile rows once, at
    creation: both bit-exact against the oracle for eb_sr and eb_pr items of
    several tiles; row-balanced items hold none."""
    t = torch()
    items, keep, expect = [], [], []
    specs = [("empty_runs", "eb_sr", 32), ("sparse_rows", "eb_pr", 128),
             ("powerlaw", "eb_sr", 256), ("long_rows", "eb_pr", 64),
             ("empty_runs", "eb_pr", 1024), ("powerlaw", "rb_sr", 0),
             ("single_entry_rows", "eb_sr", 8)]
    for name, sched, T in specs:
        pos, crd, vals, K = _csr(name)
        N = 16
        A = DeviceCSR(pos, crd, vals, K)
        Bd = dense_dev(S.gen_dense(5, K, N))
        C = t.zeros((A.M, N), dtype=t.float32, device="cuda")
        keep += [A, Bd, C]
        plan = S.make_plan(sched, tile=T) if T else S.make_plan(sched)
        items.append(S.batch_item(A.desc, Bd.data_ptr(), N, C.data_ptr(), N, N, plan))
        expect.append((pos, T if sched.startswith("eb") else 0))
    print(f"[partition] batch of {len(items)}", flush=True)
    b = S.Batch(items, "f32", stream_ptr())
    for i, (pos, T) in enumerate(expect):
        part, rows = b.partition(i)
        if T == 0:
            assert part is None and rows is None
            continue
        M = len(pos) - 1
        assert b.plan(i).tile == T
        assert np.array_equal(part, O.block_starts(pos, M, T)), f"item {i} K21"
        assert np.array_equal(rows, O.tile_rows(pos, M, T)), f"item {i} tile rows"
    b.destroy()
