"""GPU parity of the block-dense schedule (bd): BF16 / FP16 storage, blocks
of 32 rows (default) on the matrix cores when their columns span at most
256; blocks of 64 rows (tile = 64) when they span at most 192, else each
32-row half on its own; a per-row vector walk otherwise.

Tolerances (as for every 16-bit path):
  * vs FP64 of the *rounded* inputs, oracle scaled metric: <= 1e-5;
  * vs FP64 of the original FP32 inputs, norm-wise: <= 1e-2.
Two runs are bit-identical (no atomics).
"""
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S
from gpu_util import DeviceCSR, dense_dev, run_spmm

pytestmark = pytest.mark.gpu


def _rows(rng, M, K, spans):
    """CSR whose row i holds sorted unique columns drawn from spans[i] =
    (lo, hi) (half-open, clipped to K) with the given density."""
    pos, crd = [0], []
    for lo, hi, dens in spans:
        lo, hi = max(0, lo), min(K, hi)
        n = int(round((hi - lo) * dens)) if hi > lo else 0
        cols = np.sort(rng.choice(np.arange(lo, hi), n, replace=False)) if n else []
        crd.extend(int(c) for c in cols)
        pos.append(len(crd))
    pos = np.array(pos, np.int32)
    crd = np.array(crd, np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    return pos, crd, vals


def _case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "banded_bw32":  # |i - j| <= 32, every entry stored: windows of 96
        pos, crd, vals = S.gen_csr(S.GEN_BANDED, 3, 3000, 3000, bw=32)
        return pos, crd, vals, 3000
    if name == "banded_ragged":  # M not a multiple of 32, empty rows and blocks
        M = K = 2011
        spans = [(i - 20, i + 21, 0.6) for i in range(M)]
        for i in list(range(100, 180)) + [5, 6, 2010]:
            spans[i] = (0, 0, 0.0)
        return (*_rows(rng, M, K, spans), K)
    if name == "window_edges":  # blocks of width 256 (matrix cores) and 257 (vector)
        M, K = 32 * 6, 2000
        spans = []
        for b, w in enumerate([256, 257, 200, 256, 255, 400]):
            c0 = 150 * b
            for i in range(32):
                spans.append((c0, c0 + w, 0.25))
            spans[32 * b] = (c0, c0 + 1, 1.0)  # pin the window's first column
            spans[32 * b + 31] = (c0 + w - 1, c0 + w, 1.0)  # and its last
        return (*_rows(rng, M, K, spans), K)
    if name == "window_edges64":  # 64-row blocks of width 192 (one tile) and 193 (halves)
        M, K = 64 * 4 + 17, 1500
        spans = []
        for b, w in enumerate([192, 193, 150, 192, 60]):
            c0 = 250 * b
            for i in range(64):
                spans.append((c0, c0 + w, 0.3))
            spans[64 * b] = (c0, c0 + 1, 1.0)
            spans[64 * b + 63] = (c0 + w - 1, c0 + w, 1.0)
        return (*_rows(rng, M, K, spans[:M]), K)
    if name == "powerlaw_fallback":  # nothing dense: all blocks on the vector path
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 4, 3000, 3000)
        return pos, crd, vals, 3000
    if name == "long_rows_fallback":  # rows above the compensation threshold
        M, K = 64, 6000
        spans = [(0, 6000, 0.5) if i in (3, 40) else (i * 10, i * 10 + 30, 0.5)
                 for i in range(M)]
        return (*_rows(rng, M, K, spans), K)
    if name == "near_k_edge":  # windows ending at K: staged rows past K are zeros
        M, K = 96, 100
        spans = [(max(0, K - 70 + (i % 7)), K, 0.5) for i in range(M)]
        return (*_rows(rng, M, K, spans), K)
    raise KeyError(name)


CASES = ["banded_bw32", "banded_ragged", "window_edges", "window_edges64",
         "powerlaw_fallback", "long_rows_fallback", "near_k_edge"]


def _run(pos, crd, vals, B, dtype, ldb=None, ldc=None, runs=1, tile=0):
    K, N = B.shape
    A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
    ldb = ldb or N
    Bp = np.zeros((K, ldb), np.float32)
    Bp[:, :N] = B
    outs = []
    for _ in range(runs):
        C = run_spmm(A, dense_dev(Bp, dtype), N, S.make_plan("bd", tile=tile), ldb=ldb,
                     ldc=ldc, b_dtype=dtype)
        outs.append(C[: A.M, :N].cpu().numpy())
    return outs


def _check(pos, crd, vals, B, C, dtype, what):
    vr = None if vals is None else O.round_to(dtype, vals)[1]
    Br = O.round_to(dtype, B)[1]
    err_r = O.scaled_err(C, O.spmm_f64(pos, crd, vr, Br), O.spmm_absbound(pos, crd, vr, Br))
    assert err_r <= 1e-5, f"{what}: vs FP64 of rounded inputs {err_r:.3e}"
    fro, _ = O.rel_norm_errs(C, O.spmm_f64(pos, crd, vals, B))
    assert fro <= 1e-2, f"{what}: vs FP64 of FP32 inputs {fro:.3e}"


@pytest.mark.parametrize("tile", [64, 32])
@pytest.mark.parametrize("N", [256, 128, 40, 7, 300])
@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
@pytest.mark.parametrize("case", CASES)
def test_bd_parity(case, dtype, N, tile):
    pos, crd, vals, K = _case(case)
    B = np.random.default_rng(N).uniform(-1, 1, (K, N)).astype(np.float32)
    print(f"[bd] {case} {dtype} N={N} tile={tile} M={len(pos)-1} nnz={len(crd)}", flush=True)
    C, C2 = _run(pos, crd, vals, B, dtype, runs=2, tile=tile)
    _check(pos, crd, vals, B, C, dtype, f"{case}/{dtype}/N={N}")
    assert np.array_equal(C.view(np.uint32), C2.view(np.uint32)), "not deterministic"


def _disorder(name):
    """A banded matrix whose blocks break the sorted-unique column rule in
    different ways: duplicate (row, column) entries (summed, as CSR stores
    them), rows in descending column order, one row whose entries leave the
    window its first and last entries suggest; blocks in between stay sorted
    and take the matrix cores."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pos, crd, vals = S.gen_csr(S.GEN_BANDED, 8, 32 * 12, 32 * 12, bw=24)
    rows = [list(crd[pos[i]:pos[i + 1]]) for i in range(len(pos) - 1)]
    rv = [list(vals[pos[i]:pos[i + 1]]) for i in range(len(pos) - 1)]
    for i in range(32 * 1, 32 * 2, 3):  # block 1: duplicates
        k = int(rng.integers(0, len(rows[i])))
        rows[i].insert(k, rows[i][k])
        rv[i].insert(k, float(rng.uniform(-1, 1)))
    for i in range(32 * 4, 32 * 5):  # block 4: every row descending
        rows[i], rv[i] = rows[i][::-1], rv[i][::-1]
    i = 32 * 7 + 5  # block 7: one row's middle entry far outside the window
    rows[i][len(rows[i]) // 2] = 32 * 12 - 1
    i = 32 * 9 + 31  # block 9: the last row's first entry repeats as its second
    rows[i].insert(1, rows[i][0])
    rv[i].insert(1, 0.25)
    pos = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    crd = np.array([c for r in rows for c in r], np.int32)
    vals = np.array([v for r in rv for v in r], np.float32)
    return pos, crd, vals, 32 * 12


@pytest.mark.parametrize("tile", [64, 32])
@pytest.mark.parametrize("N", [256, 64])
@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_bd_unsorted_and_duplicate_columns(dtype, N, tile):
    """bd checks each block's column order before scattering it: blocks with
    duplicate, unsorted or out-of-window entries take the per-row walk and
    still give the product of the stored entries (duplicates summed)."""
    pos, crd, vals, K = _disorder("disorder")
    B = np.random.default_rng(N + 1).uniform(-1, 1, (K, N)).astype(np.float32)
    print(f"[bd] disorder {dtype} N={N} tile={tile} nnz={len(crd)}", flush=True)
    C, C2 = _run(pos, crd, vals, B, dtype, runs=2, tile=tile)
    _check(pos, crd, vals, B, C, dtype, f"disorder/{dtype}/N={N}")
    assert np.array_equal(C.view(np.uint32), C2.view(np.uint32)), "not deterministic"


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_bd_strided_unaligned_and_novals(dtype):
    """ldb = 261 (no 16-byte B loads), ldc = 300, and A without values."""
    pos, crd, _, K = _case("banded_bw32")
    N = 256
    B = np.random.default_rng(9).uniform(-1, 1, (K, N)).astype(np.float32)
    (C,) = _run(pos, crd, None, B, dtype, ldb=261, ldc=300)
    _check(pos, crd, None, B, C, dtype, f"strided/{dtype}")


def test_batch_refuses_bd():
    """bd runs standalone only: a batch item with a bd plan is refused."""
    from gpu_util import torch
    t = torch()
    pos, crd, vals = S.gen_csr(S.GEN_BANDED, 1, 500, 500, bw=8)
    A = DeviceCSR(pos, crd, vals, 500, dtype="bf16")
    B = dense_dev(S.gen_dense(2, 500, 16), "bf16")
    C = t.empty((500, 16), dtype=t.float32, device="cuda")
    item = S.batch_item(A.desc, B.data_ptr(), 16, C.data_ptr(), 16, 16, S.make_plan("bd"))
    with pytest.raises(S.SegtacoError, match="standalone"):
        S.Batch([item], dtype="bf16")
