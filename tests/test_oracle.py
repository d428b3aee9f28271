"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

* Against the reference's own code: the committed golden vectors (Eigen
  3.4.0 products, ds.h's COO->CSR and MatrixMarket reader, written by
  tests/golden/make_golden.py) and, when built, oracle/_ref live.
* Against what the mathematics fixes: hand-traced known answers of the
  partition bisection, brute force, closed forms and special cases chosen so
  that a dropped term, a wrong index or a transposed operand fails.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    d = np.load(path, allow_pickle=False)
    vals = d["vals"] if int(d["has_vals"]) else None
    return d, vals


# ---------------------------------------------------------------------------
# Golden vectors from the reference's code
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_eigen_product_bitexact(path):
    d, vals = _load(path)
    C = O.spmm_f32_seq(d["pos"], d["crd"], vals, d["B"])
    assert np.array_equal(C, d["C_eigen"]), "Eigen-order restatement differs from Eigen"


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_f64_close_to_eigen(path):
    d, vals = _load(path)
    ref = O.spmm_f64(d["pos"], d["crd"], vals, d["B"])
    scale = O.spmm_absbound(d["pos"], d["crd"], vals, d["B"])
    assert O.scaled_err(d["C_eigen"], ref, scale) < 2e-6


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_partition(path):
    d, _ = _load(path)
    M = int(d["M"])
    for T in (32, 256):
        assert np.array_equal(O.block_starts(d["pos"], M, T), d[f"block_starts_{T}"])


def test_golden_coo_duplicates_summed():
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "tiny_dups.npz"))
    # (1,0), (2,4) and (5,4) appear twice and are summed; rows 0, 3, 6 empty.
    pos, crd, vals = d["pos"], d["crd"], d["vals"]
    assert list(pos) == [0, 0, 3, 4, 4, 6, 8, 8]
    assert list(crd[pos[1]:pos[2]]) == [0, 2, 3]
    assert vals[pos[1]] == np.float32(1.25)
    assert vals[pos[2]] == np.float32(1.25)
    assert list(crd[pos[5]:pos[6]]) == [0, 4] and vals[pos[5] + 1] == np.float32(-0.375)
    # explicit zero at (4,2) is kept as a stored entry
    assert 0.0 in list(vals[pos[4]:pos[5]])


# ---------------------------------------------------------------------------
# oracle/_ref live (the reference's Eigen path compiled in place)
# ---------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ref_live_bitexact(seed):
    rng = np.random.default_rng(seed)
    M, K, N = 400, 300, [1, 5, 33][seed - 1]
    lens = rng.integers(0, 30, M)
    lens[rng.integers(0, M, 40)] = 0
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, l, replace=False)) for l in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, pos[-1]).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    c_or = O.spmm_f32_seq(pos, crd, vals, B)
    for threads in (1, 3):
        assert np.array_equal(c_or, O.ref_eigen_spmm(pos, crd, vals, B, threads=threads))


@needs_ref
def test_ref_live_coo_to_csr_matches_fixture():
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "tiny_dups.npz"))
    pos, crd, vals = O.ref_coo_to_csr(int(d["M"]), int(d["K"]), d["coo_rows"], d["coo_cols"],
                                      d["coo_vals"])
    assert np.array_equal(pos, d["pos"]) and np.array_equal(crd, d["crd"])
    assert np.array_equal(vals, d["vals"])


# ---------------------------------------------------------------------------
# Partition searches: known answers and invariants
# ---------------------------------------------------------------------------

def test_search_before_known_answers():
    # Hand-traced bisection of taco_binarySearchBefore (gpu_library.h:185-205).
    pos = np.array([0, 0, 0, 0, 0, 0, 0, 0, 5], np.int32)  # 7 empty rows, then 5
    # target 0: pos[8]=5 > 0; lo=0, hi=8; mid=4: pos[4]=0 == 0 -> return 4
    assert O.search_before(pos, 0, 8, 0) == 4
    # target 3: mid=4 (0<3) lo=4; mid=6 lo=6; mid=7 lo=7 -> return 7
    assert O.search_before(pos, 0, 8, 3) == 7
    # target >= pos[end] returns end
    assert O.search_before(pos, 0, 8, 5) == 8
    pos2 = np.array([0, 2, 2, 2, 6, 6, 9], np.int32)
    # target 2: mid=3: pos[3]=2 -> exact hit 3 (a member of the run 1..3)
    assert O.search_before(pos2, 0, 6, 2) == 3
    # target 5: mid=3 (2<5) lo=3; mid=4 (6>5) hi=4 -> return 3
    assert O.search_before(pos2, 0, 6, 5) == 3
    # target 6: mid=3 lo=3; mid=4: pos[4]=6 -> 4 (row 4 is empty; repair -> 5)
    assert O.search_before(pos2, 0, 6, 6) == 4


def _random_pos(rng, M, empty_frac=0.4):
    lens = rng.integers(1, 6, M)
    lens[rng.random(M) < empty_frac] = 0
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_search_before_invariants_and_repair(seed):
    rng = np.random.default_rng(seed)
    M = int(rng.integers(1, 60))
    pos = _random_pos(rng, M)
    nnz = int(pos[M])
    for target in range(0, nnz + 2):
        r = O.search_before(pos, 0, M, target)
        assert 0 <= r <= M
        if target >= nnz:
            assert r == M
            continue
        assert pos[r] <= target
        # r lies in the run of equal pos values ending at the row holding target
        true_row = max(i for i in range(M) if pos[i] <= target < pos[i + 1])
        assert r <= true_row and pos[r + 1] >= target
        # either r is the row itself or r sits in the run of empty rows that
        # ends at it (every pos between them equals target)
        assert r == true_row or all(pos[i] == target for i in range(r, true_row + 1))
        # the reference's repair alone (spmm_eb_sr.h:39-42) reaches the row
        rr = r
        while pos[rr + 1] == target:
            rr += 1
        assert rr == true_row


@pytest.mark.parametrize("seed", range(4))
def test_block_starts_definition(seed):
    rng = np.random.default_rng(100 + seed)
    M = int(rng.integers(1, 200))
    pos = _random_pos(rng, M, 0.3)
    for T in (1, 3, 8, 32):
        bs = O.block_starts(pos, M, T)
        nb = (int(pos[M]) + T - 1) // T
        assert len(bs) == nb + 1
        for b in range(nb + 1):
            assert bs[b] == O.search_before(pos, 0, M, b * T)
        assert bs[nb] == M  # nb*T >= nnz: the early return of the search


@pytest.mark.parametrize("seed", range(4))
def test_segment_number_brute_force(seed):
    rng = np.random.default_rng(200 + seed)
    M = int(rng.integers(1, 80))
    pos = _random_pos(rng, M)
    for e in range(int(pos[M]) + 3):
        expect = max(i for i in range(M) if pos[i] <= e) if M > 0 else 0
        assert O.segment_number(pos, M, e) == expect


# ---------------------------------------------------------------------------
# Value oracles: closed forms, special cases, order
# ---------------------------------------------------------------------------

def _rand_csr(rng, M, K, density=0.2, with_vals=True):
    dense = (rng.random((M, K)) < density)
    pos = np.concatenate([[0], np.cumsum(dense.sum(1))]).astype(np.int32)
    crd = np.nonzero(dense)[1].astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32) if with_vals else None
    return pos, crd, vals, dense


@pytest.mark.parametrize("seed", range(5))
def test_f64_vs_dense_brute_force(seed):
    rng = np.random.default_rng(300 + seed)
    M, K, N = (int(x) for x in rng.integers(1, 64, 3))
    pos, crd, vals, mask = _rand_csr(rng, M, K)
    A = np.zeros((M, K))
    A[mask] = vals
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    ref = A @ B.astype(np.float64)
    got = O.spmm_f64(pos, crd, vals, B)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    # transposed-operand / wrong-index mistakes would not match:
    if M == K and M > 1:
        assert not np.allclose(A.T @ B, ref) or np.allclose(A, A.T)


def test_special_cases():
    rng = np.random.default_rng(7)
    n, N = 37, 5
    eye_pos = np.arange(n + 1, dtype=np.int32)
    eye_crd = np.arange(n, dtype=np.int32)
    B = rng.uniform(-1, 1, (n, N)).astype(np.float32)
    # A = I -> C = B (both oracles, exactly)
    assert np.array_equal(O.spmm_f32_seq(eye_pos, eye_crd, None, B), B)
    assert np.array_equal(O.spmm_f64(eye_pos, eye_crd, None, B), B.astype(np.float64))
    # B = I -> C = dense(A)
    pos, crd, vals, mask = _rand_csr(rng, 20, n, 0.3)
    A = np.zeros((20, n), np.float32)
    A[mask] = vals
    assert np.array_equal(O.spmm_f32_seq(pos, crd, vals, np.eye(n, dtype=np.float32)), A)
    # B = ones -> row sums; empty rows -> 0
    rs = O.spmm_f64(pos, crd, vals, np.ones((n, 1), np.float32))[:, 0]
    assert np.allclose(rs, A.astype(np.float64).sum(1))
    assert np.all(rs[np.diff(pos) == 0] == 0)
    # N = 1 is SpMV
    x = rng.uniform(-1, 1, (n, 1)).astype(np.float32)
    assert np.allclose(O.spmm_f64(pos, crd, vals, x)[:, 0], A.astype(np.float64) @ x[:, 0])


def test_f32_sequential_order_known_answer():
    # Row entries in order 1, 1e8, -1e8 (B = 1): sequential FP32 gives
    # ((0 + 1) + 1e8) - 1e8 = 0, reversed order would give 1; FP64 gives 1.
    pos = np.array([0, 3], np.int32)
    crd = np.array([0, 1, 2], np.int32)
    vals = np.array([1.0, 1e8, -1e8], np.float32)
    B = np.ones((3, 1), np.float32)
    assert O.spmm_f32_seq(pos, crd, vals, B)[0, 0] == 0.0
    assert O.spmm_f64(pos, crd, vals, B)[0, 0] == 1.0


def test_error_metric_catches_planted_mistakes():
    rng = np.random.default_rng(11)
    pos, crd, vals, _ = _rand_csr(rng, 50, 40, 0.3)
    B = rng.uniform(-1, 1, (40, 6)).astype(np.float32)
    ref = O.spmm_f64(pos, crd, vals, B)
    scale = O.spmm_absbound(pos, crd, vals, B)
    good = O.spmm_f32_seq(pos, crd, vals, B)
    assert O.scaled_err(good, ref, scale) < 1e-6
    # dropped term
    bad_vals = vals.copy()
    bad_vals[len(vals) // 2] = 0
    assert O.scaled_err(O.spmm_f32_seq(pos, crd, bad_vals, B), ref, scale) > 1e-3
    # wrong (shifted) column index
    bad_crd = crd.copy()
    bad_crd[3] = (bad_crd[3] + 1) % 40
    assert O.scaled_err(O.spmm_f32_seq(pos, bad_crd, vals, B), ref, scale) > 1e-3
    # column-permuted B (invisible with the reference's B = ones)
    assert O.scaled_err(O.spmm_f32_seq(pos, crd, vals, B[:, ::-1].copy()), ref, scale) > 1e-3
    # NaN
    nan = good.copy()
    nan[0, 0] = np.nan
    assert O.scaled_err(nan, ref, scale) == float("inf")


def test_half_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 300,
                        np.array([1e-8, 6e-8, 2.98e-8, 3e-5, 65504, 65520, 1e6, -0.0, np.inf,
                                  1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9], np.float32)])
    for name, tdt in (("bf16", torch.bfloat16), ("fp16", torch.float16)):
        bits, back = O.round_to(name, x)
        tb = torch.from_numpy(x).to(tdt)
        assert np.array_equal(bits, tb.view(torch.int16).numpy().view(np.uint16))
        assert np.array_equal(back, tb.float().numpy())


# ---------------------------------------------------------------------------
# The error scale, per-entry rows and the tile partition the GPU checks use
# ---------------------------------------------------------------------------

def _dense_absbound(pos, crd, vals, B):
    """|A| |B| in float64 from a dense copy of A."""
    M = len(pos) - 1
    A = np.zeros((M, B.shape[0]))
    rows = np.repeat(np.arange(M), np.diff(pos))
    A[rows, crd] = np.abs(vals.astype(np.float64)) if vals is not None else 1.0
    return A @ np.abs(B.astype(np.float64))


def _matches(S, ref):
    return S.shape == ref.shape and np.allclose(S, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("seed", range(6))
def test_absbound_vs_dense_brute_force(seed):
    rng = np.random.default_rng(400 + seed)
    M, K, N = (int(x) for x in rng.integers(1, 70, 3))
    pos, crd, vals, _ = _rand_csr(rng, M, K, 0.25)
    empty = rng.random(M) < 0.3  # drop whole rows: empty rows must give 0
    keep = ~np.repeat(empty, np.diff(pos))
    crd, vals = crd[keep], vals[keep]
    pos = np.concatenate([[0], np.cumsum(np.where(empty, 0, np.diff(pos)))]).astype(np.int32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    for v in (vals, None):
        S = O.spmm_absbound(pos, crd, v, B)
        ref = _dense_absbound(pos, crd, v, B)
        assert _matches(S, ref)
        assert np.all(S[np.diff(pos) == 0] == 0)
        # planted mistakes in the scale that would loosen or break every
        # tolerance stated on it: doubled, signs kept (|sum| instead of
        # sum of |.|), one column of B shifted
        assert not _matches(2.0 * S, ref)
        signed = np.abs(O.spmm_f64(pos, crd, v, B))
        if np.any(S > 0):
            assert not _matches(signed, ref)
        if N > 1 and np.any(S > 0):
            assert not _matches(np.roll(S, 1, axis=1), ref) or np.allclose(S, S[:, :1])


def test_absbound_planted_scale_loosens_metric():
    """A scale twice too large halves every scaled error: an error just over
    the 1e-5 bar passes under it. So the scale must be the exact one."""
    rng = np.random.default_rng(17)
    pos, crd, vals, _ = _rand_csr(rng, 40, 30, 0.3)
    B = rng.uniform(-1, 1, (30, 4)).astype(np.float32)
    ref = O.spmm_f64(pos, crd, vals, B)
    S = O.spmm_absbound(pos, crd, vals, B)
    i, k = np.unravel_index(np.argmax(S), S.shape)
    C = ref.astype(np.float32)
    C[i, k] = np.float32(ref[i, k] + 1.5e-5 * S[i, k])
    assert O.scaled_err(C, ref, S) > 1e-5
    assert O.scaled_err(C, ref, 2.0 * S) < 1e-5


@pytest.mark.parametrize("seed", range(5))
def test_row_of_nnz_brute_force(seed):
    rng = np.random.default_rng(500 + seed)
    M = int(rng.integers(1, 300))
    pos = _random_pos(rng, M, 0.4)
    rows = O.row_of_nnz(pos, M)
    assert np.array_equal(rows, np.repeat(np.arange(M), np.diff(pos)))
    for e in range(int(pos[M])):
        assert rows[e] == O.segment_number(pos, M, e)


def _tile_rows_brute(pos, T):
    M = len(pos) - 1
    rows = np.repeat(np.arange(M), np.diff(pos))
    nt = (len(rows) + T - 1) // T
    return np.concatenate([rows[np.arange(nt) * T], [M]]).astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_tile_rows_brute_force(seed):
    rng = np.random.default_rng(600 + seed)
    M = int(rng.integers(1, 400))
    pos = _random_pos(rng, M, 0.5)
    pos_runs = pos.copy()
    for T in (1, 2, 3, 8, 32, 256):
        got = O.tile_rows(pos, M, T)
        assert np.array_equal(got, _tile_rows_brute(pos, T)), T
    # planted: the K21 search without the repair step lands inside runs of
    # empty rows; with empty rows at tile starts it must differ
    # (seven empty rows, then 5 entries: the search for entry 0 stops at row
    # 4, the hand-traced answer above, and the repair walks on to row 7)
    T = 4
    pos_runs = np.array([0, 0, 0, 0, 0, 0, 0, 0, 5], np.int32)
    no_repair = O.block_starts(pos_runs, 8, T)
    assert list(no_repair) == [4, 7, 8]
    assert list(O.tile_rows(pos_runs, 8, T)) == [7, 7, 8] == list(_tile_rows_brute(pos_runs, T))


def _fix_t0_brute(pos, T):
    """Tile by tile: the row holding the tile's first entry ends in this tile
    after starting in an earlier one -> that earlier tile."""
    M = len(pos) - 1
    rows = np.repeat(np.arange(M), np.diff(pos))
    nt = (len(rows) + T - 1) // T
    out = np.full(nt, -1, np.int32)
    for b in range(nt):
        r = rows[b * T]
        p0, p1 = int(pos[r]), int(pos[r + 1])
        if p0 < b * T and p1 <= (b + 1) * T:
            out[b] = p0 // T
    return out


@pytest.mark.parametrize("seed", range(6))
def test_fix_t0_brute_force(seed):
    rng = np.random.default_rng(700 + seed)
    M = int(rng.integers(1, 300))
    lens = rng.integers(0, 40, M)
    lens[rng.random(M) < 0.3] = 0
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    for T in (1, 3, 8, 16, 64):
        got = O.fix_t0(pos, M, T)
        ref = _fix_t0_brute(pos, T)
        assert np.array_equal(got, ref), T
    # planted: recording the start tile for every continued row (not only in
    # the tile where it ends) differs whenever a row spans three tiles
    lens = np.array([2, 20, 3], np.int64)
    pos3 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ref = _fix_t0_brute(pos3, 4)
    assert list(ref) == [-1, -1, -1, -1, -1, 0, 5]
    rows = np.repeat(np.arange(3), lens)
    planted = np.array([pos3[rows[b * 4]] // 4 if pos3[rows[b * 4]] < b * 4 else -1
                        for b in range(len(ref))], np.int32)
    assert not np.array_equal(planted, ref)
    assert np.array_equal(O.fix_t0(pos3, 3, 4), ref)
