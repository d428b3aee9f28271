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


@pytest.mark.parametrize("case", ["one_row", "one_tile", "leading_and_trailing_empty"])
def test_bit_identical_edge_structures(case):
    """A single row split over many tiles (heads and tails only), fewer
    entries than one tile, and a matrix whose first and last thousand rows
    are empty."""
    rng = np.random.default_rng(17)
    if case == "one_row":
        lens, kw = np.array([20000]), dict(tile=128)
    elif case == "one_tile":
        lens, kw = rng.integers(0, 5, 40), dict(tile=512)
    else:
        lens = rng.integers(0, 30, 8000)
        lens[:1000] = 0
        lens[-1000:] = 0
        kw = {}
    pos, crd, vals, B = _mk(rng, lens, 30000, 40)
    C_cs, C_sr, _ = _pair(pos, crd, vals, B, **kw)
    assert _bound(pos, crd, vals, B, C_cs) <= TOL
    assert np.array_equal(C_cs.view(np.uint32), C_sr.view(np.uint32))


def test_lds_overflow_units_walk():
    """One entry per row: a unit of 64 tiles x 512 entries holds 32768 interior
    rows, far beyond the LDS accumulators (1241 slots at vec 4), so every wave
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
    tiles over one-entry rows holds 1024 interior rows, which fit the 1113
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


def _vbits(vals, dtype, nnz):
    """Value bits a record carries per entry: the FP32 word, or the 16-bit
    storage word zero-extended; 1.0 in the storage type without values."""
    if vals is None:
        one = {"f32": 0x3F800000, "bf16": 0x3F80, "fp16": 0x3C00}[dtype]
        return np.full(nnz, one, np.uint32)
    if dtype == "f32":
        return np.asarray(vals, np.float32).view(np.uint32)
    return S.to_half(dtype, vals).view(np.uint16).astype(np.uint32)


def _check_preparation(pos, crd, vals, K, N, dtype="f32", **kw):
    """segtaco_spmm_prepare for eb_cs, then every array it leaves in the
    workspace checked against what the CSR itself fixes: each tile's head /
    tail rows and interior rows (non-empty rows whose entries all lie in the
    tile), each entry's in-tile row code, the per-wave sweep / walk choice,
    and the records — a walking tile's in CSR order, a sweeping tile's a
    permutation of its entries, ordered by (column, row), whose slots name
    the entries' own rows, with no slot shared by two tiles. Returns the
    resolved plan and the share of tiles that walk."""
    t = torch()
    pos = np.asarray(pos, np.int64)
    crd = np.asarray(crd, np.int64)
    A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
    bd = S.DTYPES[dtype]
    p = S.plan_resolve(A.desc, N, bd, S.make_plan("eb_cs", **kw), N, N)
    need = S.workspace_bytes(A.desc, N, p)
    ws = t.zeros(need, dtype=t.uint8, device="cuda")
    S.spmm_prepare(A.desc, bd, N, N, N, p, ws.data_ptr(), need, stream_ptr())
    t.cuda.synchronize()
    eb = S.workspace_layout(A.desc, N, p)
    cs = S.cs_workspace_layout(A.desc, N, p)
    w = ws.cpu().numpy()

    def arr(off, dt, n):
        return np.frombuffer(w, dtype=dt, count=n, offset=off).astype(np.int64)

    M, nnz, T, G, cap = len(pos) - 1, int(pos[-1]), p.tile, p.group, cs["cap"]
    nt = eb["ntiles"]
    assert nt == -(-nnz // T) and cs["units"] == -(-nt // G)
    tile_row = arr(eb["rows_off"], np.int32, nt + 1)
    flags = arr(cs["tile_flags_off"], np.uint8, nt)
    count = arr(cs["tile_count_off"], np.int32, nt)
    n_slots = arr(cs["n_slots_off"], np.int32, cs["units"])
    slot_rows = arr(cs["slot_rows_off"], np.int32, cs["units"] * cap)
    int_rows = arr(cs["int_rows_off"], np.int32, nt * T)
    codes = arr(cs["codes_off"], np.uint16, nnz)

    row = np.repeat(np.arange(M), np.diff(pos))  # row of every entry
    tile = np.arange(nnz) // T
    s = np.arange(nt) * T
    e = np.minimum(s + T, nnz)
    first, last = row[s], row[e - 1]
    assert np.array_equal(tile_row[:nt], first)
    head = pos[first] < s
    tail = (pos[last + 1] > e) & ~(head & (last == first))
    assert np.array_equal(flags & 3, head.astype(np.int64) | (tail.astype(np.int64) << 1))
    ne = np.flatnonzero(np.diff(pos) > 0)
    interior = ne[pos[ne] // T == (pos[ne + 1] - 1) // T]
    it = pos[interior] // T
    assert np.array_equal(count, np.bincount(it, minlength=nt))
    k_of = np.arange(len(interior)) - np.searchsorted(it, it)
    assert np.array_equal(int_rows[it * T + k_of], interior)
    kk = np.full(M, -1)
    kk[interior] = k_of
    in_head = head[tile] & (row == first[tile])
    in_tail = ~in_head & tail[tile] & (row == last[tile])
    exp_code = np.where(in_head, 0, np.where(in_tail, 1, 2 + kk[row]))
    assert (exp_code >= 0).all()
    assert np.array_equal(codes, exp_code)

    # the slot plan: per unit, waves of 8 tiles in order sweep while their
    # interior rows fit beside the earlier sweeping waves'
    walk = (flags & 4) != 0
    for u in range(cs["units"]):
        used = 0
        for wv in range(8):
            lo, hi = u * G + 8 * wv, min(u * G + min(8 * wv + 8, G), nt)
            if lo >= hi:
                continue
            seg = int(count[lo:hi].sum())
            fits = used + seg <= cap - 2 * G
            assert (~walk[lo:hi]).all() if fits else walk[lo:hi].all(), (u, wv)
            used += seg if fits else 0
        assert n_slots[u] == 2 * G + used <= cap

    vb = _vbits(vals, dtype, nnz)
    words = 2 if cs["packed"] else 3
    raw = arr(cs["rec_off"], np.uint32, nnz * words).reshape(nnz, words)
    if cs["packed"]:
        rcol, rcode, rval = raw[:, 0] >> 12, raw[:, 0] & 4095, raw[:, 1]
    else:
        rcol, rcode, rval = raw[:, 0], raw[:, 1], raw[:, 2]
    wt = walk[tile]
    assert np.array_equal(rcol[wt], crd[wt])
    assert np.array_equal(rval[wt], vb[wt])
    assert np.array_equal(rcode[wt], codes[wt])
    sw = ~wt
    st, sc, sv, sk = tile[sw], rcol[sw], rval[sw], rcode[sw]
    unit, g = st // G, st % G
    assert (sk < n_slots[unit]).all()
    assert ((sk >= 2 * G) | (sk == 2 * g) | (sk == 2 * g + 1)).all()
    assert (head[st] | (sk != 2 * g)).all() and (tail[st] | (sk != 2 * g + 1)).all()
    rrow = np.where(sk == 2 * g, first[st],
                    np.where(sk == 2 * g + 1, last[st], slot_rows[unit * cap + sk]))
    # a permutation of each tile's entries ...
    o1 = np.lexsort((sc, rrow, st))
    o2 = np.lexsort((crd[sw], row[sw], tile[sw]))
    assert np.array_equal(st[o1], tile[sw][o2])
    assert np.array_equal(rrow[o1], row[sw][o2])
    assert np.array_equal(sc[o1], crd[sw][o2])
    assert np.array_equal(sv[o1], vb[sw][o2])
    # ... in (column, row) order ...
    same = st[1:] == st[:-1]
    up = (sc[1:] > sc[:-1]) | ((sc[1:] == sc[:-1]) & (rrow[1:] > rrow[:-1]))
    assert (up | ~same).all()
    # ... and no interior slot serves two tiles
    isl = sk >= 2 * G
    key = unit[isl] * cap + sk[isl]
    o = np.argsort(key, kind="stable")
    dup = key[o][1:] == key[o][:-1]
    assert (st[isl][o][1:][dup] == st[isl][o][:-1][dup]).all()
    return p, float(walk.mean())


@pytest.mark.parametrize("case", ["powerlaw", "banded", "one_column_in_every_row",
                                  "one_entry_rows", "mixed", "wide", "no_values", "bf16",
                                  "empty_runs", "one_row", "one_tile"])
def test_preparation_arrays(case):
    """The preparation's arrays on structures that take each of its paths:
    the 8- and 16-key bucket sorts (power-law, banded), the whole-wave sort
    (a column shared by every row fills one bucket), walking units, units
    mixing both, 96-bit records, A without values, 16-bit values."""
    rng = np.random.default_rng(31)
    kw, K, dtype, vals_none = {}, 30000, "f32", False
    if case in ("powerlaw", "no_values", "bf16"):
        pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 5, 30000, 30000)
        vals_none = case == "no_values"
        dtype = "bf16" if case == "bf16" else "f32"
        kw = dict(tile=256) if case == "powerlaw" else {}
    elif case == "banded":
        pos, crd, vals = S.gen_csr(S.GEN_BANDED, 3, 20000, 20000, bw=8)
        K = 20000
    elif case == "one_column_in_every_row":
        lens = rng.integers(1, 6, 20000)
        pos, crd, vals, _ = _mk(rng, lens, 30000, 1)
        crd[pos[:-1]] = 0  # each row's first entry in column 0 (rows stay sorted)
        kw = dict(tile=256)  # ~85 rows, so ~85 keys in one bucket, per tile
    elif case == "one_entry_rows":
        M = 70000
        pos = np.arange(M + 1, dtype=np.int32)
        crd = rng.integers(0, 5000, M).astype(np.int32)
        vals = rng.uniform(-1, 1, M).astype(np.float32)
        K, kw = 5000, dict(tile=512)
    elif case == "mixed":
        lens = np.concatenate([np.ones(60000, np.int64), rng.integers(20, 60, 6000)])
        pos, crd, vals, _ = _mk(rng, lens, 30000, 1)
        kw = dict(tile=128)
    elif case == "empty_runs":
        # empty rows first, last and in long runs inside tiles, and a row
        # wider than the 64-row step of the row walk
        lens = rng.integers(0, 12, 9000)
        lens[:700] = 0
        lens[3000:3400] = 0
        lens[5000:5600:3] = 0
        lens[-900:] = 0
        lens[4000] = 2500
        pos, crd, vals, _ = _mk(rng, lens, 30000, 1)
        kw = dict(tile=256)
    elif case == "one_row":  # every tile inside one row: heads and tails only
        pos, crd, vals, _ = _mk(rng, np.array([20000]), 30000, 1)
        kw = dict(tile=128)
    elif case == "one_tile":  # fewer entries than one tile
        pos, crd, vals, _ = _mk(rng, rng.integers(0, 5, 40), 30000, 1)
        kw = dict(tile=512)
    else:  # wide
        K = (1 << 20) + 4099
        lens = np.concatenate([rng.integers(0, 30, 8000), np.ones(20000, np.int64)])
        pos, crd, vals, _ = _mk(rng, lens, K, 1)
        kw = dict(tile=256)
    p, walked = _check_preparation(pos, crd, None if vals_none else vals, K, 64, dtype, **kw)
    print(f"[cs prep] {case}: plan {p}, walking tiles {walked:.3f}", flush=True)
    if case == "one_entry_rows":
        # 8 tiles of 512 one-entry rows overflow the slots: every full wave
        # walks (only the last unit's lone 368-entry tile fits and sweeps)
        assert walked > 0.99
    if case == "mixed":
        assert 0.0 < walked < 1.0


def test_preparation_arrays_full_size_powerlaw():
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    p, walked = _check_preparation(pos, crd, vals, 262144, 128)
    print(f"[cs prep full] plan {p}, walking tiles {walked:.3f}", flush=True)


def test_batch_refuses_eb_cs():
    t = torch()
    pos, crd, vals = S.gen_csr(S.GEN_UNIFORM, 1, 500, 500, mean=8)
    A = DeviceCSR(pos, crd, vals, 500)
    B = dense_dev(S.gen_dense(2, 500, 16))
    C = t.empty((500, 16), dtype=t.float32, device="cuda")
    item = S.batch_item(A.desc, B.data_ptr(), 16, C.data_ptr(), 16, 16, S.make_plan("eb_cs"))
    with pytest.raises(S.SegtacoError, match="standalone"):
        S.Batch([item])
