"""CPU checks of the autotuner's host logic: row statistics, model choices,
signature and plan cache (no device work)."""
import os

import numpy as np
import pytest

import segtaco_amd as S


def _pos(lens):
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)


def test_row_stats_match_numpy():
    rng = np.random.default_rng(3)
    lens = rng.integers(0, 3000, 5000)
    lens[::7] = 0
    st = S.row_stats(_pos(lens), 9000)
    assert st.M == 5000 and st.nnz == lens.sum() and st.max_len == lens.max()
    assert st.empty_rows == (lens == 0).sum()
    assert st.mean_len == pytest.approx(lens.mean())
    assert st.cv == pytest.approx(lens.std() / lens.mean(), rel=1e-9)
    assert st.frac_nnz_long == pytest.approx(lens[lens > 1024].sum() / lens.sum())
    hist = np.array(st.hist)
    assert hist.sum() == 5000 and hist[0] == (lens == 0).sum()
    for b in range(1, 13):
        assert hist[b] == ((lens >= 2 ** (b - 1)) & (lens < 2 ** b)).sum()


def test_model_choices():
    pl = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)[0]
    p = S.tune_model(S.row_stats(pl, 262144), 128)
    assert p.schedule == S.EB_SR and p.tile == 128 and p.vec == 4
    bd = S.gen_csr(S.GEN_BANDED, 9, 100000, 100000, bw=32)[0]
    p = S.tune_model(S.row_stats(bd, 100000), 256, "bf16")
    assert p.schedule == S.RB_SR and p.vec == 8 and p.comp_threshold == 1024
    un = S.gen_csr(S.GEN_UNIFORM, 1, 4096, 4096, mean=16)[0]
    p = S.tune_model(S.row_stats(un, 4096), 32)
    assert p.schedule == S.RB_SR
    rows_per_block = 4 * (64 // p.col_lanes) * p.group
    assert rows_per_block * 16 >= 1024  # a workgroup carries ~1K entries
    empty = S.tune_model(S.row_stats(np.zeros(11, np.int32), 5), 8)
    assert empty.schedule == S.RB_SR


def test_signature_and_cache(tmp_path):
    pl = S.gen_csr(S.GEN_POWERLAW, 5, 20000, 20000)[0]
    st = S.row_stats(pl, 20000)
    sig = S.tune_signature(st, 64)
    assert sig == S.tune_signature(S.row_stats(pl.copy(), 20000), 64)
    assert sig != S.tune_signature(st, 128) and sig != S.tune_signature(st, 64, "bf16")
    S.tune_cache_clear()
    assert S.tune_cache_get(sig) is None
    assert S.tune(st, 64).as_dict() == S.tune_model(st, 64).as_dict()
    forced = S.make_plan("rb_pr", vec=4, col_lanes=16, group=4, tile=0, comp_threshold=0)
    S.tune_cache_put(sig, forced)
    assert S.tune(st, 64).as_dict() == forced.as_dict()
    path = os.path.join(tmp_path, "plans.txt")
    S.tune_cache_save(path)
    S.tune_cache_clear()
    assert S.tune_cache_get(sig) is None
    S.tune_cache_load(path)
    assert S.tune_cache_get(sig).as_dict() == forced.as_dict()
    S.tune_cache_clear()


def test_batch_rejects_bad_items_without_gpu_work():
    A = S.csr_desc(10, 10, 5, 0x1000, 0x2000, None)
    it = S.batch_item(A, 0, 4, 0x3000, 4, 8)  # B null with nnz > 0, ldb < N
    with pytest.raises(S.SegtacoError, match="bad operands"):
        S.Batch([it])
