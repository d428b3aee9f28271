"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare (both linkages of the operator surface), and its host-side
logic (plan resolution, workspace sizing, CSR validation, generators) behaves.
No compute call is made (no GPU here)."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import segtaco_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "segtaco", "*.h")))


def _declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text,
                             flags=re.M):
            names.add(m.group(1))
    return names


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_loads():
    assert "gfx950" in S.version()


def test_every_declared_symbol_is_exported():
    declared = _declared_functions()
    assert len(declared) > 90
    missing = declared - _exported()
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_operator_surface_matches_reference_names():
    # The 79 live names of the reference's gpu_kernels.cuh:11-93 (the two
    # vendor-library rows run rocSPARSE).
    names = S.legacy_names()
    assert len(names) == 79 and len(set(names)) == 79
    assert "spmm_csr_cusparse_row" in names and "spmm_csr_gpu_cusparse" in names
    for must in ("spmm_rb_sr", "spmm_rb_pr", "spmm_eb_sr", "spmm_eb_pr", "spmm_csr_gpu_taco"):
        assert must in names
    assert set(names) <= _declared_functions()


def test_cxx_linkage_aliases_exported():
    exported = _exported()
    for name in S.legacy_names():
        mangled = f"_Z{len(name)}{name}P13taco_tensor_tS0_S0_"
        assert mangled in exported, mangled
    assert "_Z20get_gpu_timer_resultv" in exported


def _desc(M=100, K=80, nnz=1000, dtype=S.DT_F32):
    return S.csr_desc(M, K, nnz, 0x1000, 0x2000, 0x3000, dtype)


def test_plan_resolution_defaults():
    A = _desc()
    p = S.plan_resolve(A, 128, S.DT_F32, S.make_plan("rb_sr"))
    assert (p.vec, p.col_lanes, p.group) == (4, 32, 1)
    assert p.comp_threshold == 0  # FP32: no compensation by default
    p = S.plan_resolve(A, 32, S.DT_F32, S.make_plan("rb_pr"))
    assert p.vec == 4 and p.col_lanes * p.group <= 64
    p = S.plan_resolve(_desc(dtype=S.DT_BF16), 256, S.DT_BF16, S.make_plan("eb_sr"), 256, 256)
    assert p.vec == 8 and p.col_lanes == 32 and p.tile > 0 and p.comp_threshold == 1024
    p = S.plan_resolve(A, 128, S.DT_F32, S.make_plan("eb_pr", group=16))
    assert p.col_lanes * p.group == 64
    p = S.plan_resolve(A, 7, S.DT_F32, S.make_plan("rb_sr"))
    assert p.vec == 1 and p.col_lanes == 8
    p = S.plan_resolve(_desc(dtype=S.DT_F16), 6, S.DT_F16, S.make_plan("rb_sr"), 6, 6)
    assert p.vec == 2


@pytest.mark.parametrize("bad", [
    dict(schedule=7), dict(schedule=0, vec=3), dict(schedule=0, col_lanes=3),
    dict(schedule=1, col_lanes=16, group=8), dict(schedule=3, col_lanes=8, group=4),
    dict(schedule=2, tile=4), dict(schedule=0, vec=8)])
def test_plan_resolution_rejects(bad):
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(_desc(), 128, S.DT_F32, S.make_plan(**bad))


def test_vec_must_divide_n():
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(_desc(), 6, S.DT_F32, S.make_plan("rb_sr", vec=4))


def test_workspace_bytes():
    A = _desc(nnz=1000)
    p = S.plan_resolve(A, 64, S.DT_F32, S.make_plan("eb_sr", tile=100))
    ntiles = 10
    need = S.workspace_bytes(A, 64, p)
    assert need >= 4 * (ntiles + 1) + 2 * 4 * ntiles * 64
    assert S.workspace_bytes(A, 64, S.plan_resolve(A, 64, S.DT_F32, S.make_plan("rb_sr"))) == 0


def test_prepared_refuses_a_workspace_not_prepared_for_the_product():
    """segtaco_spmm_prepared on an eb_cs workspace that holds no preparation
    of this product is refused on the host, before any device work (the
    pointers here are never dereferenced)."""
    A = _desc(M=1000, K=1000, nnz=5000)
    p = S.plan_resolve(A, 64, S.DT_F32, S.make_plan("eb_cs"))
    need = S.workspace_bytes(A, 64, p)
    with pytest.raises(S.SegtacoError, match="not prepared"):
        S.spmm_prepared(A, 0x10000, S.DT_F32, 64, 64, 0x20000, 64, p, 0x40000, need)
    # a preparation without a workspace is refused too
    with pytest.raises(S.SegtacoError, match="workspace"):
        S.spmm_prepare(A, S.DT_F32, 64, 64, 64, p, None, 0)


def test_validate_csr():
    pos = np.array([0, 2, 2, 5], np.int32)
    crd = np.array([0, 3, 1, 2, 4], np.int32)
    S.validate_csr(3, 5, pos, crd)
    with pytest.raises(S.SegtacoError, match="outside"):
        S.validate_csr(3, 4, pos, crd)
    with pytest.raises(S.SegtacoError, match="decreases"):
        S.validate_csr(3, 5, np.array([0, 2, 1, 5], np.int32), crd)
    with pytest.raises(S.SegtacoError, match=r"pos\[0\]"):
        S.validate_csr(3, 5, np.array([1, 2, 2, 5], np.int32), crd)


def test_validate_csr_reports_the_first_offender_across_blocks():
    """The scans run in blocks of 4096; the message names the exact index."""
    M, K = 10000, 50
    pos = np.arange(M + 1, dtype=np.int32)  # one entry per row
    crd = (np.arange(M) % K).astype(np.int32)
    S.validate_csr(M, K, pos, crd)
    for bad_at, value in ((4096, K), (9999, -1), (5000, K + 7)):
        c = crd.copy()
        c[bad_at] = value
        with pytest.raises(S.SegtacoError, match=rf"crd\[{bad_at}\] = {value} outside"):
            S.validate_csr(M, K, pos, c)
    p = pos.copy()
    p[8193] = p[8192] - 1  # decreases at row 8192
    with pytest.raises(S.SegtacoError, match="pos decreases at row 8192"):
        S.validate_csr(M, K, p, crd)


def test_legacy_rejects_malformed_operands_before_gpu_work():
    pos = np.array([0, 2, 3], np.int32)
    crd = np.array([0, 9, 1], np.int32)  # 9 >= K
    B = np.ones((4, 3), np.float32)
    with pytest.raises(S.SegtacoError, match="outside"):
        S.legacy("spmm_rb_sr", pos, crd, np.ones(3, np.float32), B)
    with pytest.raises(S.SegtacoError, match="columns"):
        S.legacy("spmm_eb_sr", pos, np.array([0, 1, 2], np.int32), None, B, K=4,
                 C=np.zeros((2, 5), np.float32))


def test_generators_deterministic_and_shaped():
    a = S.gen_csr(S.GEN_POWERLAW, 9, 5000, 5000)
    b = S.gen_csr(S.GEN_POWERLAW, 9, 5000, 5000)
    c = S.gen_csr(S.GEN_POWERLAW, 10, 5000, 5000)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])
    pos, crd, vals = a
    S.validate_csr(5000, 5000, pos, crd)
    L = np.diff(pos)
    assert 12 < L.mean() < 20 and (L == 0).any()
    for i in range(0, 5000, 97):
        row = crd[pos[i]:pos[i + 1]]
        assert np.all(np.diff(row) > 0)
    assert vals.min() >= -1 and vals.max() <= 1
    pb, cb, _ = S.gen_csr(S.GEN_BANDED, 1, 500, 500, bw=32)
    assert int(pb[-1]) == 65 * 500 - 2 * sum(range(33))
    pu, _, _ = S.gen_csr(S.GEN_UNIFORM, 1, 64, 4096, mean=16)
    assert np.all(np.diff(pu) == 16)
    d1 = S.gen_dense(3, 100, 7)
    assert np.array_equal(d1, S.gen_dense(3, 100, 7)) and abs(d1.mean()) < 0.1


def test_eb_cs_plan_resolution():
    A = _desc(M=262144, K=262144, nnz=4204209)
    p = S.plan_resolve(A, 128, S.DT_F32, S.make_plan("eb_cs"))
    # one 64-tile unit per compute unit (256 on MI355X): 4204209 / 16384 -> 257,
    # 16359 tiles in 256 units
    assert (p.vec, p.col_lanes, p.group, p.tile, p.comp_threshold) == (4, 8, 64, 257, 0)
    assert S.workspace_bytes(A, 128, p) > S.workspace_bytes(A, 128, S.plan_resolve(
        A, 128, S.DT_F32, S.make_plan("eb_sr", tile=264)))
    # 16-bit storage: vector capped at 4; eb_cs never compensates (the
    # default resolves to 0, an explicit threshold is refused)
    B16 = _desc(M=262144, K=262144, nnz=4204209, dtype=S.DT_BF16)
    p = S.plan_resolve(B16, 256, S.DT_BF16, S.make_plan("eb_cs"), 256, 256)
    assert p.vec == 4 and p.col_lanes == 8 and p.comp_threshold == 0
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(B16, 256, S.DT_BF16, S.make_plan("eb_cs", comp_threshold=1024), 256, 256)
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(A, 128, S.DT_F32, S.make_plan("eb_cs", col_lanes=16))
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(A, 128, S.DT_F32, S.make_plan("eb_cs", tile=1024))
    with pytest.raises(S.SegtacoError):
        S.plan_resolve(A, 128, S.DT_F32, S.make_plan("eb_cs", group=128))
    # small matrices: the tile floor
    p = S.plan_resolve(_desc(M=100, K=80, nnz=1000), 64, S.DT_F32, S.make_plan("eb_cs"))
    assert p.tile == 32


def test_nnz_bound_leaves_room_for_32bit_tile_positions():
    """Entry positions are 32-bit and kernels form positions up to one tile
    (at most 2^22) plus a staging chunk past a tile's start: the engine takes
    nnz up to 2^31 - 1 - 2^23 and refuses more with UNSUPPORTED (4)."""
    limit = 0x7FFFFFFF - (1 << 23)
    for sched in ("rb_sr", "eb_sr", "eb_pr"):
        p = S.plan_resolve(_desc(M=1 << 26, K=1 << 20, nnz=limit), 32, S.DT_F32,
                           S.make_plan(sched, **({"tile": 1 << 22} if sched != "rb_sr" else {})))
        assert p.schedule == S.SCHEDULES[sched]
        with pytest.raises(S.SegtacoError) as e:
            S.plan_resolve(_desc(M=1 << 26, K=1 << 20, nnz=limit + 1), 32, S.DT_F32,
                           S.make_plan(sched))
        assert e.value.status == 4 and "nnz" in str(e.value)


def test_eb_sr_default_tile_follows_product_size():
    """eb_sr's default tile is 128 entries, halved (down to 8) while the
    product would fill fewer than two 4-wave workgroups per compute unit (256
    on MI355X, also the answer without a device): a lane group walks its tile
    serially, so a small product needs short tiles to spread over the device.
    Explicit tiles are kept."""
    def tile(nnz, N):
        return S.plan_resolve(_desc(M=max(1, nnz // 16), K=1000, nnz=nnz), N, S.DT_F32,
                              S.make_plan("eb_sr")).tile

    # N = 128 FP32: 32 lanes per entry, 2 lane groups per wave, 8 tiles per workgroup
    want = 2 * 256 * 8
    assert tile(4204209, 128) == 128           # the power-law headline keeps 128
    assert tile(128 * want, 128) == 128        # exactly enough tiles at 128
    assert tile(128 * (want - 1), 128) == 64  # one tile short at 128
    assert tile(15591, 128) == 8               # floor
    # N = 16 FP32: 4 lanes per entry, 16 groups per wave, 64 tiles per workgroup
    assert tile(100000, 16) == 8
    assert tile(64 * 2 * 256 * 64, 16) == 64
    # the smallest tile T with ceil(nnz / T) >= want, or 8
    for nnz in (1, 1000, 50000, 300000, 2000000):
        t = tile(nnz, 128)
        assert t in (8, 16, 32, 64, 128)
        assert t == 128 or -(-nnz // t) >= want or t == 8
        assert t == 8 or -(-nnz // (t // 2)) >= want or t == 128
    assert S.plan_resolve(_desc(nnz=15591), 128, S.DT_F32,
                          S.make_plan("eb_sr", tile=512)).tile == 512


def test_bd_plan_resolution():
    """bd: 16-bit storage only, 32-row blocks by default (64 on request),
    lanes and groups fixed at 1."""
    B16 = _desc(M=1000, K=1000, nnz=20000, dtype=S.DT_BF16)
    p = S.plan_resolve(B16, 256, S.DT_BF16, S.make_plan("bd"), 256, 256)
    assert (p.tile, p.col_lanes, p.group, p.comp_threshold) == (32, 1, 1, 1024)
    assert S.workspace_bytes(B16, 256, p) == 0
    assert S.plan_resolve(B16, 256, S.DT_BF16, S.make_plan("bd", tile=64), 256, 256).tile == 64
    with pytest.raises(S.SegtacoError):  # FP32 has no faster matrix-core path here
        S.plan_resolve(_desc(M=1000, K=1000, nnz=20000), 64, S.DT_F32, S.make_plan("bd"))
    for bad in (dict(tile=128), dict(tile=48), dict(col_lanes=4), dict(group=2)):
        with pytest.raises(S.SegtacoError):
            S.plan_resolve(B16, 256, S.DT_BF16, S.make_plan("bd", **bad), 256, 256)


def test_tune_cache_takes_batch_schedules_only():
    """The plan cache feeds batches and the server: standalone-only schedules
    are refused."""
    for name in ("eb_cs", "bd"):
        with pytest.raises(S.SegtacoError):
            S.tune_cache_put(0x1234, S.make_plan(name))
