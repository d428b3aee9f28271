"""SDDMM: out[p] = s_p * dot(X[row(p)], Y[crd[p]]).

CPU: the FP64 oracle is pinned by an independent dense brute force
((X Y^T) masked by the pattern, numpy float64) and special cases; planted
mistakes fail the metric. GPU: segtaco_sddmm against the oracle for FP32 /
BF16 / FP16 storage, ragged and empty rows, long rows, J not a multiple of
the vector width, strided X/Y, run-to-run determinism, and the full-size
power-law pattern.
"""
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S

TOL = 1e-5


def _pattern(rng, lens, K):
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in lens]
                         or [np.zeros(0, int)]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    return pos, crd, vals


def _dense_ref(pos, crd, vals, X, Y):
    rows = np.repeat(np.arange(len(pos) - 1), np.diff(pos))
    full = X.astype(np.float64) @ Y.astype(np.float64).T
    v = np.ones(len(crd)) if vals is None else vals.astype(np.float64)
    return full[rows, crd] * v


def _err(out, ref, scale):
    return float(np.max(np.abs(out - ref) / (scale + 1e-30))) if len(ref) else 0.0


@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_dense_brute_force(seed):
    rng = np.random.default_rng(seed)
    M, K, J = 40, 35, [1, 7, 32, 65][seed]
    lens = rng.integers(0, 12, M)
    lens[3] = 0
    pos, crd, vals = _pattern(rng, lens, K)
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    got = O.sddmm_f64(pos, crd, vals, X, Y)
    assert np.allclose(got, _dense_ref(pos, crd, vals, X, Y), rtol=1e-12, atol=1e-12)
    assert np.allclose(O.sddmm_f64(pos, crd, None, X, Y), _dense_ref(pos, crd, None, X, Y),
                       rtol=1e-12, atol=1e-12)


def test_oracle_special_cases():
    rng = np.random.default_rng(9)
    n, J = 30, 16
    X = rng.uniform(-1, 1, (n, J)).astype(np.float32)
    diag_pos = np.arange(n + 1, dtype=np.int32)
    diag_crd = np.arange(n, dtype=np.int32)
    # diagonal pattern with Y = X: squared row norms
    sq = O.sddmm_f64(diag_pos, diag_crd, None, X, X)
    assert np.allclose(sq, (X.astype(np.float64) ** 2).sum(1), rtol=1e-12)
    # J = 1: s_p * x_i * y_k
    pos, crd, vals = _pattern(rng, rng.integers(0, 5, n), n)
    x = rng.uniform(-1, 1, (n, 1)).astype(np.float32)
    y = rng.uniform(-1, 1, (n, 1)).astype(np.float32)
    rows = np.repeat(np.arange(n), np.diff(pos))
    assert np.allclose(O.sddmm_f64(pos, crd, vals, x, y),
                       vals.astype(np.float64) * x[rows, 0] * y[crd, 0], rtol=1e-12)
    # the scale bounds the value
    scale = O.sddmm_f64(pos, crd, vals, x, y, absval=True)
    assert np.all(np.abs(O.sddmm_f64(pos, crd, vals, x, y)) <= scale + 1e-15)


def test_metric_catches_planted_mistakes():
    rng = np.random.default_rng(3)
    M, K, J = 50, 60, 24
    pos, crd, vals = _pattern(rng, rng.integers(1, 9, M), K)
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    ref = O.sddmm_f64(pos, crd, vals, X, Y)
    scale = O.sddmm_f64(pos, crd, vals, X, Y, absval=True)
    good = ref.astype(np.float32)
    assert _err(good, ref, scale) < 1e-6
    # transposed roles (Y's row from the output row, X's from the column)
    swapped = O.sddmm_f64(pos, crd, vals, Y[:M], X[:K] if K <= M else
                          np.vstack([X, X])[:K]).astype(np.float32)
    assert _err(swapped, ref, scale) > 1e-3
    # one dropped term of the dot product
    Xd = X.copy()
    Xd[:, 5] = 0
    assert _err(O.sddmm_f64(pos, crd, vals, Xd, Y).astype(np.float32), ref, scale) > 1e-3
    # shifted column index
    bad = crd.copy()
    bad[7] = (bad[7] + 1) % K
    assert _err(O.sddmm_f64(pos, bad, vals, X, Y).astype(np.float32), ref, scale) > 1e-3


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------

def _run(pos, crd, vals, X, Y, dtype="f32", ldx=None, ldy=None, reps=1):
    from gpu_util import DeviceCSR, dense_dev, stream_ptr, torch
    t = torch()
    M, J = X.shape
    K = Y.shape[0]
    ldx = ldx or J
    ldy = ldy or J
    Xp = np.zeros((M, ldx), np.float32)
    Xp[:, :J] = X
    Yp = np.zeros((K, ldy), np.float32)
    Yp[:, :J] = Y
    A = DeviceCSR(pos, crd, vals, K, dtype=dtype)
    Xd, Yd = dense_dev(Xp, dtype), dense_dev(Yp, dtype)
    nnz = int(pos[-1])
    ws_bytes = S.sddmm_workspace_bytes(A.desc)
    ws = t.empty(max(ws_bytes, 1), dtype=t.uint8, device="cuda")
    outs = []
    for _ in range(reps):
        out = t.full((max(nnz, 1),), float("nan"), dtype=t.float32, device="cuda")
        S.sddmm(A.desc, Xd.data_ptr(), ldx, Yd.data_ptr(), ldy, J, S.DTYPES[dtype],
                out.data_ptr(), ws.data_ptr(), ws_bytes, stream_ptr())
        t.cuda.synchronize()
        outs.append(out.cpu().numpy()[:nnz])
    return outs


def _case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "ragged_j128":
        lens = rng.integers(0, 30, 3000)
        lens[100:400] = 0
        return (*_pattern(rng, lens, 2500), 2500, 128)
    if name == "long_rows_j64":
        lens = np.array([0, 5, 9000, 0, 0, 3, 2048, 1])
        return (*_pattern(rng, lens, 12000), 12000, 64)
    if name == "j37_odd":
        return (*_pattern(rng, rng.integers(0, 20, 1500), 1400), 1400, 37)
    if name == "j1":
        return (*_pattern(rng, rng.integers(0, 50, 1000), 900), 900, 1)
    if name == "j300_multi_chunk":
        return (*_pattern(rng, rng.integers(0, 16, 800), 700), 700, 300)
    if name == "j8_narrow":  # 2 lanes x 4 (FP32) or 1 lane x 8 per entry
        return (*_pattern(rng, rng.integers(0, 40, 900), 800), 800, 8)
    if name == "all_empty":
        return (*_pattern(rng, np.zeros(64, int), 10), 10, 8)
    raise KeyError(name)


CASES = ["ragged_j128", "long_rows_j64", "j37_odd", "j1", "j8_narrow", "j300_multi_chunk",
         "all_empty"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp16"])
def test_sddmm_parity(name, dtype):
    pos, crd, vals, K, J = _case(name)
    rng = np.random.default_rng(1)
    M = len(pos) - 1
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    if dtype != "f32":  # compare against the oracle on the rounded inputs
        vals, X, Y = (O.round_to(dtype, a)[1] for a in (vals, X, Y))
    out = _run(pos, crd, vals, X, Y, dtype)[0]
    ref = O.sddmm_f64(pos, crd, vals, X, Y)
    scale = O.sddmm_f64(pos, crd, vals, X, Y, absval=True)
    err = _err(out, ref, scale)
    print(f"[sddmm] {name} {dtype}: err {err:.2e}", flush=True)
    assert err <= TOL


@pytest.mark.gpu
def test_sddmm_strided_novals_and_deterministic():
    pos, crd, _, K, J = _case("ragged_j128")
    rng = np.random.default_rng(2)
    M = len(pos) - 1
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    outs = _run(pos, crd, None, X, Y, "f32", ldx=J + 4, ldy=J + 12, reps=3)
    ref = O.sddmm_f64(pos, crd, None, X, Y)
    scale = O.sddmm_f64(pos, crd, None, X, Y, absval=True)
    assert _err(outs[0], ref, scale) <= TOL
    assert all(np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)) for o in outs[1:])


@pytest.mark.gpu
def test_sddmm_full_size_powerlaw():
    """The headline pattern (262144^2 power-law, 4.2M entries) with J = 128."""
    pos, crd, vals = S.gen_csr(S.GEN_POWERLAW, 2, 262144, 262144)
    X = S.gen_dense(21, 262144, 128)
    Y = S.gen_dense(22, 262144, 128)
    out = _run(pos, crd, vals, X, Y)[0]
    ref = O.sddmm_f64(pos, crd, vals, X, Y)
    scale = O.sddmm_f64(pos, crd, vals, X, Y, absval=True)
    assert _err(out, ref, scale) <= TOL


@pytest.mark.gpu
def test_sddmm_x_beyond_32bit_offsets():
    """M > 2^24 rows of X: loads take the 64-bit offset path; entries sit in
    the first and the last rows."""
    rng = np.random.default_rng(8)
    M, K, J = (1 << 24) + 300, 500, 4
    rows = np.sort(np.concatenate([rng.choice(M, 2000, replace=False), [0, M - 1]]))
    rows = np.unique(rows)
    lens = np.zeros(M, np.int64)
    lens[rows] = rng.integers(1, 6, len(rows))
    pos = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(lens[r]), replace=False)) for r in rows]
                         ).astype(np.int32)
    vals = rng.uniform(-1, 1, len(crd)).astype(np.float32)
    X = rng.uniform(-1, 1, (M, J)).astype(np.float32)
    Y = rng.uniform(-1, 1, (K, J)).astype(np.float32)
    print(f"[sddmm] M={M} J={J} nnz={len(crd)}", flush=True)
    out = _run(pos, crd, vals, X, Y)[0]
    ref = O.sddmm_f64(pos, crd, vals, X, Y)
    scale = O.sddmm_f64(pos, crd, vals, X, Y, absval=True)
    assert _err(out, ref, scale) <= TOL
