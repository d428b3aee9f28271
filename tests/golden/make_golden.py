"""Writes tests/golden/*.npz from the reference's own CPU code.

Run from the repository root in a container where the reference tree is
present and oracle/_ref/libsegtaco_ref.so is built (`make -C oracle`):

    python tests/golden/make_golden.py

Every expected output comes from the reference code compiled in place by
oracle/Makefile: ds.h's MatrixMarket reader (read_mtx_file, ds.h:183-268) and
COO->CSR builder (to_eigen_csr, ds.h:380-397), and Eigen 3.4.0's sparse x
dense product (SparseDenseProduct.h:109-147). Inputs are seeded numpy draws.
The fixtures are data only (inputs and expected outputs).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

MTX = "/root/reference/matrix_subset/192bit/192bit.mtx"


def _uniform_csr(rng, M, K, per_row):
    pos = np.arange(M + 1, dtype=np.int32) * per_row
    crd = np.concatenate([np.sort(rng.choice(K, per_row, replace=False)) for _ in range(M)])
    return pos.astype(np.int32), crd.astype(np.int32)


def _powerlaw_csr(rng, M, K, mean, sigma, cap):
    mu = np.log(mean) - sigma * sigma / 2
    L = np.minimum(np.rint(np.exp(mu + sigma * rng.standard_normal(M))), cap).astype(np.int64)
    L = np.minimum(L, K)
    pos = np.concatenate([[0], np.cumsum(L)]).astype(np.int32)
    crd = np.concatenate([np.sort(rng.choice(K, int(l), replace=False)) for l in L] or
                         [np.zeros(0, np.int64)]).astype(np.int32)
    return pos, crd


def case_tiny(rng):
    # COO with duplicates and explicit zeros: to_eigen_csr sums duplicates and
    # keeps zeros; rows 0, 3, 6 empty (leading, middle, trailing).
    rows = np.array([1, 1, 2, 1, 4, 4, 5, 2, 5, 5, 1], dtype=np.int32)
    cols = np.array([0, 3, 4, 0, 1, 2, 4, 4, 0, 4, 2], dtype=np.int32)
    vals = np.array([.5, -1, .25, .75, 2, 0, -.5, 1, 3, .125, -2], dtype=np.float32)
    M, K, N = 7, 5, 3
    pos, crd, v = O.ref_coo_to_csr(M, K, rows, cols, vals)
    B = _grid(rng, (K, N))
    return dict(coo_rows=rows, coo_cols=cols, coo_vals=vals, M=M, K=K, pos=pos, crd=crd,
                vals=v, B=B)


def _grid(rng, shape):
    # U[-1, 1] on a 2^-12 grid: products carry up to 24 significant bits, so
    # FP32 sums still round, while the fixture stays compressible.
    return (rng.integers(-4096, 4097, shape) / 4096.0).astype(np.float32)


def case_matrix(pos, crd, K, N, rng, with_vals=True):
    nnz = int(pos[-1])
    vals = _grid(rng, nnz) if with_vals else None
    B = _grid(rng, (K, N))
    return dict(M=len(pos) - 1, K=K, pos=pos, crd=crd, vals=vals, B=B)


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libsegtaco_ref.so is not built")
    rng = np.random.default_rng(20240917)
    cases = {}
    cases["tiny_dups"] = case_tiny(rng)
    p, c = _uniform_csr(rng, 256, 256, 16)
    cases["uniform_256_n32"] = case_matrix(p, c, 256, 32, rng)
    p, c = _powerlaw_csr(rng, 1500, 1500, 16, 1.5, 8192)
    cases["powerlaw_1500_n7"] = case_matrix(p, c, 1500, 7, rng)
    p, c = _powerlaw_csr(rng, 300, 3000, 40, 1.5, 3000)
    cases["longrows_300_n4_novals"] = case_matrix(p, c, 3000, 4, rng, with_vals=False)
    if os.path.exists(MTX):
        nr, nc, pos, crd = O.ref_read_mtx(MTX)
        head = 1500
        hp = pos[: head + 1].copy()
        hc = crd[: int(hp[-1])].copy()
        cases["mtx192bit_head1500_n2"] = case_matrix(hp, hc, nc, 2, rng)
        cases["mtx192bit_head1500_n2"].update(
            mtx_nrow=nr, mtx_ncol=nc, mtx_nnz=int(pos[-1]),
            mtx_pos_sum=int(pos.astype(np.int64).sum()),
            mtx_crd_sum=int(crd.astype(np.int64).sum()))
    for name, d in cases.items():
        vals = d["vals"]
        B = d["B"]
        C1 = O.ref_eigen_spmm(d["pos"], d["crd"], vals, B, K=d["K"], threads=1)
        C4 = O.ref_eigen_spmm(d["pos"], d["crd"], vals, B, K=d["K"], threads=4)
        assert np.array_equal(C1, C4)
        out = {k: np.asarray(v) for k, v in d.items() if v is not None}
        out["has_vals"] = np.int32(vals is not None)
        out["C_eigen"] = C1
        for T in (32, 256):
            out[f"block_starts_{T}"] = O.block_starts(d["pos"], d["M"], T)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
