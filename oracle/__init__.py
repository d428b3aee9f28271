"""CPU oracle for the CSR SpMM path — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` (as the checker) and bench.py's
``cpu_baseline`` leg import this package. The product (``segtaco_amd``) never
does.

Two libraries back it:

* ``liboracle.so`` (oracle.c): a plain-C restatement of the reference's
  partition searches (gpu_library.h:138-152, 185-215, 236-240) and of the
  Eigen product order (eigen-3.4.0/Eigen/src/SparseCore/
  SparseDenseProduct.h:141-146), plus an FP64 closed form.
* ``_ref/libsegtaco_ref.so`` (ref_shim.cpp): the reference's own CPU code —
  ds.h's MatrixMarket reader and COO->CSR builder, and the vendored Eigen
  3.4.0 sparse x dense product — compiled from the reference tree in place.
  Optional: absent when it was not built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u16p = ctypes.POINTER(ctypes.c_uint16)


def _ptr(a, typ):
    if a is None:
        return None
    return a.ctypes.data_as(typ)


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} missing (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.or_search_before.restype = ctypes.c_int32
        L.or_search_before.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.or_block_starts.restype = None
        L.or_block_starts.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i32p]
        L.or_segment_number.restype = ctypes.c_int32
        L.or_segment_number.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32]
        L.or_row_of_nnz.restype = None
        L.or_row_of_nnz.argtypes = [_i32p, ctypes.c_int32, _i32p]
        for fn in ("or_tile_rows", "or_fix_t0"):
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       _i32p]
        L.or_spmm_f32_seq.restype = None
        L.or_spmm_f32_seq.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f32p,
                                      _f32p, ctypes.c_int64, _f32p, ctypes.c_int64]
        L.or_spmm_f64.restype = None
        L.or_spmm_f64.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f32p,
                                  _f32p, ctypes.c_int64, _f64p, ctypes.c_int64]
        L.or_spmm_absbound.restype = None
        L.or_spmm_absbound.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f32p,
                                       _f32p, ctypes.c_int64, _f64p, ctypes.c_int64]
        L.or_sddmm_f64.restype = None
        L.or_sddmm_f64.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f32p, _f32p,
                                   ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_int32, _f64p]
        L.or_max_scaled_err.restype = ctypes.c_double
        L.or_max_scaled_err.argtypes = [ctypes.c_int64, _f32p, _f64p, _f64p]
        L.or_round_array.restype = None
        L.or_round_array.argtypes = [ctypes.c_int32, ctypes.c_int64, _f32p, _u16p, _f32p]
        _LIB = L
    return _LIB


def ref_lib():
    """Load _ref/libsegtaco_ref.so, or return None when it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libsegtaco_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_eigen_spmm.restype = None
        R.ref_eigen_spmm.argtypes = [ctypes.c_int32] * 4 + [_i32p, _i32p, _f32p, _f32p, _f32p,
                                                          ctypes.c_int32]
        R.ref_eigen_spmm_timed.restype = None
        R.ref_eigen_spmm_timed.argtypes = [ctypes.c_int32] * 4 + [
            _i32p, _i32p, _f32p, _f32p, _f32p, ctypes.c_int32, ctypes.c_int32, _f64p]
        R.ref_coo_to_csr.restype = ctypes.c_int32
        R.ref_coo_to_csr.argtypes = [ctypes.c_int32] * 3 + [
            _i32p, _i32p, _f32p, ctypes.POINTER(_i32p), ctypes.POINTER(_i32p),
            ctypes.POINTER(_f32p)]
        R.ref_read_mtx.restype = ctypes.c_int32
        R.ref_read_mtx.argtypes = [ctypes.c_char_p, _i32p, _i32p, ctypes.POINTER(_i32p),
                                   ctypes.POINTER(_i32p)]
        for fn in ("ref_compare_array", "ref_compare_array_my"):
            getattr(R, fn).restype = ctypes.c_float
            getattr(R, fn).argtypes = [_f32p, _f32p, ctypes.c_size_t]
        R.ref_free.restype = None
        R.ref_free.argtypes = [ctypes.c_void_p]
        R.ref_max_threads.restype = ctypes.c_int32
        R.ref_max_threads.argtypes = []
        _REF = R
    return _REF


# ----------------------------------------------------------------------------
# Partition searches (integer; bit-exact parity targets)
# ----------------------------------------------------------------------------

def search_before(pos: np.ndarray, start: int, end: int, target: int) -> int:
    """taco_binarySearchBefore (gpu_library.h:185-205)."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    return int(lib().or_search_before(_ptr(pos, _i32p), start, end, target))


def block_starts(pos: np.ndarray, M: int, T: int, num_blocks: int | None = None) -> np.ndarray:
    """K21 partition: out[b] = search_before(pos, 0, M, b*T), b = 0..num_blocks
    (gpu_library.h:206-215, 236-240). num_blocks defaults to ceil(nnz/T)."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    nnz = int(pos[M])
    if num_blocks is None:
        num_blocks = (nnz + T - 1) // T
    out = np.empty(num_blocks + 1, dtype=np.int32)
    lib().or_block_starts(_ptr(pos, _i32p), M, T, num_blocks, _ptr(out, _i32p))
    return out


def segment_number(pos: np.ndarray, M: int, elem: int) -> int:
    """binary_search_segment_number (gpu_library.h:138-152)."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    return int(lib().or_segment_number(_ptr(pos, _i32p), M, elem))


def row_of_nnz(pos: np.ndarray, M: int) -> np.ndarray:
    """Row of every stored entry (the COO row array)."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    rows = np.empty(max(int(pos[M]), 1), dtype=np.int32)
    lib().or_row_of_nnz(_ptr(pos, _i32p), M, _ptr(rows, _i32p))
    return rows[: int(pos[M])]


def tile_rows(pos: np.ndarray, M: int, T: int) -> np.ndarray:
    """First row of each of the ceil(nnz/T) element-balanced tiles: the K21
    search plus the repair loop of spmm_eb_sr.h:39-42; entry ntiles is M."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    nt = (int(pos[M]) + T - 1) // T
    out = np.empty(nt + 1, dtype=np.int32)
    lib().or_tile_rows(_ptr(pos, _i32p), M, T, nt, _ptr(out, _i32p))
    return out


def fix_t0(pos: np.ndarray, M: int, T: int) -> np.ndarray:
    """Per tile: the tile where the row ending in it started, if that row
    started in an earlier tile, else -1."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    nt = (int(pos[M]) + T - 1) // T
    out = np.empty(max(nt, 1), dtype=np.int32)
    lib().or_fix_t0(_ptr(pos, _i32p), M, T, nt, _ptr(out, _i32p))
    return out[:nt]


# ----------------------------------------------------------------------------
# SpMM value oracles
# ----------------------------------------------------------------------------

def _csr_args(pos, crd, vals, B):
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    crd = np.ascontiguousarray(crd, dtype=np.int32)
    if crd.size == 0:
        crd = np.zeros(1, dtype=np.int32)
    if vals is not None:
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        if vals.size == 0:
            vals = np.zeros(1, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    return pos, crd, vals, B


def spmm_f32_seq(pos, crd, vals, B) -> np.ndarray:
    """Eigen-order FP32 product (SparseDenseProduct.h:141-146)."""
    M = len(pos) - 1
    pos, crd, vals, B = _csr_args(pos, crd, vals, B)
    N = B.shape[1]
    C = np.empty((M, N), dtype=np.float32)
    lib().or_spmm_f32_seq(M, N, _ptr(pos, _i32p), _ptr(crd, _i32p), _ptr(vals, _f32p),
                          _ptr(B, _f32p), N, _ptr(C, _f32p), N)
    return C


def spmm_f64(pos, crd, vals, B) -> np.ndarray:
    """Closed form in double."""
    M = len(pos) - 1
    pos, crd, vals, B = _csr_args(pos, crd, vals, B)
    N = B.shape[1]
    C = np.empty((M, N), dtype=np.float64)
    lib().or_spmm_f64(M, N, _ptr(pos, _i32p), _ptr(crd, _i32p), _ptr(vals, _f32p),
                      _ptr(B, _f32p), N, _ptr(C, _f64p), N)
    return C


def sddmm_f64(pos, crd, vals, X, Y, absval: bool = False) -> np.ndarray:
    """SDDMM in double: out[p] = s_p * dot(X[row(p)], Y[crd[p]]); with absval,
    the error scale |s_p| * sum |X||Y|."""
    M = len(pos) - 1
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    crd = np.ascontiguousarray(crd, dtype=np.int32)
    if crd.size == 0:
        crd = np.zeros(1, dtype=np.int32)
    vals = None if vals is None else np.ascontiguousarray(vals, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    J = X.shape[1]
    nnz = int(pos[-1]) if M > 0 else 0
    out = np.empty(max(nnz, 1), dtype=np.float64)
    lib().or_sddmm_f64(M, J, _ptr(pos, _i32p), _ptr(crd, _i32p), _ptr(vals, _f32p),
                       _ptr(X, _f32p), J, _ptr(Y, _f32p), Y.shape[1], int(absval),
                       _ptr(out, _f64p))
    return out[:nnz]


def spmm_absbound(pos, crd, vals, B) -> np.ndarray:
    M = len(pos) - 1
    pos, crd, vals, B = _csr_args(pos, crd, vals, B)
    N = B.shape[1]
    S = np.empty((M, N), dtype=np.float64)
    lib().or_spmm_absbound(M, N, _ptr(pos, _i32p), _ptr(crd, _i32p), _ptr(vals, _f32p),
                           _ptr(B, _f32p), N, _ptr(S, _f64p), N)
    return S


def scaled_err(C: np.ndarray, ref: np.ndarray, scale: np.ndarray) -> float:
    """max |C - ref| / (sum_p |v_p||B[crd_p,k]| + 1e-30): the error metric the
    parity tolerances are stated in."""
    C = np.ascontiguousarray(C, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    scale = np.ascontiguousarray(scale, dtype=np.float64)
    assert C.size == ref.size == scale.size
    if C.size == 0:
        return 0.0
    return float(lib().or_max_scaled_err(C.size, _ptr(C, _f32p), _ptr(ref, _f64p),
                                         _ptr(scale, _f64p)))


def rel_norm_errs(C: np.ndarray, ref: np.ndarray):
    """(||C - ref||_F / ||ref||_F, max|C - ref| / max|ref|): norm-wise error
    measures for storage-precision checks, where an element whose magnitude is
    below the storage format's subnormal range may legitimately lose all of
    its own digits."""
    C = np.asarray(C, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = C - ref
    fro = float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))
    mx = float(np.abs(d).max() / max(np.abs(ref).max(), 1e-300)) if ref.size else 0.0
    return fro, mx


def round_to(dtype: str, x: np.ndarray):
    """Round float32 -> ('bf16'|'fp16') with RNE; returns (uint16 bits, widened float32)."""
    code = {"bf16": 1, "fp16": 2}[dtype]
    x = np.ascontiguousarray(x, dtype=np.float32)
    bits = np.empty(x.shape, dtype=np.uint16)
    back = np.empty(x.shape, dtype=np.float32)
    lib().or_round_array(code, x.size, _ptr(x, _f32p), _ptr(bits, _u16p), _ptr(back, _f32p))
    return bits, back


# ----------------------------------------------------------------------------
# The reference's own CPU code (oracle/_ref)
# ----------------------------------------------------------------------------

def ref_eigen_spmm(pos, crd, vals, B, K: int | None = None, threads: int = 1) -> np.ndarray:
    """The reference's Eigen CPU SpMM (EigenCSR x EigenRowMajor)."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    M = len(pos) - 1
    if vals is None:
        vals = np.ones(int(pos[M]), dtype=np.float32)
    pos, crd, vals, B = _csr_args(pos, crd, vals, B)
    K = B.shape[0] if K is None else K
    N = B.shape[1]
    C = np.empty((M, N), dtype=np.float32)
    R.ref_eigen_spmm(M, K, N, int(pos[M]), _ptr(pos, _i32p), _ptr(crd, _i32p),
                     _ptr(vals, _f32p), _ptr(B, _f32p), _ptr(C, _f32p), threads)
    return C


def ref_eigen_spmm_timed(pos, crd, vals, B, threads: int, trials: int):
    """Per-trial seconds of the reference's Eigen SpMM (one warm-up first)."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    M = len(pos) - 1
    pos, crd, vals, B = _csr_args(pos, crd, vals, B)
    K, N = B.shape
    C = np.empty((M, N), dtype=np.float32)
    times = np.zeros(trials, dtype=np.float64)
    R.ref_eigen_spmm_timed(M, K, N, int(pos[M]), _ptr(pos, _i32p), _ptr(crd, _i32p),
                           _ptr(vals, _f32p), _ptr(B, _f32p), _ptr(C, _f32p), threads,
                           trials, _ptr(times, _f64p))
    return times, C


def ref_coo_to_csr(m: int, n: int, rows, cols, vals):
    """ds.h:380-397 to_eigen_csr: sort, sum duplicates, compress."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    p_pos, p_crd, p_val = _i32p(), _i32p(), _f32p()
    nnz = R.ref_coo_to_csr(m, n, len(rows), _ptr(rows, _i32p), _ptr(cols, _i32p),
                           _ptr(vals, _f32p), ctypes.byref(p_pos), ctypes.byref(p_crd),
                           ctypes.byref(p_val))
    pos = np.ctypeslib.as_array(p_pos, shape=(m + 1,)).copy()
    crd = np.ctypeslib.as_array(p_crd, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(p_val, shape=(max(nnz, 1),))[:nnz].copy()
    for p in (p_pos, p_crd, p_val):
        R.ref_free(ctypes.cast(p, ctypes.c_void_p))
    return pos, crd, val


def ref_read_mtx(path: str):
    """ds.h:183-268 read_mtx_file: returns (nrow, ncol, pos, crd)."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    nr, nc = ctypes.c_int32(), ctypes.c_int32()
    p_pos, p_crd = _i32p(), _i32p()
    nnz = R.ref_read_mtx(path.encode(), ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(p_pos),
                         ctypes.byref(p_crd))
    if nnz < 0:
        raise FileNotFoundError(path)
    pos = np.ctypeslib.as_array(p_pos, shape=(nr.value + 1,)).copy()
    crd = np.ctypeslib.as_array(p_crd, shape=(max(nnz, 1),))[:nnz].copy()
    R.ref_free(ctypes.cast(p_pos, ctypes.c_void_p))
    R.ref_free(ctypes.cast(p_crd, ctypes.c_void_p))
    return nr.value, nc.value, pos, crd


def ref_compare_array(x, y) -> float:
    """ds.h:463-477 compare_array, compiled from the reference."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    y = np.ascontiguousarray(y, dtype=np.float32).ravel()
    return float(R.ref_compare_array(_ptr(x, _f32p), _ptr(y, _f32p), x.size))


def ref_compare_array_my(x, y) -> float:
    """ds.h:478-492 compare_array_my, compiled from the reference (prints
    each mismatch to stdout)."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libsegtaco_ref.so not built")
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    y = np.ascontiguousarray(y, dtype=np.float32).ravel()
    return float(R.ref_compare_array_my(_ptr(x, _f32p), _ptr(y, _f32p), x.size))
