"""segtaco_amd — Python binding of libsegtaco, the gfx950 CSR SpMM engine.

Everything here is a thin ctypes layer over the C-ABI declared in
``include/segtaco/*.h``; all compute runs in the HIP library. There is no
CPU fallback: importing works without a GPU, but every compute call goes
through libsegtaco.so and raises if it is missing.

Two layers, mirroring the C-ABI:

* :func:`legacy` runs an operator-surface entry point (``spmm_rb_sr``,
  ``spmm_rb_pr``, ``spmm_eb_sr``, ``spmm_eb_pr``, ``spmm_csr_gpu_taco`` and the
  other names of the reference's gpu_kernels.cuh) on host numpy operands, with
  the reference's copy-in / compute / copy-out contract.
* :func:`spmm` runs the device-resident API on device pointers (integers) with
  an explicit :class:`Plan`, workspace and stream.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SEGTACO_LIB points at another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("SEGTACO_LIB") or os.path.join(_HERE, "libsegtaco.so")

# Schedules / dtypes / status codes (include/segtaco/spmm.h)
RB_SR, RB_PR, EB_SR, EB_PR, EB_CS, BD = 0, 1, 2, 3, 4, 5
SCHEDULES = {"rb_sr": RB_SR, "rb_pr": RB_PR, "eb_sr": EB_SR, "eb_pr": EB_PR, "eb_cs": EB_CS,
             "bd": BD}
DT_F32, DT_BF16, DT_F16 = 0, 1, 2
DTYPES = {"f32": DT_F32, "bf16": DT_BF16, "fp16": DT_F16}
OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_CAPACITY, ERR_STATE = 1, 2, 3, 4, 5, 6
ERRMODE_EXIT, ERRMODE_RECORD = 0, 1
GEN_UNIFORM, GEN_POWERLAW, GEN_BANDED = 0, 1, 2

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_vp = ctypes.c_void_p


class SegtacoError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"segtaco error {status}: {msg}")
        self.status = status


class Plan(ctypes.Structure):
    """segtaco_plan_t: schedule + (vec, col_lanes, group, tile, comp_threshold).
    Zero fields are resolved from the operands."""
    _fields_ = [("schedule", ctypes.c_int32), ("vec", ctypes.c_int32),
                ("col_lanes", ctypes.c_int32), ("group", ctypes.c_int32),
                ("tile", ctypes.c_int32), ("comp_threshold", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}

    def __repr__(self):
        return f"Plan({self.as_dict()})"


def make_plan(schedule, vec=0, col_lanes=0, group=0, tile=0, comp_threshold=-1) -> Plan:
    if isinstance(schedule, str):
        schedule = SCHEDULES[schedule]
    return Plan(schedule, vec, col_lanes, group, tile, comp_threshold)


class CSR(ctypes.Structure):
    """segtaco_csr_t: device CSR descriptor (pointers are device addresses)."""
    _fields_ = [("M", ctypes.c_int32), ("K", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("vals_dtype", ctypes.c_int32), ("pos", _vp), ("crd", _vp), ("vals", _vp)]


HIST_BINS = 24


class RowStats(ctypes.Structure):
    """segtaco_row_stats_t: row-length statistics the tuner reads."""
    _fields_ = [("M", ctypes.c_int32), ("K", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("max_len", ctypes.c_int32), ("empty_rows", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("mean_len", ctypes.c_double),
                ("cv", ctypes.c_double), ("frac_nnz_long", ctypes.c_double),
                ("hist", ctypes.c_int64 * HIST_BINS)]


class BatchItem(ctypes.Structure):
    """segtaco_batch_item_t."""
    _fields_ = [("A", CSR), ("B", _vp), ("ldb", ctypes.c_int64), ("C", _vp),
                ("ldc", ctypes.c_int64), ("N", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("plan", Plan), ("pos_host", _i32p)]


class ServeConfig(ctypes.Structure):
    """segtaco_serve_config_t."""
    _fields_ = [("cache_bytes", ctypes.c_size_t), ("inflight_bytes", ctypes.c_size_t),
                ("max_batch", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("dtype", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    """segtaco_serve_stats_t."""
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("rejected", ctypes.c_int64), ("ticks", ctypes.c_int64),
                ("cache_hits", ctypes.c_int64), ("cache_misses", ctypes.c_int64),
                ("evictions", ctypes.c_int64), ("uploaded_bytes", ctypes.c_int64),
                ("mean_batch", ctypes.c_double), ("p50_us", ctypes.c_double),
                ("p99_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("gflops", ctypes.c_double),
                ("wall_s", ctypes.c_double), ("max_concurrent_ticks", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: (getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}


class TacoTensor(ctypes.Structure):
    """taco_tensor_t (taco_tensor_t.h:6-15), identical layout."""
    _fields_ = [("order", ctypes.c_int32), ("dimensions", _i32p), ("csize", ctypes.c_int32),
                ("mode_ordering", _i32p), ("mode_types", _vp),
                ("indices", ctypes.POINTER(ctypes.POINTER(_i32p))), ("vals", _f32p),
                ("vals_size", ctypes.c_int32)]


class MtxInfo(ctypes.Structure):
    """segtaco_mtx_info_t."""
    _fields_ = [("coordinate", ctypes.c_int32), ("field", ctypes.c_int32),
                ("symmetry", ctypes.c_int32), ("entries", ctypes.c_int32)]


MTX_REFERENCE_TRUNCATE = 1

_LIB = None


def lib() -> ctypes.CDLL:
    """Load libsegtaco.so (raises if it was not built).

    Mixing with PyTorch in one process: torch bundles its own HIP/HSA runtime
    and libsegtaco binds to whichever libamdhip64.so.7 is loaded first, so
    import torch before the first call into this module; loading libsegtaco
    first and torch afterwards leaves two HSA runtimes in the process."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libsegtaco.so not built ({LIB_PATH}); run `python -m segtaco_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.segtaco_set_error_mode.restype = ctypes.c_int32
    L.segtaco_set_error_mode.argtypes = [ctypes.c_int32]
    L.segtaco_last_error.restype = ctypes.c_int32
    L.segtaco_last_error_string.restype = ctypes.c_char_p
    L.segtaco_clear_error.restype = None
    L.segtaco_version.restype = ctypes.c_char_p
    L.segtaco_device_count.restype = ctypes.c_int32
    L.get_gpu_timer_result.restype = ctypes.c_float
    L.segtaco_legacy_call.restype = None
    L.segtaco_legacy_call.argtypes = [ctypes.c_int32, ctypes.POINTER(TacoTensor),
                                      ctypes.POINTER(TacoTensor), ctypes.POINTER(TacoTensor)]
    L.segtaco_legacy_count.restype = ctypes.c_int32
    L.segtaco_legacy_name.restype = ctypes.c_char_p
    L.segtaco_legacy_name.argtypes = [ctypes.c_int32]
    L.segtaco_plan_resolve.restype = ctypes.c_int32
    L.segtaco_plan_resolve.argtypes = [ctypes.POINTER(CSR), ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Plan)]
    L.segtaco_spmm_workspace_bytes.restype = ctypes.c_size_t
    L.segtaco_spmm_workspace_bytes.argtypes = [ctypes.POINTER(CSR), ctypes.c_int32,
                                               ctypes.POINTER(Plan)]
    L.segtaco_spmm.restype = ctypes.c_int32
    L.segtaco_spmm.argtypes = [ctypes.POINTER(CSR), _vp, ctypes.c_int32, ctypes.c_int64,
                               ctypes.c_int32, _vp, ctypes.c_int64, ctypes.POINTER(Plan), _vp,
                               ctypes.c_size_t, _vp]
    L.segtaco_spmm_prepare.restype = ctypes.c_int32
    L.segtaco_spmm_prepare.argtypes = [ctypes.POINTER(CSR), ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Plan), _vp,
                                       ctypes.c_size_t, _vp]
    L.segtaco_spmm_prepared.restype = ctypes.c_int32
    L.segtaco_spmm_prepared.argtypes = L.segtaco_spmm.argtypes
    L.segtaco_spmm_noalloc.restype = ctypes.c_int32
    L.segtaco_spmm_noalloc.argtypes = [ctypes.POINTER(CSR), _vp, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_int32, _vp, ctypes.c_int64,
                                       ctypes.POINTER(Plan), _vp]
    L.segtaco_spmm_stage_times.restype = ctypes.c_int32
    L.segtaco_spmm_stage_times.argtypes = [ctypes.POINTER(CSR), _vp, ctypes.c_int32,
                                           ctypes.c_int64, ctypes.c_int32, _vp, ctypes.c_int64,
                                           ctypes.POINTER(Plan), _vp, ctypes.c_size_t, _vp,
                                           ctypes.c_int32, _f32p]
    L.segtaco_block_starts.restype = ctypes.c_int32
    L.segtaco_block_starts.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       _vp, _vp]
    L.segtaco_tile_rows.restype = ctypes.c_int32
    L.segtaco_tile_rows.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, _vp, _vp, _vp]
    L.segtaco_spmm_workspace_layout.restype = ctypes.c_int32
    L.segtaco_spmm_workspace_layout.argtypes = [ctypes.POINTER(CSR), ctypes.c_int32,
                                                ctypes.POINTER(Plan),
                                                ctypes.POINTER(ctypes.c_int64)]
    L.segtaco_cs_workspace_layout.restype = ctypes.c_int32
    L.segtaco_cs_workspace_layout.argtypes = [ctypes.POINTER(CSR), ctypes.c_int32,
                                              ctypes.POINTER(Plan),
                                              ctypes.POINTER(ctypes.c_int64)]
    L.segtaco_entry_rows.restype = ctypes.c_int32
    L.segtaco_entry_rows.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                                     _vp, _vp]
    L.segtaco_batch_partition.restype = ctypes.c_int32
    L.segtaco_batch_partition.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
    L.segtaco_csr_validate_host.restype = ctypes.c_int32
    L.segtaco_csr_validate_host.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            _i32p, _i32p]
    L.segtaco_device_malloc.restype = ctypes.c_int32
    L.segtaco_device_malloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    L.segtaco_device_free.restype = ctypes.c_int32
    L.segtaco_device_free.argtypes = [_vp]
    L.segtaco_memcpy.restype = ctypes.c_int32
    L.segtaco_memcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int32]
    L.segtaco_device_synchronize.restype = ctypes.c_int32
    L.segtaco_stream_synchronize.restype = ctypes.c_int32
    L.segtaco_stream_synchronize.argtypes = [_vp]
    L.segtaco_gen_pos.restype = ctypes.c_int64
    L.segtaco_gen_pos.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int32, ctypes.c_int32, _i32p]
    L.segtaco_gen_fill.restype = ctypes.c_int32
    L.segtaco_gen_fill.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f32p]
    L.segtaco_gen_dense.restype = ctypes.c_int32
    L.segtaco_gen_dense.argtypes = [ctypes.c_uint64, ctypes.c_int64, _f32p]
    L.segtaco_convert_f32.restype = ctypes.c_int32
    L.segtaco_convert_f32.argtypes = [ctypes.c_int32, ctypes.c_int64, _f32p, _u16p]
    L.segtaco_row_stats.restype = ctypes.c_int32
    L.segtaco_row_stats.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(RowStats)]
    L.segtaco_tune_model.restype = ctypes.c_int32
    L.segtaco_tune_model.argtypes = [ctypes.POINTER(RowStats), ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(Plan)]
    L.segtaco_tune.restype = ctypes.c_int32
    L.segtaco_tune.argtypes = [ctypes.POINTER(RowStats), ctypes.c_int32, ctypes.c_int32,
                               ctypes.POINTER(Plan)]
    L.segtaco_tune_measure.restype = ctypes.c_int32
    L.segtaco_tune_measure.argtypes = [ctypes.POINTER(CSR), ctypes.POINTER(RowStats), _vp,
                                       ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _vp,
                                       ctypes.c_int64, ctypes.c_int32, _vp, ctypes.POINTER(Plan),
                                       ctypes.POINTER(ctypes.c_float)]
    L.segtaco_tune_signature.restype = ctypes.c_uint64
    L.segtaco_tune_signature.argtypes = [ctypes.POINTER(RowStats), ctypes.c_int32,
                                         ctypes.c_int32]
    L.segtaco_tune_cache_get.restype = ctypes.c_int32
    L.segtaco_tune_cache_get.argtypes = [ctypes.c_uint64, ctypes.POINTER(Plan)]
    L.segtaco_tune_cache_put.restype = ctypes.c_int32
    L.segtaco_tune_cache_put.argtypes = [ctypes.c_uint64, ctypes.POINTER(Plan)]
    L.segtaco_tune_cache_clear.restype = ctypes.c_int32
    L.segtaco_tune_cache_save.restype = ctypes.c_int32
    L.segtaco_tune_cache_save.argtypes = [ctypes.c_char_p]
    L.segtaco_tune_cache_load.restype = ctypes.c_int32
    L.segtaco_tune_cache_load.argtypes = [ctypes.c_char_p]
    L.segtaco_batch_create.restype = ctypes.c_int32
    L.segtaco_batch_create.argtypes = [ctypes.POINTER(BatchItem), ctypes.c_int32,
                                       ctypes.c_int32, _vp, ctypes.POINTER(_vp)]
    for fn in ("segtaco_batch_run", "segtaco_batch_replay"):
        getattr(L, fn).restype = ctypes.c_int32
        getattr(L, fn).argtypes = [_vp, _vp]
    L.segtaco_batch_capture.restype = ctypes.c_int32
    L.segtaco_batch_capture.argtypes = [_vp]
    L.segtaco_batch_plan.restype = ctypes.c_int32
    L.segtaco_batch_plan.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(Plan)]
    L.segtaco_batch_units.restype = ctypes.c_int64
    L.segtaco_batch_units.argtypes = [_vp]
    L.segtaco_batch_workspace_bytes.restype = ctypes.c_size_t
    L.segtaco_batch_workspace_bytes.argtypes = [_vp]
    L.segtaco_batch_destroy.restype = ctypes.c_int32
    L.segtaco_batch_destroy.argtypes = [_vp]
    L.segtaco_server_create.restype = ctypes.c_int32
    L.segtaco_server_create.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(_vp)]
    L.segtaco_server_add_matrix.restype = ctypes.c_int32
    L.segtaco_server_add_matrix.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            _i32p, _i32p, _vp, ctypes.POINTER(ctypes.c_int32)]
    L.segtaco_server_submit.restype = ctypes.c_int32
    L.segtaco_server_submit.argtypes = [_vp, ctypes.c_int32, _vp, ctypes.c_int64, ctypes.c_int32,
                                        _vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.segtaco_server_wait.restype = ctypes.c_int32
    L.segtaco_server_wait.argtypes = [_vp, ctypes.c_int64, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_int32)]
    L.segtaco_server_drain.restype = ctypes.c_int32
    L.segtaco_server_drain.argtypes = [_vp, ctypes.c_double]
    L.segtaco_server_replay.restype = ctypes.c_int32
    L.segtaco_server_replay.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_vp),
                                        ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_vp),
                                        ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_double]
    L.segtaco_server_stats.restype = ctypes.c_int32
    L.segtaco_server_stats.argtypes = [_vp, ctypes.POINTER(ServeStats)]
    L.segtaco_server_reset_stats.restype = ctypes.c_int32
    L.segtaco_server_reset_stats.argtypes = [_vp]
    L.segtaco_server_plan.restype = ctypes.c_int32
    L.segtaco_server_plan.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Plan)]
    L.segtaco_server_destroy.restype = ctypes.c_int32
    L.segtaco_server_destroy.argtypes = [_vp]
    L.segtaco_server_debug_fail_tick.restype = ctypes.c_int32
    L.segtaco_server_debug_fail_tick.argtypes = [_vp, ctypes.c_int64]
    _i32pp = ctypes.POINTER(_i32p)
    L.segtaco_mtx_read.restype = ctypes.c_int32
    L.segtaco_mtx_read.argtypes = [ctypes.c_char_p, ctypes.c_int32, _i32p, _i32p, _i32p, _i32pp,
                                   _i32pp, ctypes.POINTER(MtxInfo)]
    L.segtaco_mtx_read_csr.restype = ctypes.c_int32
    L.segtaco_mtx_read_csr.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64, _i32p,
                                       _i32p, _i32p, _i32pp, _i32pp, ctypes.POINTER(_f32p)]
    L.segtaco_fill_reference_vals.restype = None
    L.segtaco_fill_reference_vals.argtypes = [ctypes.c_uint64, ctypes.c_int64, _f32p]
    L.segtaco_coo_to_csr.restype = ctypes.c_int32
    L.segtaco_coo_to_csr.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i32p, _i32p,
                                     _f32p, _i32p, _i32p, _f32p, _i32p]
    L.segtaco_host_free.restype = None
    L.segtaco_host_free.argtypes = [_vp]
    for fn in ("segtaco_compare_array", "segtaco_compare_array_my"):
        getattr(L, fn).restype = ctypes.c_float
        getattr(L, fn).argtypes = [_f32p, _f32p, ctypes.c_size_t]
    L.segtaco_compare_matrices.restype = ctypes.c_float
    L.segtaco_compare_matrices.argtypes = [ctypes.POINTER(TacoTensor), ctypes.POINTER(TacoTensor)]
    L.segtaco_sddmm_workspace_bytes.restype = ctypes.c_size_t
    L.segtaco_sddmm_workspace_bytes.argtypes = [ctypes.POINTER(CSR)]
    L.segtaco_sddmm.restype = ctypes.c_int32
    L.segtaco_sddmm.argtypes = [ctypes.POINTER(CSR), _vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_int32, _vp, _vp, ctypes.c_size_t, _vp]
    L.segtaco_vendor_spmv.restype = ctypes.c_int32
    L.segtaco_vendor_spmv.argtypes = [ctypes.POINTER(CSR), _vp, _vp, ctypes.c_int32, _vp,
                                      ctypes.POINTER(ctypes.c_float)]
    L.segtaco_vendor_spmm.restype = ctypes.c_int32
    L.segtaco_vendor_spmm.argtypes = [ctypes.POINTER(CSR), _vp, ctypes.c_int64, ctypes.c_int32,
                                      _vp, ctypes.c_int64, ctypes.c_int32, _vp, _f32p]
    _LIB = L
    return L


def check(status: int) -> None:
    if status != OK:
        L = lib()
        msg = (L.segtaco_last_error_string() or b"").decode()
        raise SegtacoError(status, msg)


def version() -> str:
    return lib().segtaco_version().decode()


def device_count() -> int:
    """GPUs visible to libsegtaco's HIP runtime (negative HIP error code on failure)."""
    return int(lib().segtaco_device_count())


# ---------------------------------------------------------------------------
# Synthetic operands (host, numpy)
# ---------------------------------------------------------------------------

def gen_csr(kind: int, seed: int, M: int, K: int, mean: float = 16.0, sigma: float = 1.5,
            cap: int = 8192, bw: int = 32, values: bool = True):
    """Seeded synthetic CSR: (pos int32[M+1], crd int32[nnz], vals float32[nnz])."""
    L = lib()
    pos = np.empty(M + 1, dtype=np.int32)
    nnz = L.segtaco_gen_pos(kind, seed, M, K, mean, sigma, cap, bw,
                            pos.ctypes.data_as(_i32p))
    if nnz < 0:
        raise SegtacoError(1, "segtaco_gen_pos failed (bad arguments or nnz > INT32_MAX)")
    crd = np.empty(max(nnz, 1), dtype=np.int32)
    vals = np.empty(max(nnz, 1), dtype=np.float32) if values else None
    check(L.segtaco_gen_fill(kind, seed, M, K, bw, pos.ctypes.data_as(_i32p),
                             crd.ctypes.data_as(_i32p),
                             vals.ctypes.data_as(_f32p) if values else None))
    return pos, crd[:nnz], (vals[:nnz] if values else None)


def gen_dense(seed: int, rows: int, cols: int) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float32)
    check(lib().segtaco_gen_dense(seed, rows * cols, out.ctypes.data_as(_f32p)))
    return out


def to_half(dtype: str | int, x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16/FP16 bit patterns (uint16), round to nearest even."""
    code = DTYPES[dtype] if isinstance(dtype, str) else dtype
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    check(lib().segtaco_convert_f32(code, x.size, x.ctypes.data_as(_f32p),
                                    out.ctypes.data_as(_u16p)))
    return out


def validate_csr(M: int, K: int, pos: np.ndarray, crd: np.ndarray) -> None:
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    crd = np.ascontiguousarray(crd, dtype=np.int32)
    nnz = int(pos[M]) if M > 0 else 0
    crd_p = crd.ctypes.data_as(_i32p) if crd.size else None
    check(lib().segtaco_csr_validate_host(M, K, nnz, pos.ctypes.data_as(_i32p), crd_p))


# ---------------------------------------------------------------------------
# Operator surface (host operands)
# ---------------------------------------------------------------------------

def legacy_names() -> list[str]:
    L = lib()
    return [L.segtaco_legacy_name(i).decode() for i in range(L.segtaco_legacy_count())]


def _legacy_id(name: str) -> int:
    names = legacy_names()
    if name not in names:
        raise KeyError(f"unknown operator {name}")
    return names.index(name)


@dataclass
class _Keep:
    """Keeps the ctypes buffers of a taco_tensor_t alive."""
    objs: list


def _taco_dense(arr: np.ndarray, keep: list) -> TacoTensor:
    dims = (ctypes.c_int32 * 2)(arr.shape[0], arr.shape[1])
    order = (ctypes.c_int32 * 2)(0, 1)
    keep += [dims, order, arr]
    t = TacoTensor()
    t.order = 2
    t.dimensions = ctypes.cast(dims, _i32p)
    t.mode_ordering = ctypes.cast(order, _i32p)
    t.vals = arr.ctypes.data_as(_f32p)
    t.vals_size = arr.size
    return t


def _taco_csr(M: int, K: int, pos, crd, vals, keep: list) -> TacoTensor:
    dims = (ctypes.c_int32 * 2)(M, K)
    idx1 = (_i32p * 2)(pos.ctypes.data_as(_i32p), crd.ctypes.data_as(_i32p))
    idx = (ctypes.POINTER(_i32p) * 2)(None, ctypes.cast(idx1, ctypes.POINTER(_i32p)))
    keep += [dims, idx1, idx, pos, crd, vals]
    t = TacoTensor()
    t.order = 2
    t.dimensions = ctypes.cast(dims, _i32p)
    t.indices = ctypes.cast(idx, ctypes.POINTER(ctypes.POINTER(_i32p)))
    t.vals = vals.ctypes.data_as(_f32p) if vals is not None else None
    t.vals_size = int(pos[M]) if M > 0 else 0
    return t


def legacy(name: str, pos, crd, vals, B: np.ndarray, K: int | None = None,
           C: np.ndarray | None = None):
    """Run operator-surface entry point `name` on host operands.

    Returns (C float32 [M, N], elapsed ms from get_gpu_timer_result()).
    Errors raise SegtacoError (the library runs in record mode here)."""
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    crd = np.ascontiguousarray(crd, dtype=np.int32)
    if crd.size == 0:
        crd = np.zeros(1, dtype=np.int32)
    if vals is not None:
        vals = np.ascontiguousarray(vals, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M = len(pos) - 1
    K = B.shape[0] if K is None else K
    N = B.shape[1]
    if C is None:
        C = np.full((M, N), np.nan, dtype=np.float32)
    keep: list = []
    tA = _taco_csr(M, K, pos, crd, vals, keep)
    tB = _taco_dense(B, keep)
    tC = _taco_dense(C, keep)
    prev = L.segtaco_set_error_mode(ERRMODE_RECORD)
    try:
        L.segtaco_clear_error()
        L.segtaco_legacy_call(_legacy_id(name), ctypes.byref(tC), ctypes.byref(tA),
                              ctypes.byref(tB))
        check(L.segtaco_last_error())
    finally:
        L.segtaco_set_error_mode(prev)
    return C, float(L.get_gpu_timer_result())


# ---------------------------------------------------------------------------
# Operand construction and comparators (include/segtaco/io.h)
# ---------------------------------------------------------------------------

def _take(ptr, n, dtype):
    """Copy a library-allocated array into numpy and free it."""
    L = lib()
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dtype, copy=True) \
        if n > 0 else np.zeros(0, dtype=dtype)
    L.segtaco_host_free(ctypes.cast(ptr, _vp))
    return out


def mtx_read(path: str, flags: int = 0):
    """MatrixMarket pattern -> (nrow, ncol, rows, cols, info), sorted 0-based
    coordinates (read_mtx_file, ds.h:183-268)."""
    L = lib()
    m, n, k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    R, C = _i32p(), _i32p()
    info = MtxInfo()
    check(L.segtaco_mtx_read(path.encode(), flags, ctypes.byref(m), ctypes.byref(n),
                             ctypes.byref(k), ctypes.byref(R), ctypes.byref(C),
                             ctypes.byref(info)))
    rows = _take(R, k.value, np.int32)
    cols = _take(C, k.value, np.int32)
    return m.value, n.value, rows, cols, info


def mtx_read_csr(path: str, flags: int = 0, seed: int = 0):
    """new_get_matrix + to_eigen_csr: (M, K, pos, crd, vals). seed = 2**64 - 1
    fills ones instead of the {0, 0.1, 0.2} reference values."""
    L = lib()
    M, K, nnz = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    P, D, V = _i32p(), _i32p(), _f32p()
    check(L.segtaco_mtx_read_csr(path.encode(), flags, ctypes.c_uint64(seed), ctypes.byref(M),
                                 ctypes.byref(K), ctypes.byref(nnz), ctypes.byref(P),
                                 ctypes.byref(D), ctypes.byref(V)))
    pos = _take(P, M.value + 1, np.int32)
    return M.value, K.value, pos, _take(D, nnz.value, np.int32), _take(V, nnz.value, np.float32)


def fill_reference_vals(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().segtaco_fill_reference_vals(ctypes.c_uint64(seed), n, out.ctypes.data_as(_f32p))
    return out


def coo_to_csr(m: int, n: int, rows, cols, vals=None):
    """COO -> compressed CSR (to_eigen_csr, ds.h:380-397): (pos, crd, vals)."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    k = len(rows)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.float32)
    pos = np.empty(m + 1, dtype=np.int32)
    crd = np.empty(max(k, 1), dtype=np.int32)
    ov = np.empty(max(k, 1), dtype=np.float32)
    out = ctypes.c_int32()
    check(lib().segtaco_coo_to_csr(m, n, k, rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p),
                                   v.ctypes.data_as(_f32p) if v is not None else None,
                                   pos.ctypes.data_as(_i32p), crd.ctypes.data_as(_i32p),
                                   ov.ctypes.data_as(_f32p), ctypes.byref(out)))
    c = out.value
    return pos, crd[:c].copy(), (ov[:c].copy() if v is not None else None)


def compare_array(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    y = np.ascontiguousarray(y, dtype=np.float32).ravel()
    return float(lib().segtaco_compare_array(x.ctypes.data_as(_f32p), y.ctypes.data_as(_f32p),
                                             x.size))


def compare_array_my(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    y = np.ascontiguousarray(y, dtype=np.float32).ravel()
    return float(lib().segtaco_compare_array_my(x.ctypes.data_as(_f32p),
                                                y.ctypes.data_as(_f32p), x.size))


def compare_matrices(a: np.ndarray, b: np.ndarray) -> float:
    keep: list = []
    ta = _taco_dense(np.ascontiguousarray(a, dtype=np.float32), keep)
    tb = _taco_dense(np.ascontiguousarray(b, dtype=np.float32), keep)
    return float(lib().segtaco_compare_matrices(ctypes.byref(ta), ctypes.byref(tb)))


def vendor_spmv(A: "CSR", x_ptr: int, y_ptr: int, alg: int = 0, stream: int | None = None,
                timed: bool = False):
    """rocSPARSE CSR SpMV y = A x on device operands (comparison row). Returns
    the compute-stage ms when timed (synchronous), else None (asynchronous)."""
    ms = ctypes.c_float(0.0)
    check(lib().segtaco_vendor_spmv(ctypes.byref(A), x_ptr, y_ptr, alg, stream or None,
                                    ctypes.byref(ms) if timed else None))
    return float(ms.value) if timed else None


def vendor_spmm(A: "CSR", B_ptr: int, ldb: int, N: int, C_ptr: int, ldc: int, alg: int = 0,
                stream: int | None = None, timed: bool = False):
    """rocSPARSE CSR SpMM on device operands (comparison row). Returns the
    compute-stage ms when timed (synchronous), else None (asynchronous)."""
    ms = ctypes.c_float(0.0)
    check(lib().segtaco_vendor_spmm(ctypes.byref(A), B_ptr, ldb, N, C_ptr, ldc, alg,
                                    stream or None, ctypes.byref(ms) if timed else None))
    return float(ms.value) if timed else None


# ---------------------------------------------------------------------------
# Device-resident API (device pointers as integers)
# ---------------------------------------------------------------------------

def csr_desc(M: int, K: int, nnz: int, pos_ptr: int, crd_ptr: int, vals_ptr: int | None,
             vals_dtype: int = DT_F32) -> CSR:
    return CSR(M, K, nnz, vals_dtype, pos_ptr, crd_ptr, vals_ptr or None)


def plan_resolve(A: CSR, N: int, b_dtype: int, plan: Plan, ldb: int | None = None,
                 ldc: int | None = None) -> Plan:
    p = Plan(*[getattr(plan, f) for f, _ in Plan._fields_])
    check(lib().segtaco_plan_resolve(ctypes.byref(A), N, b_dtype, ldb or N, ldc or N,
                                     ctypes.byref(p)))
    return p


def workspace_bytes(A: CSR, N: int, plan: Plan) -> int:
    return int(lib().segtaco_spmm_workspace_bytes(ctypes.byref(A), N, ctypes.byref(plan)))


def spmm(A: CSR, B_ptr: int, b_dtype: int, ldb: int, N: int, C_ptr: int, ldc: int,
         plan: Plan, ws_ptr: int | None, ws_bytes: int, stream: int | None = None) -> None:
    """C = A * B on device memory (asynchronous on `stream`)."""
    check(lib().segtaco_spmm(ctypes.byref(A), B_ptr, b_dtype, ldb, N, C_ptr, ldc,
                             ctypes.byref(plan), ws_ptr or None, ws_bytes, stream or None))


def spmm_prepare(A: CSR, b_dtype: int, ldb: int, N: int, ldc: int, plan: Plan,
                 ws_ptr: int | None, ws_bytes: int, stream: int | None = None) -> None:
    """Writes what a product with A computes from A alone (eb_cs: slot plan
    and column-sorted records) into the workspace (asynchronous)."""
    check(lib().segtaco_spmm_prepare(ctypes.byref(A), b_dtype, ldb, N, ldc, ctypes.byref(plan),
                                     ws_ptr or None, ws_bytes, stream or None))


def spmm_prepared(A: CSR, B_ptr: int, b_dtype: int, ldb: int, N: int, C_ptr: int, ldc: int,
                  plan: Plan, ws_ptr: int | None, ws_bytes: int,
                  stream: int | None = None) -> None:
    """spmm on a workspace prepared for this product (skips the preparation)."""
    check(lib().segtaco_spmm_prepared(ctypes.byref(A), B_ptr, b_dtype, ldb, N, C_ptr, ldc,
                                      ctypes.byref(plan), ws_ptr or None, ws_bytes,
                                      stream or None))


NOALLOC_WORKSPACE_BYTES = 128 << 20


def spmm_noalloc(A: CSR, B_ptr: int, b_dtype: int, ldb: int, N: int, C_ptr: int, ldc: int,
                 plan: Plan, stream: int | None = None) -> None:
    """C = A * B with the library's fixed device workspace (no allocation;
    calls overlapping on one GPU race)."""
    check(lib().segtaco_spmm_noalloc(ctypes.byref(A), B_ptr, b_dtype, ldb, N, C_ptr, ldc,
                                     ctypes.byref(plan), stream or None))


def spmm_stage_times(A: CSR, B_ptr: int, b_dtype: int, ldb: int, N: int, C_ptr: int, ldc: int,
                     plan: Plan, ws_ptr: int | None, ws_bytes: int, stream: int | None = None,
                     reps: int = 10) -> tuple[float, float, float]:
    """Mean (partition, main kernel, fix-up) milliseconds over `reps` runs,
    timed with HIP events on `stream` (synchronous)."""
    out = (ctypes.c_float * 3)()
    check(lib().segtaco_spmm_stage_times(ctypes.byref(A), B_ptr, b_dtype, ldb, N, C_ptr, ldc,
                                         ctypes.byref(plan), ws_ptr or None, ws_bytes,
                                         stream or None, reps, out))
    return float(out[0]), float(out[1]), float(out[2])


def sddmm_workspace_bytes(S: CSR) -> int:
    return int(lib().segtaco_sddmm_workspace_bytes(ctypes.byref(S)))


def sddmm(S: CSR, X_ptr: int, ldx: int, Y_ptr: int, ldy: int, J: int, dtype: int, out_ptr: int,
          ws_ptr: int | None, ws_bytes: int, stream: int | None = None) -> None:
    """out[p] = s_p * dot(X[row(p)], Y[crd[p]]) on device memory (asynchronous)."""
    check(lib().segtaco_sddmm(ctypes.byref(S), X_ptr, ldx, Y_ptr, ldy, J, dtype, out_ptr,
                              ws_ptr or None, ws_bytes, stream or None))


def block_starts(pos_ptr: int, M: int, T: int, num_blocks: int, out_ptr: int,
                 stream: int | None = None) -> None:
    """K21 partition on device (gpu_library.h:206-215)."""
    check(lib().segtaco_block_starts(pos_ptr, M, T, num_blocks, out_ptr, stream or None))


ROWS_SCATTER = 0
ROWS_SEARCH = 1


def tile_rows(pos_ptr: int, M: int, T: int, ntiles: int, rows_ptr: int, method: int,
              part_ptr: int | None = None, stream: int | None = None) -> None:
    """First row of every element-balanced tile (K21 search + repair), on
    device; method ROWS_SCATTER (standalone calls, SDDMM) or ROWS_SEARCH
    (batches, server; also the K21 array into part_ptr)."""
    check(lib().segtaco_tile_rows(pos_ptr, M, T, ntiles, method, rows_ptr, part_ptr or None,
                                  stream or None))


def workspace_layout(A: CSR, N: int, plan: Plan) -> dict:
    """Where an element-balanced call keeps its partition in the workspace."""
    out = (ctypes.c_int64 * 5)()
    check(lib().segtaco_spmm_workspace_layout(ctypes.byref(A), N, ctypes.byref(plan), out))
    return dict(ntiles=int(out[0]), fix_off=int(out[1]), rows_off=int(out[2]),
                head_off=int(out[3]), tail_off=int(out[4]))


def cs_workspace_layout(A: CSR, N: int, plan: Plan) -> dict:
    """Where an eb_cs call keeps its preparation in the workspace."""
    out = (ctypes.c_int64 * 10)()
    check(lib().segtaco_cs_workspace_layout(ctypes.byref(A), N, ctypes.byref(plan), out))
    keys = ("units", "cap", "packed", "rec_off", "slot_rows_off", "int_rows_off",
            "codes_off", "n_slots_off", "tile_count_off", "tile_flags_off")
    return {k: int(v) for k, v in zip(keys, out)}


def entry_rows(pos_ptr: int, M: int, nnz: int, T: int, tile_rows_ptr: int, rows_ptr: int,
               stream: int | None = None) -> None:
    """Row of every stored entry as eb_pr's walk derives it, on device."""
    check(lib().segtaco_entry_rows(pos_ptr, M, nnz, T, tile_rows_ptr, rows_ptr,
                                   stream or None))


def device_to_host(ptr: int, n: int, dtype=np.int32) -> np.ndarray:
    """Synchronous copy of n elements from device memory."""
    out = np.empty(n, dtype=dtype)
    if n:
        check(lib().segtaco_memcpy(out.ctypes.data, ptr, out.nbytes, 1))
    return out


# ---------------------------------------------------------------------------
# Autotuner (include/segtaco/autotune.h)
# ---------------------------------------------------------------------------

def row_stats(pos: np.ndarray, K: int) -> RowStats:
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    st = RowStats()
    check(lib().segtaco_row_stats(pos.ctypes.data_as(_i32p), len(pos) - 1, K, ctypes.byref(st)))
    return st


def tune_model(st: RowStats, N: int, dtype: str | int = "f32") -> Plan:
    dt = DTYPES[dtype] if isinstance(dtype, str) else dtype
    p = Plan()
    check(lib().segtaco_tune_model(ctypes.byref(st), N, dt, ctypes.byref(p)))
    return p


def tune(st: RowStats, N: int, dtype: str | int = "f32") -> Plan:
    """Cached measurement for the statistics' signature, else the model."""
    dt = DTYPES[dtype] if isinstance(dtype, str) else dtype
    p = Plan()
    check(lib().segtaco_tune(ctypes.byref(st), N, dt, ctypes.byref(p)))
    return p


def tune_signature(st: RowStats, N: int, dtype: str | int = "f32") -> int:
    dt = DTYPES[dtype] if isinstance(dtype, str) else dtype
    return int(lib().segtaco_tune_signature(ctypes.byref(st), N, dt))


def tune_cache_get(sig: int):
    p = Plan()
    return p if lib().segtaco_tune_cache_get(sig, ctypes.byref(p)) == OK else None


def tune_cache_put(sig: int, plan: Plan) -> None:
    check(lib().segtaco_tune_cache_put(sig, ctypes.byref(plan)))


def tune_cache_clear() -> None:
    check(lib().segtaco_tune_cache_clear())


def tune_cache_save(path: str) -> None:
    check(lib().segtaco_tune_cache_save(path.encode()))


def tune_cache_load(path: str) -> None:
    check(lib().segtaco_tune_cache_load(path.encode()))


def tune_measure(A: CSR, st: RowStats, B_ptr: int, b_dtype: int, ldb: int, N: int, C_ptr: int,
                 ldc: int, reps: int = 3, stream: int | None = None):
    """Times candidate plans on the device (eb_cs among them); returns (best
    plan, its ms). The plan cache keeps the best plan batches and the server
    can run (rb_sr .. eb_pr), which differs from the result when eb_cs wins."""
    p = Plan()
    ms = ctypes.c_float()
    check(lib().segtaco_tune_measure(ctypes.byref(A), ctypes.byref(st), B_ptr, b_dtype, ldb, N,
                                     C_ptr, ldc, reps, stream or None, ctypes.byref(p),
                                     ctypes.byref(ms)))
    return p, float(ms.value)


# ---------------------------------------------------------------------------
# Batched persistent launch (include/segtaco/batch.h)
# ---------------------------------------------------------------------------

SCHEDULE_AUTO = -1


def batch_item(A: CSR, B_ptr: int, ldb: int, C_ptr: int, ldc: int, N: int,
               plan: Plan | None = None, pos_host: np.ndarray | None = None) -> BatchItem:
    it = BatchItem()
    it.A = A
    it.B = B_ptr
    it.ldb = ldb
    it.C = C_ptr
    it.ldc = ldc
    it.N = N
    it.plan = plan if plan is not None else make_plan(0, comp_threshold=-1)
    if plan is None:
        it.plan.schedule = SCHEDULE_AUTO
    if pos_host is not None:
        it.pos_host = pos_host.ctypes.data_as(_i32p)
    return it


class Batch:
    """Owns a segtaco_batch_t: create once, run / capture / replay many times."""

    def __init__(self, items: list, dtype: str | int = "f32", stream: int | None = None):
        self.dtype = DTYPES[dtype] if isinstance(dtype, str) else dtype
        arr = (BatchItem * max(len(items), 1))(*items)
        self._items = arr
        h = _vp()
        check(lib().segtaco_batch_create(arr, len(items), self.dtype, stream or None,
                                         ctypes.byref(h)))
        self.handle = h
        self.n = len(items)

    def run(self, stream: int | None = None) -> None:
        check(lib().segtaco_batch_run(self.handle, stream or None))

    def capture(self) -> None:
        check(lib().segtaco_batch_capture(self.handle))

    def replay(self, stream: int | None = None) -> None:
        check(lib().segtaco_batch_replay(self.handle, stream or None))

    def plan(self, i: int) -> Plan:
        p = Plan()
        check(lib().segtaco_batch_plan(self.handle, i, ctypes.byref(p)))
        return p

    def units(self) -> int:
        return int(lib().segtaco_batch_units(self.handle))

    def workspace_bytes(self) -> int:
        return int(lib().segtaco_batch_workspace_bytes(self.handle))

    def partition(self, i: int):
        """Item i's (K21 array, tile rows) as the batch holds them, copied to
        the host; (None, None) for a row-balanced item."""
        nt = ctypes.c_int32()
        part, rows = _vp(), _vp()
        check(lib().segtaco_batch_partition(self.handle, i, ctypes.byref(nt), ctypes.byref(part),
                                            ctypes.byref(rows)))
        if nt.value == 0:
            return None, None
        return (device_to_host(part.value, nt.value + 1),
                device_to_host(rows.value, nt.value + 1))

    def destroy(self) -> None:
        if self.handle:
            lib().segtaco_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Continuous-batching server (include/segtaco/serve.h)
# ---------------------------------------------------------------------------

class Server:
    """Owns a segtaco_server_t. B and C of every request are device pointers."""

    def __init__(self, cache_bytes: int, inflight_bytes: int, max_batch: int = 64,
                 streams: int = 2, dtype: str | int = "f32"):
        self.dtype = DTYPES[dtype] if isinstance(dtype, str) else dtype
        cfg = ServeConfig(cache_bytes, inflight_bytes, max_batch, streams, self.dtype, 0)
        h = _vp()
        check(lib().segtaco_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h

    def add_matrix(self, pos: np.ndarray, crd: np.ndarray, vals, K: int) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        crd = np.ascontiguousarray(crd, dtype=np.int32)
        if crd.size == 0:
            crd = np.zeros(1, dtype=np.int32)
        M = len(pos) - 1
        vptr = None
        if vals is not None:
            v = (np.ascontiguousarray(vals, dtype=np.float32) if self.dtype == DT_F32
                 else to_half(self.dtype, vals))
            vptr = v.ctypes.data
        mid = ctypes.c_int32()
        check(lib().segtaco_server_add_matrix(self.handle, M, K, int(pos[M]) if M else 0,
                                              pos.ctypes.data_as(_i32p),
                                              crd.ctypes.data_as(_i32p), vptr,
                                              ctypes.byref(mid)))
        return int(mid.value)

    def submit(self, mid: int, B_ptr: int, ldb: int, N: int, C_ptr: int, ldc: int) -> int:
        rid = ctypes.c_int64()
        check(lib().segtaco_server_submit(self.handle, mid, B_ptr, ldb, N, C_ptr, ldc,
                                          ctypes.byref(rid)))
        return int(rid.value)

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = ctypes.c_int32()
        check(lib().segtaco_server_wait(self.handle, rid, timeout, ctypes.byref(st)))
        return int(st.value)

    def drain(self, timeout: float = -1.0) -> None:
        check(lib().segtaco_server_drain(self.handle, timeout))

    def replay(self, arrival_us, mids, B_ptrs, ldbs, Ns, C_ptrs, ldcs, timeout: float = 600.0):
        n = len(mids)
        arr = np.ascontiguousarray(arrival_us, dtype=np.float64)
        m = np.ascontiguousarray(mids, dtype=np.int32)
        lb = np.ascontiguousarray(ldbs, dtype=np.int64)
        nn = np.ascontiguousarray(Ns, dtype=np.int32)
        lc = np.ascontiguousarray(ldcs, dtype=np.int64)
        bp = (_vp * n)(*B_ptrs)
        cp = (_vp * n)(*C_ptrs)
        ids = np.zeros(n, dtype=np.int64)
        check(lib().segtaco_server_replay(
            self.handle, n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            m.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), bp,
            lb.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            nn.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cp,
            lc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), timeout))
        return ids

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().segtaco_server_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self) -> None:
        check(lib().segtaco_server_reset_stats(self.handle))

    def debug_fail_tick(self, ticks_ahead: int = 1) -> None:
        """Error-path testing: that tick's launch reports a HIP error without
        running."""
        check(lib().segtaco_server_debug_fail_tick(self.handle, ticks_ahead))

    def plan(self, mid: int, N: int) -> Plan:
        p = Plan()
        check(lib().segtaco_server_plan(self.handle, mid, N, ctypes.byref(p)))
        return p

    def destroy(self) -> None:
        if self.handle:
            lib().segtaco_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
