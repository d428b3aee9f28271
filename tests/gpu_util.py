"""Device plumbing for the GPU tests: torch tensors hold device memory, every
computation goes through libsegtaco's C-ABI (segtaco_amd)."""
from __future__ import annotations

import numpy as np

import segtaco_amd as S


def torch():
    import torch as _t
    return _t


def dev(a: np.ndarray):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).to("cuda")


def stream_ptr() -> int:
    return torch().cuda.current_stream().cuda_stream


class DeviceCSR:
    """pos/crd/vals resident on the device (torch tensors kept alive)."""

    def __init__(self, pos, crd, vals, K: int, dtype: str = "f32"):
        t = torch()
        self.M = len(pos) - 1
        self.K = K
        self.nnz = int(pos[-1]) if self.M > 0 else 0
        self.dtype = S.DTYPES[dtype]
        self.pos = dev(np.asarray(pos, np.int32))
        self.crd = dev(np.asarray(crd, np.int32) if len(crd) else np.zeros(1, np.int32))
        if vals is None:
            self.vals = None
        elif dtype == "f32":
            self.vals = dev(np.asarray(vals, np.float32))
        else:
            self.vals = dev(S.to_half(dtype, vals).view(np.int16))
        self.desc = S.csr_desc(self.M, K, self.nnz, self.pos.data_ptr(), self.crd.data_ptr(),
                               self.vals.data_ptr() if self.vals is not None else None,
                               self.dtype)
        self._t = t


def dense_dev(B: np.ndarray, dtype: str = "f32"):
    if dtype == "f32":
        return dev(np.asarray(B, np.float32))
    return dev(S.to_half(dtype, B).view(np.int16))


def run_spmm(A: DeviceCSR, Bd, N: int, plan: S.Plan, ldb: int | None = None,
             ldc: int | None = None, b_dtype: str = "f32", C=None, sync=True):
    """C = A*B through segtaco_spmm; returns the device C tensor (M x ldc)."""
    t = torch()
    ldb = ldb or N
    ldc = ldc or N
    if C is None:
        C = t.full((max(A.M, 1), ldc), float("nan"), dtype=t.float32, device="cuda")
    bd = S.DTYPES[b_dtype]
    ws_bytes = S.workspace_bytes(A.desc, N, S.plan_resolve(A.desc, N, bd, plan, ldb, ldc))
    ws = t.empty(max(ws_bytes, 1), dtype=t.uint8, device="cuda")
    S.spmm(A.desc, Bd.data_ptr(), bd, ldb, N, C.data_ptr(), ldc, plan, ws.data_ptr(), ws_bytes,
           stream_ptr())
    if sync:
        t.cuda.synchronize()
    return C
