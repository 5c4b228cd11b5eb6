"""MI355X drop-in for the reference's ``gpu_flexdtw`` module (gpu_flexdtw.py).

Public API (same names and arguments as the reference; device arrays are ROCm
torch tensors instead of CuPy arrays):

    is_gpu_available() -> bool                                   gpu_flexdtw.py:86-92
    cost_matrix_from_features_gpu(F1, F2, dtype=torch.float64)   gpu_flexdtw.py:95-109
    cost_matrix_to_gpu_f32(C)                                    gpu_flexdtw.py:112-116
    cost_matrix_to_gpu(C, dtype=torch.float64)
    gather_c_at_global_indices(C_dev, rows, cols)                gpu_flexdtw.py:119-131
    flexdtw_chunk_from_global_C(C_dev, L2_global, r0, r1, c0, c1, steps,
                                weights, buffer=1.0)             gpu_flexdtw.py:288-402
        -> best_cost, wp (2, n), D f64, P int32, B int8, debug

Differences from the reference, all deliberate:
  * the DP runs in IEEE FP64 (the reference kernel is FP32 fast-math), so D,
    P, B and wp equal FlexDTW.flexdtw exactly;
  * flexdtw_chunk_from_global_C returns the documented 6-tuple (the reference
    returns only (D, P, B) at gpu_flexdtw.py:402);
  * indices are 64-bit (gpu_flexdtw.py:188 overflows past 2^31 cells);
  * cost_matrix_from_features_gpu returns FP64 unless another dtype is asked.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .engine import StepWeights

__all__ = ["is_gpu_available", "cost_matrix_from_features_gpu", "cost_matrix_to_gpu_f32",
           "cost_matrix_to_gpu", "gather_c_at_global_indices", "flexdtw_chunk_from_global_C",
           "DIAG_THREADS"]

# Kept for API compatibility (gpu_flexdtw.py:139).  The HIP kernels pick their
# own launch shape: one workgroup of up to 1024 threads per tile, 2-8 columns
# per lane.  Setting this has no effect.
DIAG_THREADS: int = 256


def _torch():
    import torch
    return torch


def _device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("parflexdtw_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def is_gpu_available() -> bool:
    """True when a ROCm device is visible and the HIP library loads."""
    try:
        torch = _torch()
        if not torch.cuda.is_available():
            return False
        lib()
        return True
    except Exception:
        return False


def cost_matrix_to_gpu(C, dtype=None):
    """Upload a host cost matrix as a contiguous device tensor (FP64 default)."""
    torch = _torch()
    dtype = torch.float64 if dtype is None else dtype
    if isinstance(C, torch.Tensor):
        return C.to(device=_device(), dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(C, dtype=np.float64)).to(
        device=_device(), dtype=dtype).contiguous()


def cost_matrix_to_gpu_f32(C):
    """gpu_flexdtw.py:112-116: contiguous float32 device copy of C."""
    return cost_matrix_to_gpu(C, dtype=_torch().float32)


def cost_matrix_from_features_gpu(F1, F2, dtype=None):
    """C = 1 - L2norm(F1)^T L2norm(F2) on the GPU (gpu_flexdtw.py:95-109).

    F1, F2: (K, L1), (K, L2) host arrays or device tensors.  dtype float64
    (default): the K products are summed in increasing order with separately
    rounded operations (bit-identical to parflexdtw_amd.synth.cost_matrix_det).
    dtype bfloat16 or float32: the normalised features are rounded to BF16 and
    multiplied on the MFMA units with FP32 accumulation (precision form)."""
    torch = _torch()
    dtype = torch.float64 if dtype is None else dtype
    dev = _device()
    F1d = cost_matrix_to_gpu(F1, torch.float64)
    F2d = cost_matrix_to_gpu(F2, torch.float64)
    if F1d.dim() != 2 or F2d.dim() != 2 or F1d.shape[0] != F2d.shape[0]:
        raise ValueError("F1 and F2 must be (K, L1) and (K, L2) with the same K")
    K, N = F1d.shape
    M = F2d.shape[1]
    codes = {torch.float64: _lib.PF_F64, torch.float32: _lib.PF_F32,
             torch.bfloat16: _lib.PF_BF16}
    if dtype not in codes:
        raise ValueError(f"unsupported cost dtype {dtype}")
    code = codes[dtype]
    C = torch.empty((N, M), dtype=dtype, device=dev)
    nbytes = ctypes.c_int64()
    check(lib().pf_cost_workspace(code, K, N, M, ctypes.byref(nbytes)), "pf_cost_workspace")
    work = torch.empty(-(-nbytes.value // 8), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().pf_cost(code, F1d.data_ptr(), F2d.data_ptr(), K, N, M, C.data_ptr(), M,
                        work.data_ptr(), stream), "pf_cost")
    return C


def gather_c_at_global_indices(C_dev, global_rows, global_cols):
    """gpu_flexdtw.py:119-131: C_dev[gr, gc] as host float64 (64-bit indices)."""
    torch = _torch()
    gr = torch.as_tensor(np.asarray(global_rows, dtype=np.int64), device=C_dev.device)
    gc = torch.as_tensor(np.asarray(global_cols, dtype=np.int64), device=C_dev.device)
    return C_dev[gr, gc].to(torch.float64).cpu().numpy()


def flexdtw_chunk_from_global_C(C_dev, L2_global, r0, r1, c0, c1, steps, weights, buffer=1.0):
    """FlexDTW on the chunk C_dev[r0:r1, c0:c1] (gpu_flexdtw.py:288-402).

    Returns (best_cost, wp (2, n) int64, D (R, W) f64, P int32, B int8, debug)
    with debug the candidate list of find_best_endpoint (FlexDTW.py:1098-1148)."""
    torch = _torch()
    R, W = int(r1) - int(r0), int(c1) - int(c0)
    if R <= 0 or W <= 0:
        raise ValueError(f"Empty chunk: R={R}, C={W}")
    if C_dev.dim() != 2 or int(L2_global) != C_dev.shape[1]:
        raise ValueError("L2_global must equal C_dev.shape[1]")
    if r0 < 0 or c0 < 0 or r1 > C_dev.shape[0] or c1 > C_dev.shape[1]:
        raise ValueError("chunk lies outside C_dev")
    sw = StepWeights(steps, weights)
    Cd = C_dev
    if Cd.dtype != torch.float64 or not Cd.is_contiguous():
        Cd = Cd.to(torch.float64).contiguous()
    dev = Cd.device
    D = torch.empty((R, W), dtype=torch.float64, device=dev)
    P = torch.empty((R, W), dtype=torch.int32, device=dev)
    B = torch.empty((R, W), dtype=torch.int8, device=dev)
    buf = int(buffer)
    ncand = max(W - buf, 0) + max(R - 1 - buf, 0)
    avg = torch.empty(max(ncand, 1), dtype=torch.float64, device=dev)
    best = torch.empty(3, dtype=torch.float64, device=dev)
    path = torch.empty((R + W + 2) * 2, dtype=torch.int64, device=dev)
    plen = torch.empty(1, dtype=torch.int64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = lib()
    check(L.pf_tile_full(Cd.data_ptr(), int(L2_global), int(r0), int(r1), int(c0), int(c1),
                         sw.sp, sw.wp, sw.K, D.data_ptr(), P.data_ptr(), B.data_ptr(), st),
          "pf_tile_full")
    check(L.pf_tile_endpoint(D.data_ptr(), P.data_ptr(), R, W, buf, avg.data_ptr(),
                             best.data_ptr(), st), "pf_tile_endpoint")
    check(L.pf_tile_backtrace(B.data_ptr(), R, W, sw.sp, sw.K, best.data_ptr(), path.data_ptr(),
                              plen.data_ptr(), st), "pf_tile_backtrace")
    Dh = D.cpu().numpy()
    Ph = P.cpu().numpy()
    Bh = B.cpu().numpy()
    avg_h = avg.cpu().numpy()[:ncand]
    best_h = best.cpu().numpy()
    n = int(plen.cpu().numpy()[0])
    wp = path.cpu().numpy()[:2 * n].reshape(n, 2).T.copy()
    n_row = max(W - buf, 0)
    debug = []
    for k in range(ncand):
        if k < n_row:
            r, c = R - 1, buf + k
            debug.append((c - W + 1, float(avg_h[k]), r, c))
        else:
            r, c = R - 2 - (k - n_row), W - 1
            debug.append((R - 1 - r, float(avg_h[k]), r, c))
    return float(best_h[0]), wp, Dh, Ph, Bh, debug
