"""FP64 matrix-core GEMM of libxrhip (xr_gemm_f64) on torch tensors."""
import torch

from ._lib import context, ptr


def gemm_f64(A, B, C, transA=0, transB=0, alpha=1.0, beta=0.0):
    """C = alpha op(A) op(B) + beta C in place, for 2-D or batched 3-D
    contiguous float64 CUDA tensors.  transA = 1: A holds op(A)^T."""
    batched = C.dim() == 3
    A3 = A if A.dim() == 3 else A.unsqueeze(0)
    B3 = B if B.dim() == 3 else B.unsqueeze(0)
    C3 = C if batched else C.unsqueeze(0)
    for t in (A3, B3, C3):
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("gemm_f64 needs contiguous float64 CUDA tensors")
    nb = C3.shape[0]
    M, N = C3.shape[1], C3.shape[2]
    K = A3.shape[1] if transA else A3.shape[2]
    if (A3.shape[2] if transA else A3.shape[1]) != M:
        raise ValueError("A does not match C")
    if (B3.shape[2] if transB else B3.shape[1]) != K or (B3.shape[1] if transB else B3.shape[2]) != N:
        raise ValueError("B does not match A and C")
    sA = A3.shape[1] * A3.shape[2] if A3.shape[0] > 1 else 0
    sB = B3.shape[1] * B3.shape[2] if B3.shape[0] > 1 else 0
    ctx = context(C.device.index)
    ctx.call("xr_gemm_f64", int(transA), int(transB), M, N, K, float(alpha), ptr(A3), A3.shape[2], sA,
             ptr(B3), B3.shape[2], sB, float(beta), ptr(C3), N, M * N, nb)
    return C


MODES = {"f64": 0, "split_f32": 1, "split_i8": 2}


def mm(A, B, transA=0, transB=0, alpha=1.0, out=None, beta=0.0, split=False, mode=None):
    """out = alpha op(A) op(B) + beta out for 2-D float64 CUDA tensors that
    are row-major with any leading dimension (views of wider buffers work
    without copies).  transA = 1 uses A^T, transB = 1 uses B^T; split=True
    (or "f32") runs the split-FP32 product, split="i8" the split-int8 one.
    mode ('f64', 'split_f32', 'split_i8'), when given, runs the product as
    the two-electron contractions do in that precision mode (xr_gemm_mode:
    split-int8 only where it pays)."""
    entry = {False: "xr_gemm_f64", True: "xr_gemm_split_f32", "f32": "xr_gemm_split_f32",
             "i8": "xr_gemm_split_i8"}.get(split)
    if entry is None:
        raise ValueError("split must be False, True, 'f32' or 'i8'")
    if mode is not None and mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    for t in (A, B):
        if t.dim() != 2 or t.dtype != torch.float64 or not t.is_cuda:
            raise ValueError("mm needs 2-D float64 CUDA tensors")
    if A.stride(1) != 1:
        A = A.contiguous()
    if B.stride(1) != 1:
        B = B.contiguous()
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    N = B.shape[0] if transB else B.shape[1]
    if (B.shape[1] if transB else B.shape[0]) != K:
        raise ValueError("inner dimensions differ")
    if out is None:
        out = torch.empty(M, N, dtype=torch.float64, device=A.device)
        beta = 0.0
    if out.stride(1) != 1 or tuple(out.shape) != (M, N):
        raise ValueError("out must be row-major (M, N)")
    ctx = context(A.device.index)
    if mode is not None:
        ctx.call("xr_gemm_mode", MODES[mode], int(transA), int(transB), M, N, K, float(alpha), ptr(A), A.stride(0),
                 0, ptr(B), B.stride(0), 0, float(beta), ptr(out), out.stride(0), 0, 1)
        return out
    ctx.call(entry, int(transA), int(transB), M, N, K,
             float(alpha), ptr(A), A.stride(0), 0, ptr(B), B.stride(0), 0, float(beta), ptr(out),
             out.stride(0), 0, 1)
    return out


def syrk(A, trans=0, alpha=1.0, out=None):
    """alpha A A^T (trans = 0) or alpha A^T A (trans = 1) for a 2-D row-major
    float64 CUDA tensor: the upper tiles on the matrix cores, the lower
    triangle mirrored (xr_gemm_syrk; the result is exactly symmetric)."""
    if A.dim() != 2 or A.dtype != torch.float64 or not A.is_cuda:
        raise ValueError("syrk needs a 2-D float64 CUDA tensor")
    if A.stride(1) != 1:
        A = A.contiguous()
    n, K = (A.shape[1], A.shape[0]) if trans else (A.shape[0], A.shape[1])
    if out is None:
        out = torch.empty(n, n, dtype=torch.float64, device=A.device)
    if out.stride(1) != 1 or tuple(out.shape) != (n, n):
        raise ValueError("out must be row-major (n, n)")
    context(A.device.index).call("xr_gemm_syrk", int(trans), n, K, float(alpha), ptr(A), A.stride(0), ptr(out),
                                 out.stride(0))
    return out


def permute4(src, perm, out=None, accumulate=False):
    """out = src.permute(perm) as a contiguous tensor (or out += it) for a
    contiguous 4-index float64 CUDA tensor, one library pass (xr_permute4)."""
    import numpy as np
    import ctypes
    if src.dim() != 4 or not src.is_contiguous() or src.dtype != torch.float64:
        raise ValueError("permute4 needs a contiguous 4-index float64 tensor")
    shape = tuple(src.shape[p] for p in perm)
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs out")
        out = torch.empty(shape, dtype=torch.float64, device=src.device)
    elif out.numel() != src.numel() or not out.is_contiguous():
        raise ValueError("out must be contiguous with the permuted shape's size")
    dims = np.asarray(src.shape, dtype=np.int64)
    pp = np.asarray(perm, dtype=np.int32)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    context(src.device.index).call("xr_permute4", ptr(src), P(dims), P(pp), ptr(out), 1 if accumulate else 0)
    return out
