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
