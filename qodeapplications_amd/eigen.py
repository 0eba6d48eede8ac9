"""Lowest eigenpairs of the symmetric XR matrix, on the GPU in FP64.

Small matrices use a dense symmetric eigendecomposition; large ones a block
Davidson iteration whose matrix products run on libxrhip's FP64 GEMM.
"""
import torch

from .gemm import gemm_f64

DENSE_MAX = 6144


def lowest_eigenpairs(H, k=1, tol=1e-10, max_iter=300, block=None):
    """(values[k], vectors[n, k]) of the symmetric device matrix H."""
    n = H.shape[0]
    if n <= DENSE_MAX:
        w, v = torch.linalg.eigh(H)
        return w[:k].clone(), v[:, :k].contiguous()
    return davidson(lambda X, Y: gemm_f64(H, X, Y), torch.diagonal(H).clone(), n, k, tol, max_iter,
                    block, H.device)


def davidson(matmul, diag, n, k=1, tol=1e-10, max_iter=300, block=None, device=None,
             max_space=None):
    """Block Davidson with the diagonal preconditioner.  matmul(X, Y) writes
    H X into Y for X of shape (n, b)."""
    b = block or max(2 * k, 4)
    max_space = max_space or max(8 * b, 40)
    dt = torch.float64
    # start from the unit vectors of the lowest diagonal entries
    idx = torch.argsort(diag)[:b]
    V = torch.zeros(n, b, dtype=dt, device=device)
    V[idx, torch.arange(b, device=device)] = 1.0
    AV = torch.empty_like(V)
    matmul(V.contiguous(), AV)
    theta = None
    for it in range(max_iter):
        S = V.T @ AV
        S = 0.5 * (S + S.T)
        w, y = torch.linalg.eigh(S)
        theta = w[:k]
        X = V @ y[:, :b]
        R = AV @ y[:, :b] - X * w[:b]
        rn = torch.linalg.norm(R[:, :k], dim=0)
        if float(rn.max()) < tol:
            return theta.clone(), X[:, :k].contiguous()
        denom = w[:b][None, :] - diag[:, None]
        denom = torch.where(denom.abs() < 1e-8, torch.full_like(denom, 1e-8), denom)
        T = R / denom
        if V.shape[1] + b > max_space:
            V, AV = X.contiguous(), (AV @ y[:, :b]).contiguous()
        for _ in range(2):
            T = T - V @ (V.T @ T)
        Q, Rq = torch.linalg.qr(T)
        keep = torch.abs(torch.diagonal(Rq)) > 1e-12
        Q = Q[:, keep]
        if Q.shape[1] == 0:
            return theta.clone(), X[:, :k].contiguous()
        Q = Q - V @ (V.T @ Q)
        Q, _ = torch.linalg.qr(Q)
        AQ = torch.empty_like(Q)
        matmul(Q.contiguous(), AQ)
        V = torch.cat([V, Q], dim=1)
        AV = torch.cat([AV, AQ], dim=1)
    raise RuntimeError(f"Davidson did not converge in {max_iter} iterations (residual {float(rn.max()):.2e})")
