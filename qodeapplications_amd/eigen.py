"""Lowest eigenpairs of the symmetric XR matrix, on the GPU in FP64.

Small matrices use a dense symmetric eigendecomposition.  Large ones use a
block Davidson iteration whose products with H run on libxrhip's FP64 GEMM.
For tensor-product XR matrices the preconditioner is the exact inverse of
the monomer part H0 = sum_m H1_m (a Kronecker sum, inverted through the
small eigendecompositions of the H1_m), and the start vectors are the lowest
product states of H0.
"""
import torch

from .gemm import gemm_f64

DENSE_MAX = 6144


class KroneckerSumPreconditioner:
    """(H0 - theta)^-1 for H0 = H1_1 (x) 1 + 1 (x) H1_2 (+ ...), applied in the
    product eigenbasis of the monomer blocks."""

    def __init__(self, H1):
        self.w, self.U = [], []
        for h in H1:
            w, U = torch.linalg.eigh(0.5 * (h + h.T))
            self.w.append(w)
            self.U.append(U)
        self.dims = [h.shape[0] for h in H1]
        lam = self.w[0]
        for w in self.w[1:]:
            lam = (lam[:, None] + w[None, :]).reshape(-1)
        self.lam = lam          # eigenvalues of H0 in product order

    def _to_eig(self, R, transpose):
        # R: (n, b); apply U_m^T (transpose) or U_m along every fragment axis
        b = R.shape[1]
        T = R.T.reshape([b] + self.dims)
        for ax, U in enumerate(self.U):
            M = U.T if transpose else U
            T = torch.movedim(torch.tensordot(M, T, dims=([1], [ax + 1])), 0, ax + 1)
        return T.reshape(b, -1).T.contiguous()

    def __call__(self, R, theta):
        Y = self._to_eig(R, True)
        den = self.lam[:, None] - theta[None, :]
        den = torch.where(den.abs() < 1e-8, torch.full_like(den, 1e-8), den)
        return self._to_eig(Y / den, False)

    def guess(self, b):
        idx = torch.argsort(self.lam)[:b]
        E = torch.zeros(self.lam.shape[0], b, dtype=torch.float64, device=self.lam.device)
        E[idx, torch.arange(b, device=E.device)] = 1.0
        return self._to_eig(E, False)


def lowest_eigenpairs(H, k=1, tol=1e-9, max_iter=400, H1=None, block=None):
    """(values[k], vectors[n, k]) of the symmetric device matrix H.  H1: the
    monomer blocks when H is a tensor-product XR matrix (preconditioner)."""
    n = H.shape[0]
    if n <= DENSE_MAX:
        w, v = torch.linalg.eigh(0.5 * (H + H.T))
        return w[:k].clone(), v[:, :k].contiguous()
    pre = KroneckerSumPreconditioner(H1) if H1 is not None else None
    return davidson(lambda X, Y: gemm_f64(H, X, Y), torch.diagonal(H).clone(), n, k, tol,
                    max_iter, block, H.device, precond=pre)


def davidson(matmul, diag, n, k=1, tol=1e-9, max_iter=400, block=None, device=None,
             max_space=None, precond=None):
    """Block Davidson.  matmul(X, Y) writes H X into Y for X of shape (n, b);
    precond(R, theta) approximates (H - theta)^-1 R (default: diagonal).
    Converged when every residual norm is below tol * max(1, |theta|)."""
    b = block or max(2 * k, 4)
    max_space = max_space or max(10 * b, 48)
    dt = torch.float64
    if precond is not None:
        V = precond.guess(b)
    else:
        idx = torch.argsort(diag)[:b]
        V = torch.zeros(n, b, dtype=dt, device=device)
        V[idx, torch.arange(b, device=device)] = 1.0
    V, _ = torch.linalg.qr(V)
    V = V.contiguous()
    AV = torch.empty(V.shape, dtype=dt, device=device)
    matmul(V, AV)
    for it in range(max_iter):
        S = V.T @ AV
        S = 0.5 * (S + S.T)
        w, y = torch.linalg.eigh(S)
        theta = w[:b]
        X = V @ y[:, :b]
        AX = AV @ y[:, :b]
        R = AX - X * theta
        rn = torch.linalg.norm(R[:, :k], dim=0)
        if bool((rn <= tol * torch.clamp(theta[:k].abs(), min=1.0)).all()):
            return theta[:k].clone(), X[:, :k].contiguous()
        if precond is not None:
            T = precond(R, theta)
        else:
            den = theta[None, :] - diag[:, None]
            den = torch.where(den.abs() < 1e-8, torch.full_like(den, 1e-8), den)
            T = R / den
        if V.shape[1] + b > max_space:      # thick restart on the current Ritz vectors
            V, AV = X.contiguous(), AX.contiguous()
        for _ in range(2):
            T = T - V @ (V.T @ T)
        Q, Rq = torch.linalg.qr(T)
        keep = torch.abs(torch.diagonal(Rq)) > 1e-10 * float(torch.abs(torch.diagonal(Rq)).max() + 1e-300)
        Q = Q[:, keep]
        if Q.shape[1] == 0:
            return theta[:k].clone(), X[:, :k].contiguous()
        Q = Q - V @ (V.T @ Q)
        Q, _ = torch.linalg.qr(Q)
        Q = Q.contiguous()
        AQ = torch.empty(Q.shape, dtype=dt, device=Q.device)
        matmul(Q, AQ)
        V = torch.cat([V, Q], dim=1)
        AV = torch.cat([AV, AQ], dim=1)
    raise RuntimeError(f"Davidson did not converge in {max_iter} iterations "
                       f"(residual {float(rn.max()):.2e})")
