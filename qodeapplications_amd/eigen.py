"""Lowest eigenpairs of the symmetric XR matrix, on the GPU in FP64.

Small matrices use a dense symmetric eigendecomposition.  Large ones use a
block Davidson iteration whose products with H run on libxrhip's FP64 GEMM.
For tensor-product XR matrices the preconditioner is the exact inverse of
the monomer part H0 = sum_m H1_m (a Kronecker sum, inverted through the
small eigendecompositions of the H1_m), and the start vectors are the lowest
product states of H0.
"""
import numpy as np
import torch

from ._lib import context, ptr
from .gemm import gemm_f64, mm

DENSE_MAX = 6144
PRECOND_FLOOR = 1e-8     # |denominator| floor of the diagonal preconditioners


def upload(a, device):
    """Host array -> device tensor without waiting for the work already
    queued on the stream (a pageable copy would): staged in pinned memory and
    copied on the stream, which keeps the staging block until the copy ran."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.float32:
        t = t.double()
    return t.pin_memory().to(device, non_blocking=True)


def precond_divide(X, lam, theta, sign=1.0, out=None):
    """out[i, j] = X[i, j] / d, d = sign (lam[i] - theta[j]) with |d| floored
    at PRECOND_FLOOR (to +PRECOND_FLOOR), one library pass (xr_precond_divide);
    out may be X."""
    if out is None:
        out = torch.empty_like(X)
    n, b = X.shape
    context(X.device.index).call("xr_precond_divide", n, b, ptr(X), X.stride(0), ptr(out), out.stride(0),
                                 ptr(lam), ptr(theta.contiguous()), float(sign), PRECOND_FLOOR)
    return out


def residual(X, AX, theta, k):
    """(R, norms): R = AX - X diag(theta) and the norms of R's first k
    columns, one library pass (xr_residual; fixed-order column sums)."""
    n, b = X.shape
    R = torch.empty(n, b, dtype=X.dtype, device=X.device)
    norms = torch.empty(k, dtype=X.dtype, device=X.device)
    context(X.device.index).call("xr_residual", n, b, k, ptr(X), X.stride(0), ptr(AX), AX.stride(0),
                                 ptr(theta.contiguous()), ptr(R), b, ptr(norms))
    return R, norms


def column_sumsq(A):
    """sum_i A[i, j]^2 per column, one library pass (xr_column_norms)."""
    n, b = A.shape
    out = torch.empty(b, dtype=A.dtype, device=A.device)
    context(A.device.index).call("xr_column_norms", n, b, ptr(A), A.stride(0), ptr(out), 1)
    return out


def store_rows(dst, A):
    """dst (b x n, row stride dst.stride(0)) = A^T for A (n x b): a block of
    vectors written into a basis kept one vector per row (xr_transpose)."""
    n, b = A.shape
    context(A.device.index).call("xr_transpose", n, b, ptr(A), A.stride(0), ptr(dst), dst.stride(0))


def _narrow(b):
    """The library's column-reduction kernels take blocks of at most 32."""
    return b <= 32


class KroneckerSumPreconditioner:
    """(H0 - theta)^-1 for H0 = H1_1 (x) 1 + 1 (x) H1_2 (+ ...), applied in the
    product eigenbasis of the monomer blocks.  With the states' charges
    given, each monomer block is diagonalised per charge (it conserves
    charge), so every eigenvector has a definite charge (qcol)."""

    def __init__(self, H1, charges=None):
        self.w, self.U, self.qcol = [], [], []
        for m, h in enumerate(H1):
            hs = 0.5 * (h + h.T)
            if charges is None:
                w, U = torch.linalg.eigh(hs)
                q = np.zeros(h.shape[0], dtype=np.int64)
            else:
                qm = np.asarray(charges[m], dtype=np.int64)
                U = torch.zeros_like(hs)
                w = torch.empty(h.shape[0], dtype=torch.float64, device=h.device)
                q = np.empty(h.shape[0], dtype=np.int64)
                col = 0
                for c in np.unique(qm):
                    idx = torch.as_tensor(np.nonzero(qm == c)[0], device=h.device)
                    wc, Uc = torch.linalg.eigh(hs[idx][:, idx])
                    nc = idx.numel()
                    U[idx, col:col + nc] = Uc
                    w[col:col + nc] = wc
                    q[col:col + nc] = c
                    col += nc
            self.w.append(w)
            self.U.append(U)
            self.qcol.append(q)
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
        return self._to_eig(precond_divide(Y, self.lam, theta, out=Y), False)

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


def _orthonormalize(T, V=None, passes=2, Vt=None, extra=None):
    """Columns of T made orthonormal (and orthogonal to V's columns) by
    projected Cholesky QR; the tiny b x b factorisations run on the host.
    The projection is repeated only when it removed more than 3/4 of a
    column's squared norm (the Daniel-Gragg-Kaufman-Stewart test), since
    each pass reads all of V.  Vt: the basis stored one vector per row
    instead of V.  Returns the kept columns (rank-revealing through the
    Cholesky pivots).  Each pass makes one device-to-host transfer; extra (a
    device vector) rides along with the first, and (T, extra on the host)
    is returned when it is given."""
    n0 = None
    ex = None

    def done(T):
        return T if extra is None else (T, ex)

    for p in range(passes):
        n0_t = None
        if (V is not None and V.shape[1]) or (Vt is not None and Vt.shape[0]):
            if p == 0:
                n0_t = column_sumsq(T) if _narrow(T.shape[1]) and T.stride(1) == 1 else (T * T).sum(0)
            if Vt is not None:          # basis stored one vector per row
                mm(Vt, mm(Vt, T), transA=1, alpha=-1.0, out=T, beta=1.0)
            else:
                mm(V, mm(V, T, transA=1), alpha=-1.0, out=T, beta=1.0)
        nb = T.shape[1]
        parts = [mm(T, T, transA=1).reshape(-1)]
        if n0_t is not None:
            parts.append(n0_t)
        if p == 0 and extra is not None:
            parts.append(extra.reshape(-1).to(T.dtype))
        h = torch.cat(parts).cpu().numpy() if len(parts) > 1 else parts[0].cpu().numpy()
        G = h[:nb * nb].reshape(nb, nb)
        off = nb * nb
        if n0_t is not None:
            n0 = h[off:off + nb].copy()
            off += nb
        if p == 0 and extra is not None:
            ex = h[off:].copy()
        d = np.sqrt(np.maximum(np.diag(G), 0.0))
        keep = d > 1e-10 * max(float(d.max()), 1e-300)
        if not keep.any():
            return done(T[:, :0])
        if not keep.all():
            T = T[:, torch.as_tensor(np.nonzero(keep)[0], device=T.device)].contiguous()
            G = G[np.ix_(keep, keep)]
            if n0 is not None:
                n0 = n0[keep]
        try:
            L = np.linalg.cholesky(G)
        except np.linalg.LinAlgError:
            w, U = np.linalg.eigh(G)
            good = w > 1e-14 * w.max()
            Minv = U[:, good] / np.sqrt(w[good])
            T = mm(T, upload(Minv, T.device))
            n0 = None
            continue
        Linv = np.ascontiguousarray(np.linalg.inv(L).T)
        T = mm(T, upload(Linv, T.device))
        if n0 is not None and np.all(np.diag(G) > 0.25 * n0):
            break
    return done(T.contiguous())


def davidson(matmul, diag, n, k=1, tol=1e-9, max_iter=400, block=None, device=None,
             max_space=None, precond=None, start=None, keep=8, olsen=False, strict=True,
             state=None, stop_above=None, stop_margin=4.0, info=None):
    """Block Davidson.  matmul(X, Y) writes H X into Y for X of shape (n, b);
    precond(R, theta) approximates (H - theta)^-1 R (default: diagonal).
    Converged when every residual norm is below tol * max(1, |theta|), or
    when it has stalled at the rounding floor within 100x of that.  A full
    subspace restarts on its `keep` lowest Ritz vectors.  olsen: apply
    Olsen's correction to the preconditioned residual.  strict=False returns
    the current Ritz pairs after max_iter instead of raising.  state: a dict
    that receives the subspace on return and, when it already holds one (same
    n and block), continues from it instead of from the start vectors.
    stop_above: give up (info["abandoned"] = True) as soon as every wanted
    Ritz value minus stop_margin times its residual norm lies above this
    energy, i.e. when these roots can no longer be among the lowest ones of
    a larger problem (a heuristic: an eigenvalue lies within one residual
    norm of each Ritz value, but early in the iteration not necessarily the
    lowest one, hence the margin).  info: a dict that receives the iteration count, the last
    residual norms and the abandoned / converged flags (neither set: the
    call ran out of iterations).  The subspace lives in preallocated device
    buffers; the Rayleigh-Ritz eigenproblem and the b x b factorisations are
    solved on the host."""
    b = block or max(2 * k, 4)
    max_space = max_space or max(8 * b, 24)
    if start is not None:
        max_space = max(max_space, start.shape[1] + b)
    dt = torch.float64
    if state is not None and state.get("b") == b and state.get("n") == n:
        Vt, AVt, m, S = state["Vt"], state["AVt"], state["m"], state["S"]
        max_space = Vt.shape[0] - b
    else:
        if start is not None:
            V0 = start
        elif precond is not None and hasattr(precond, "guess"):
            V0 = precond.guess(b)
        else:
            idx = torch.argsort(diag)[:b]
            V0 = torch.zeros(n, b, dtype=dt, device=device)
            V0[idx, torch.arange(b, device=device)] = 1.0
        # basis and its image stored one vector per row: the tall-skinny
        # products then read contiguous vectors
        Vt = torch.empty(max_space + b, n, dtype=dt, device=device)
        AVt = torch.empty(max_space + b, n, dtype=dt, device=device)
        V0 = _orthonormalize(V0.contiguous())
        m = V0.shape[1]
        store_rows(Vt[:m], V0)
        AX0 = torch.empty(n, m, dtype=dt, device=device)
        matmul(V0, AX0)
        store_rows(AVt[:m], AX0)
        S = mm(V0, AX0, transA=1).cpu().numpy()

    def keep_state():
        if state is not None:
            state.update(Vt=Vt, AVt=AVt, m=m, S=S, b=b, n=n, hist=hist)

    def converged():
        if info is not None:
            info["converged"] = True
        keep_state()
        return theta[:k].clone(), X[:, :k].contiguous()

    hist = list(state.get("hist", [])) if state is not None and state.get("n") == n else []
    for it in range(max_iter):
        S = 0.5 * (S + S.T)
        w, y = np.linalg.eigh(S)
        nb = min(b, m)
        yb = upload(y[:, :nb], device)
        theta = upload(w[:nb], device)
        X = mm(Vt[:m], yb, transA=1)
        AX = mm(AVt[:m], yb, transA=1)
        # the residual norms reach the host with the orthonormalisation's
        # first transfer (one synchronisation fewer per iteration); the
        # correction computed meanwhile is dropped when they show convergence
        if _narrow(nb):
            R, rn_dev = residual(X, AX, theta, min(k, nb))
        else:
            R = AX - X * theta
            rn_dev = torch.linalg.norm(R[:, :k], dim=0)
        target = tol * np.maximum(np.abs(w[:k]), 1.0)
        if precond is not None:
            T = precond(R, theta)
            if olsen:
                # Olsen's correction: remove the part of M^-1 r along M^-1 x,
                # so the new direction is not dominated by the Ritz vector itself
                PX = precond(X, theta)
                eps = (X * T).sum(0) / (X * PX).sum(0)
                T = T - PX * eps[None, :]
        else:
            T = precond_divide(R, diag, theta, sign=-1.0)     # R / (theta - diag), floored
        restart = m + nb > max_space
        if restart:                         # thick restart on the lowest Ritz vectors
            nk = min(m, max(2 * nb, keep))
            yk = upload(y[:, :nk], device)
            Xk, AXk = mm(Vt[:m], yk, transA=1), mm(AVt[:m], yk, transA=1)
            store_rows(Vt[:nk], Xk)
            store_rows(AVt[:nk], AXk)
            S = np.diag(w[:nk])
            m = nk
        Q, rn = _orthonormalize(T.contiguous(), Vt=Vt[:m], extra=rn_dev)
        if info is not None:
            info.update(iters=it + 1, rn=rn, abandoned=False, converged=False)
        if np.all(rn <= target):
            return converged()
        if stop_above is not None and np.all(w[:k] - stop_margin * rn > stop_above):
            if info is not None:
                info["abandoned"] = True
            keep_state()
            return theta[:k].clone(), X[:, :k].contiguous()
        # rounding floor: the residual is within 100x of the target and has
        # stopped falling over the last 10 iterations
        hist.append(float(np.max(rn / target)))
        if len(hist) > 20 and hist[-1] < 100.0 and min(hist[-10:]) > 0.5 * min(hist[:-10]):
            return converged()
        q = Q.shape[1]
        if q == 0:
            return converged()
        AQ = torch.empty(n, q, dtype=dt, device=device)
        matmul(Q, AQ)
        # grow the projected matrix by the new rows / columns: one product
        # over the extended basis gives V^T A Q and Q^T A Q in one transfer
        store_rows(Vt[m:m + q], Q)
        store_rows(AVt[m:m + q], AQ)
        VA = mm(Vt[:m + q], AQ).cpu().numpy()
        S = np.block([[S, VA[:m]], [VA[:m].T, VA[m:]]])
        m += q
    if not strict:
        keep_state()
        return theta[:k].clone(), X[:, :k].contiguous()
    raise RuntimeError(f"Davidson did not converge in {max_iter} iterations "
                       f"(residual {float(rn.max()):.2e})")
