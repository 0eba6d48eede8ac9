"""XR-CCSD: coupled cluster with fragment-state singles and doubles on the
XR Hamiltonian (general-XRCC/excitonic.py:28-53, which hands the monomer
and pair blocks to Qode's hierarchical-fluctuation CCSD: H_012_operator,
SD_operator, inv_diagE, Baker-Campbell-Hausdorff).

The reference state |0> is state 0 of every fragment (the first of each
fragment's state list) and

    T = sum_m sum_{i!=0} t_m[i] |i><0|_m + sum_{m<n} sum_{i,j!=0} t_mn[i, j] |i><0|_m |j><0|_n.

Transition operators of different fragments commute and |i><0| |i'><0| = 0
on one fragment, so exp(+-T) is the finite series sum_{k<=N} (+-T)^k / k!.
The equations are E = <0| H e^T |0> and <mu| e^-T H e^T |0> = 0 for every
single and double excitation mu, solved by Jacobi steps with the
monomer-diagonal denominators (inv_diagE) and DIIS.

Everything is evaluated in the tensor-product space with the matrix-free
XR operator (xr_operator.XROperator): the singles of T as axis GEMMs with
the matrix t_m e_0^T, the doubles as pair packs, one outer-product GEMM
t_mn x (the (0, 0) row of the packed vector), and an unpack-add, all on the
library.  For a dimer, singles and doubles span every excitation, so E is
an exact eigenvalue of the XR matrix (the lowest when the iteration starts
in its basin).
"""
import math

import numpy as np
import torch

from .gemm import mm
from .xr_operator import XROperator, _axis_apply


class Amplitudes:
    """singles[m]: (d_m,) with singles[m][0] = 0; doubles[(m, n)]: (d_m, d_n)
    with row 0 and column 0 zero (m < n)."""

    def __init__(self, dims, device):
        self.dims = list(dims)
        self.singles = [torch.zeros(d, dtype=torch.float64, device=device) for d in dims]
        self.doubles = {(m, n): torch.zeros(dims[m], dims[n], dtype=torch.float64, device=device)
                        for m in range(len(dims)) for n in range(m + 1, len(dims))}

    def flat(self):
        parts = [t[1:] for t in self.singles] + [t[1:, 1:].reshape(-1) for t in self.doubles.values()]
        return torch.cat(parts)

    def set_flat(self, x):
        o = 0
        for t in self.singles:
            k = t.numel() - 1
            t[1:] = x[o:o + k]
            o += k
        for t in self.doubles.values():
            r, c = t.shape[0] - 1, t.shape[1] - 1
            t[1:, 1:] = x[o:o + r * c].view(r, c)
            o += r * c


class XRCCSD:
    """op: the XR operator (monomer and pair terms); iterate() solves the
    CCSD equations from T = 0."""

    def __init__(self, op):
        self.op = op
        self.dims = op.dims
        self.nf = op.nf
        self.n = op.n
        self.device = op.device
        self.strides = [int(np.prod(self.dims[m + 1:])) for m in range(self.nf)]
        h0 = [torch.diagonal(h) for h in op.H1]
        self.D1 = [(d - d[0]) for d in h0]

    def apply_T(self, amp, X, sign=1.0):
        """sign * T X for a product-space vector X (n x 1)."""
        Y = torch.zeros_like(X)
        dims = self.dims
        for m, t in enumerate(amp.singles):
            A = torch.zeros(dims[m], dims[m], dtype=torch.float64, device=self.device)
            A[:, 0] = sign * t
            pre = int(np.prod(dims[:m]))
            post = int(np.prod(dims[m + 1:]))
            _axis_apply(A, X, Y, pre, dims[m], post, beta=1.0)
        for (m, n), t in amp.doubles.items():
            D = dims[m] * dims[n]
            R = self.n // D
            Pk = torch.empty(D, R, dtype=torch.float64, device=self.device)
            self.op._pack(dims, 1, m, n, None, D, X, Pk, 0)
            Z = mm(t.reshape(D, 1), Pk[0:1], alpha=sign)
            self.op._pack(dims, 1, m, n, None, D, Y, Z, 1, 1.0)
        return Y

    def exp_T(self, amp, X, sign=1.0):
        """exp(sign T) X by the finite series (T is nilpotent of order N + 1)."""
        out = X.clone()
        term = X
        for k in range(1, self.nf + 1):
            term = self.apply_T(amp, term, sign) / k
            out = out + term
        return out

    def transformed(self, amp):
        """W = e^-T H e^T |0>."""
        ref = torch.zeros(self.n, 1, dtype=torch.float64, device=self.device)
        ref[0, 0] = 1.0
        psi = self.exp_T(amp, ref)
        Hpsi = torch.empty_like(psi)
        self.op.matvec(psi.contiguous(), Hpsi)
        return self.exp_T(amp, Hpsi, -1.0)

    def residuals(self, W):
        """(E, singles residuals, doubles residuals) read from W."""
        w = W.view(-1)
        E = w[0]
        r1 = [w[torch.arange(d, device=self.device) * self.strides[m]] for m, d in enumerate(self.dims)]
        r2 = {}
        for m in range(self.nf):
            for n in range(m + 1, self.nf):
                i = torch.arange(self.dims[m], device=self.device) * self.strides[m]
                j = torch.arange(self.dims[n], device=self.device) * self.strides[n]
                r2[(m, n)] = w[i[:, None] + j[None, :]]
        return E, r1, r2

    def iterate(self, thresh=1e-10, max_iter=200, diis_start=0, diis_size=8, log=None):
        amp = Amplitudes(self.dims, self.device)
        den = Amplitudes(self.dims, self.device)
        for m, d in enumerate(self.D1):
            den.singles[m].copy_(d)
        for (m, n), t in den.doubles.items():
            t.copy_(self.D1[m][:, None] + self.D1[n][None, :])
        dvec = den.flat()
        dvec = torch.where(dvec.abs() < 1e-8, torch.full_like(dvec, 1e-8), dvec)
        hist_t, hist_r = [], []
        E = None
        for it in range(max_iter):
            W = self.transformed(amp)
            E_t, r1, r2 = self.residuals(W)
            res = Amplitudes(self.dims, self.device)
            for m, r in enumerate(r1):
                res.singles[m].copy_(r)
                res.singles[m][0] = 0.0
            for key, r in r2.items():
                res.doubles[key].copy_(r)
                res.doubles[key][0, :] = 0.0
                res.doubles[key][:, 0] = 0.0
            rv = res.flat()
            E = float(E_t)
            rmax = float(rv.abs().max()) if rv.numel() else 0.0
            if log is not None:
                log(f"XR-CCSD iteration {it}: E = {E:.12f}, max residual {rmax:.2e}")
            if rmax < thresh:
                return E, amp, it
            t_new = amp.flat() - rv / dvec
            if it >= diis_start:
                hist_t.append(t_new.clone())
                hist_r.append((rv / dvec).clone())
                if len(hist_t) > diis_size:
                    hist_t.pop(0)
                    hist_r.pop(0)
                if len(hist_t) >= 2:
                    t_new = _diis(hist_t, hist_r)
            amp.set_flat(t_new)
        raise RuntimeError(f"XR-CCSD did not converge in {max_iter} iterations (E = {E})")


def _diis(ts, rs):
    k = len(ts)
    R = torch.stack(rs)
    B = np.zeros((k + 1, k + 1))
    B[:k, :k] = (R @ R.T).cpu().numpy()
    B[k, :k] = B[:k, k] = -1.0
    rhs = np.zeros(k + 1)
    rhs[k] = -1.0
    try:
        c = np.linalg.solve(B, rhs)[:k]
    except np.linalg.LinAlgError:
        return ts[-1]
    c = torch.as_tensor(c, dtype=torch.float64, device=R.device)
    return (c[:, None] * torch.stack(ts)).sum(0)


def ccsd(H, out=None, resources=None, diis_start=0, charges=None, thresh=1e-10, max_iter=200):
    """Drop-in for general-XRCC/excitonic.py:ccsd: H = (monomer_Hamiltonians,
    dimer_Couplings) as numpy arrays or device tensors, dimer_Couplings[m][n]
    the (d_m d_n)^2 pair matrix for m < n.  Returns (E, T) with T the
    Amplitudes.  out: an object with .log(...) (optional); resources is
    accepted for the signature and unused (all work is on the GPU);
    charges: per fragment, the charge of each state (lets the operator cut
    the pair matrices into charge blocks; default one block)."""
    H1, H2 = H[0], H[1]
    dev = torch.device("cuda", 0)
    t = lambda a: a.to(dev, torch.float64) if isinstance(a, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    H1 = [t(h) for h in H1]
    nf = len(H1)
    pairs = {(m, n): t(H2[m][n]) for m in range(nf) for n in range(m + 1, nf) if H2[m][n] is not None}
    if charges is None:
        charges = [np.zeros(h.shape[0], dtype=np.int64) for h in H1]
    op = XROperator(H1, pairs, charges)
    log = out.log if out is not None and hasattr(out, "log") else None
    if log is not None:
        log("Performing coupled-cluster iterations ...")
    E, amp, _ = XRCCSD(op).iterate(thresh=thresh, max_iter=max_iter, diis_start=diis_start, log=log)
    if log is not None:
        log("CCSD Energy =", E, "Hartree")
    return E, amp
