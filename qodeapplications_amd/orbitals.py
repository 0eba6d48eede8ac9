"""Fragment-blocked spin-orbital integrals as differentiable functions of
intra-fragment orbital rotations (for dE/dkappa).

The spatial integrals h (and the per-nucleus potentials U) are rotated with
U = exp(K), K = blockdiag(X_m - X_m^T) over the fragments, V (chemist
notation) with U on all four indices, then cut into the spin-orbital
blocks the contractions use (V_pqrs = (1/4)(<pq|rs> - <pq|sr>), alpha
spin orbitals of a fragment first).  The rotation follows the orbital
update of StateSpaceOptimizer/orbital_solver.py:420-437 to first order
(U = 1 + kappa at kappa = 0); the gradient is with respect to the
antisymmetric generator.  Everything is torch on the device, so autograd
carries the integral-block adjoints of the XR energy back to kappa.
"""
import itertools

import numpy as np
import torch

from .gemm import mm


class _LeadingAxisProduct(torch.autograd.Function):
    """Y[i..., p] = sum_a X[a, i...] R[a, p] for X viewed as (n, rest): the
    leading axis is contracted with R and the new axis appended last, so k
    applications transform all k axes of a tensor (R^T A R for k = 2, the
    4-index transform of V for k = 4).  Forward and backward run on the
    library's FP64 GEMM (dX = R dY^T, dR = X dY)."""

    @staticmethod
    def forward(ctx, X, R):
        n = R.shape[0]
        Xm = X.reshape(n, -1).contiguous()
        ctx.save_for_backward(Xm, R)
        ctx.shape = X.shape
        Y = mm(Xm, R.contiguous(), transA=1)                 # (rest, n)
        return Y.view(*X.shape[1:], R.shape[1])

    @staticmethod
    def backward(ctx, dY):
        Xm, R = ctx.saved_tensors
        dYm = dY.reshape(-1, R.shape[1]).contiguous()
        dX = mm(R.contiguous(), dYm, transB=1).view(ctx.shape) if ctx.needs_input_grad[0] else None
        dR = mm(Xm, dYm) if ctx.needs_input_grad[1] else None
        return dX, dR


def transform(A, R):
    """A with every axis transformed by R: sum A[a, b, ...] R[a, p] R[b, q] ..."""
    Y = A
    for _ in range(A.dim()):
        Y = _LeadingAxisProduct.apply(Y, R)
    return Y


class OrbitalRotation:
    def __init__(self, frag_orbs, h_sp, V_sp, U_sp=None, device=0):
        dev = torch.device("cuda", device) if isinstance(device, int) else device
        self.device = dev
        self.frag_orbs = list(frag_orbs)
        self.off = np.concatenate([[0], np.cumsum(self.frag_orbs)]).astype(int)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        self.h = t(h_sp)
        self.V = t(V_sp)
        self.U = None if U_sp is None else t(U_sp)

    def zero_kappas(self):
        return [torch.zeros(n, n, dtype=torch.float64, device=self.device, requires_grad=True)
                for n in self.frag_orbs]

    def rotated(self, kappas, at_zero=False):
        """Rotated h, V, U.  at_zero: the caller guarantees every kappa is
        zero (the gradient point), where exp(K) = 1 + K holds exactly in
        value and first derivative, so the matrix exponential is skipped."""
        n = int(self.off[-1])
        parts = []
        for m, X in enumerate(kappas):
            parts.append(X - X.T)
        K = torch.block_diag(*parts)
        if at_zero:
            R = K + torch.eye(n, dtype=torch.float64, device=self.device)
        else:
            R = torch.linalg.matrix_exp(K)
        h = transform(self.h, R)
        V = transform(self.V, R)
        U = None if self.U is None else torch.stack([transform(u, R) for u in self.U])
        return h, V, U

    def _sl(self, m):
        return slice(int(self.off[m]), int(self.off[m + 1]))

    def one_body(self, A, m1, m2):
        n1, n2 = self.frag_orbs[m1], self.frag_orbs[m2]
        blk = A[self._sl(m1), self._sl(m2)]
        z = torch.zeros(n1, n2, dtype=torch.float64, device=self.device)
        return torch.cat([torch.cat([blk, z], 1), torch.cat([z, blk], 1)], 0)

    def two_body(self, V, ms):
        n = [self.frag_orbs[m] for m in ms]
        s = [self._sl(m) for m in ms]
        C = V[s[0], s[2], s[1], s[3]].permute(0, 2, 1, 3)
        X = V[s[0], s[3], s[1], s[2]].permute(0, 2, 3, 1)
        zero = torch.zeros(n[0], n[1], n[2], n[3], dtype=torch.float64, device=self.device)
        cube = {}
        for sp in itertools.product((0, 1), repeat=4):
            blk = zero
            if sp[0] == sp[2] and sp[1] == sp[3]:
                blk = blk + C
            if sp[0] == sp[3] and sp[1] == sp[2]:
                blk = blk - X
            cube[sp] = 0.25 * blk
        # assemble the (2n0, 2n1, 2n2, 2n3) block from the 16 spin cases
        def cat(level, prefix):
            if level == 4:
                return cube[prefix]
            return torch.cat([cat(level + 1, prefix + (b,)) for b in (0, 1)], dim=level)
        return cat(0, ())

    def blocks(self, kappas, keys, at_zero=False):
        """Differentiable blocks for IntegralCache keys ('T', m, n), ('U', a, m, n),
        ('V', m, n, o, p)."""
        h, V, U = self.rotated(kappas, at_zero)
        out = {}
        for key in keys:
            if key[0] == "T":
                out[key] = self.one_body(h, *key[1:])
            elif key[0] == "U":
                a, m1, m2 = key[1:]
                if U is None:
                    n1, n2 = self.frag_orbs[m1], self.frag_orbs[m2]
                    out[key] = torch.zeros(2 * n1, 2 * n2, dtype=torch.float64, device=self.device)
                else:
                    out[key] = self.one_body(U[a], m1, m2)
            else:
                out[key] = self.two_body(V, key[1:])
        return out


def integral_keys(nf):
    """Every integral block the monomer and pair terms of nf fragments read
    (IntegralCache keys)."""
    keys = [("T", m, n) for m in range(nf) for n in range(nf)]
    keys += [("U", a, m, n) for a in range(nf) for m in range(nf) for n in range(nf)]
    v = [(m, m, m, m) for m in range(nf)]
    for m in range(nf):
        for n in range(m + 1, nf):
            v += [(m, m, n, n), (n, n, m, m), (m, m, m, n), (m, n, n, n), (n, n, n, m), (n, m, m, m),
                  (m, n, m, n)]
    keys += [("V",) + k for k in v]
    return keys


def dimer_integral_keys():
    return integral_keys(2)


def energy_and_gradients(system, rotation, lagrangian=True):
    """E, dE/dz (per fragment and charge) and dE/dkappa (per fragment,
    antisymmetric n_m x n_m) at kappa = 0, for an XRSystem whose integrals
    are `rotation`'s."""
    from .grad import XREnergy
    kappas = rotation.zero_kappas()
    keys = integral_keys(len(system.frags))
    blocks = rotation.blocks(kappas, keys, at_zero=True)
    z_keys = [(m, c) for m, F in enumerate(system.frags) for c in F.charges]
    zs = [system.frags[m].blocks[c].coeffs.clone().requires_grad_(True) for m, c in z_keys]
    E = XREnergy.apply(system, lagrangian, z_keys, keys, *zs, *[blocks[k] for k in keys])
    gz_and_k = torch.autograd.grad(E, zs + kappas)
    gz = {k: g for k, g in zip(z_keys, gz_and_k[:len(zs)])}
    gk = list(gz_and_k[len(zs):])
    return float(E.detach()), gz, gk
