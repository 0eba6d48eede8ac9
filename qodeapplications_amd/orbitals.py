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

    def blocks(self, kappas, keys, at_zero=False):
        """Differentiable blocks for IntegralCache keys ('T', m, n), ('U', a, m, n),
        ('V', m, n, o, p), cut from the rotated spatial integrals by the
        library (xr_spin_blocks; the adjoint xr_spin_blocks_adjoint carries
        the block gradients back to h, V and U)."""
        h, V, U = self.rotated(kappas, at_zero)
        out = {}
        lib = []
        for key in keys:
            if key[0] == "U" and U is None:
                n1, n2 = self.frag_orbs[key[2]], self.frag_orbs[key[3]]
                out[key] = torch.zeros(2 * n1, 2 * n2, dtype=torch.float64, device=self.device)
            else:
                lib.append(key)
        if lib:
            desc = _spin_desc(lib, self.frag_orbs, self.off)
            res = _SpinBlocks.apply(desc, h, V, U if U is not None else h.new_zeros(1, 1, 1))
            out.update(zip(lib, res))
        return out


def _spin_desc(keys, frag_orbs, off):
    """Launch arguments of xr_spin_blocks for the keys: kinds, dims, offsets,
    the source of each block (-1 = h, -2 = V, a >= 0 = U[a]) and its shape."""
    k = len(keys)
    kinds = np.zeros(k, dtype=np.int32)
    dims = np.zeros((k, 4), dtype=np.int32)
    offs = np.zeros((k, 4), dtype=np.int32)
    srcs, shapes = [], []
    for i, key in enumerate(keys):
        ms = key[1:] if key[0] in ("T", "V") else key[2:]
        kinds[i] = 2 if key[0] == "V" else 1
        for a, m in enumerate(ms):
            dims[i, a] = frag_orbs[m]
            offs[i, a] = off[m]
        srcs.append(-2 if key[0] == "V" else (-1 if key[0] == "T" else key[1]))
        shapes.append(tuple(2 * frag_orbs[m] for m in ms))
    return {"k": k, "kinds": kinds, "dims": dims.reshape(-1), "offs": offs.reshape(-1), "srcs": srcs,
            "shapes": shapes}


class _SpinBlocks(torch.autograd.Function):
    """Spin-orbital blocks of (h, V, U) for a key list (forward one library
    call, backward one library call accumulating into dh, dV, dU)."""

    @staticmethod
    def forward(ctx, desc, h, V, U):
        from ._lib import context
        dev = h.device
        outs = [torch.empty(sh, dtype=torch.float64, device=dev) for sh in desc["shapes"]]
        # the pointer arrays stay referenced through the call
        sp = np.asarray([(V if s == -2 else h if s == -1 else U[s]).data_ptr() for s in desc["srcs"]],
                        dtype=np.uint64)
        op = np.asarray([o.data_ptr() for o in outs], dtype=np.uint64)
        c = context(dev.index)
        c.call("xr_spin_blocks", desc["k"], _np_ptr(desc["kinds"]), _np_ptr(desc["dims"]), _np_ptr(desc["offs"]),
               _np_ptr(sp), _np_ptr(op), h.shape[0])
        ctx.desc = desc
        ctx.shapes = (h.shape, V.shape, U.shape)
        ctx.dev = dev
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        from ._lib import context
        desc = ctx.desc
        dh = torch.zeros(ctx.shapes[0], dtype=torch.float64, device=ctx.dev)
        dV = torch.zeros(ctx.shapes[1], dtype=torch.float64, device=ctx.dev)
        dU = torch.zeros(ctx.shapes[2], dtype=torch.float64, device=ctx.dev)
        gs = [None if g is None else g.contiguous() for g in grads]
        gp = np.asarray([0 if g is None else g.data_ptr() for g in gs], dtype=np.uint64)
        dp = np.asarray([(dV if s == -2 else dh if s == -1 else dU[s]).data_ptr() for s in desc["srcs"]],
                        dtype=np.uint64)
        c = context(ctx.dev.index)
        c.call("xr_spin_blocks_adjoint", desc["k"], _np_ptr(desc["kinds"]), _np_ptr(desc["dims"]),
               _np_ptr(desc["offs"]), _np_ptr(gp), _np_ptr(dp), ctx.shapes[0][0])
        return None, dh, dV, dU


def _np_ptr(a):
    import ctypes
    return ctypes.c_void_p(a.ctypes.data)


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
