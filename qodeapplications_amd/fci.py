"""Fragment FCI in the determinant space of one charge block: the sigma
vector H psi and the lowest fragment states, the upstream step that
produces the CI vectors the density build consumes
(general-XRCC/atomic_states/LMO_frozen_Hamiltonian.c:358 compute_HPsi,
driven by LMO_frozen_Hamiltonian_wrapper.py through a Lanczos solver).

H = sum_pq h_pq p+ q + sum_pqrs V_pqrs p+ q+ s r over the fragment's spin
orbitals (alpha first), the convention of the XR contractions
(general-XRCC/H_contractions.c: monomer_2e contracts V[p,q,r,s] with
<p+ q+ s r>), i.e. V = (1/4)(<pq|rs> - <pq|sr>) for the physicist's
integrals <pq|rs>.  The reference's compute_HPsi takes <pq|rs> itself and
antisymmetrises inside (:380-440); antisymmetrize() converts.

The sigma runs the factorised density kernels backwards: with the state
matrix Z on both sides and Ebar = 1, the adjoint of the one-body density
contracted with h and of the monomer sigma with V gives
dE/dz_bra[k, P] = sum_Q <P| H |Q> Z[k, Q], i.e. (H Z^T)^T, with the same
gathers and GEMMs as the forward terms (no determinant-pair loop).
"""
import numpy as np
import torch

from ._lib import context, ptr
from .density import ChargeBlock
from .eigen import davidson, upload


def antisymmetrize(g):
    """V = (1/4)(g[p,q,r,s] - g[p,q,s,r]) from physicist's integrals g = <pq|rs>."""
    g = np.asarray(g, dtype=np.float64)
    return 0.25 * (g - g.transpose(0, 1, 3, 2))


class Hamiltonian:
    """H of one fragment for one electron count: h (dim x dim) and V (dim^4)
    over the spin orbitals, the determinants of n_elec electrons with the
    n_core lowest spatial orbitals of each spin filled (lexical order)."""

    def __init__(self, n_orbs, n_elec, h, V, n_core=0, device=0):
        from .formats import lex_configs
        self.n_orbs, self.n_elec, self.n_core = int(n_orbs), int(n_elec), int(n_core)
        self.configs = lex_configs(self.n_orbs, self.n_elec, self.n_core)
        dev = torch.device("cuda", device) if isinstance(device, int) else device
        self.device = dev
        self.h = torch.as_tensor(np.ascontiguousarray(h, dtype=np.float64), device=dev)
        self.V = torch.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dev)
        dim = 2 * self.n_orbs
        if tuple(self.h.shape) != (dim, dim) or tuple(self.V.shape) != (dim,) * 4:
            raise ValueError(f"h must be {dim} x {dim} and V {dim}^4")
        self.n_configs = self.configs.shape[0]
        self._block = ChargeBlock(self.n_orbs, self.n_core, self.configs,
                                  np.zeros((1, self.n_configs)), device=dev)
        self._diag = None

    def sigma(self, Z, chunk=64):
        """(H Z^T)^T for Z (n_vec x n_configs) on the device, `chunk` vectors
        per pass (the one-body adjoint's delta_IJ h weight is chunk^2 dim^2)."""
        Z = Z.contiguous()
        if Z.shape[0] <= chunk:
            return self._sigma(Z)
        return torch.cat([self._sigma(Z[i:i + chunk]) for i in range(0, Z.shape[0], chunk)])

    def _sigma(self, Z):
        Z = Z.contiguous()
        nv = Z.shape[0]
        out = torch.zeros_like(Z)
        if nv == 0 or self.n_elec == 0:      # H annihilates the vacuum
            return out
        ctx = context(self.device.index)
        dim = 2 * self.n_orbs
        b = self._block
        # one-body: Rbar[I, J, p, q] = delta_IJ h[p, q]
        Rb = torch.zeros(nv, nv, dim, dim, dtype=torch.float64, device=self.device)
        Rb.view(nv * nv, dim * dim)[::nv + 1] = self.h.reshape(1, -1)
        ctx.call("xr_density_adjoint", b"ca", self.n_orbs, self.n_core, self.n_elec, nv, self.n_configs, ptr(Z),
                 nv, self.n_configs, ptr(Z), ptr(Rb), ptr(out), None)
        if self.n_elec >= 2:
            Eb = torch.eye(nv, dtype=torch.float64, device=self.device)
            ctx.call("xr_monomer_2e_adjoint", self.n_orbs, self.n_core, self.n_elec, nv, self.n_configs, ptr(Z),
                     nv, self.n_configs, ptr(Z), ptr(self.V), ptr(Eb), ptr(out), None, None)
        return out

    def diagonal(self):
        """<P| H |P> of every determinant."""
        if self._diag is None:
            h = self.h.cpu().numpy()
            V = self.V.cpu().numpy()
            occ = self.configs
            d = h[occ, occ].sum(axis=1)
            if self.n_elec >= 2:
                # <P| p+ q+ s r |P> for p != q occupied: p+ q+ q p = n_p n_q (r = p,
                # s = q: V[p, q, p, q]) and p+ q+ p q = -n_p n_q (V[p, q, q, p])
                J = np.einsum("pqpq->pq", V) - np.einsum("pqqp->pq", V)
                np.fill_diagonal(J, 0.0)
                d = d + J[occ[:, :, None], occ[:, None, :]].sum(axis=(1, 2))
            self._diag = upload(d, self.device)
        return self._diag

    def lowest(self, k=1, tol=1e-9, max_iter=400):
        """(energies[k], states (k x n_configs)) of the lowest roots: Davidson
        with the diagonal preconditioner on the sigma above."""
        n = self.n_configs
        if n <= 2048:
            E = torch.eye(n, dtype=torch.float64, device=self.device)
            H = self.sigma(E)
            w, v = torch.linalg.eigh(0.5 * (H + H.T))
            return w[:k], v[:, :k].T.contiguous()

        def matmul(X, Y):
            Y.copy_(self.sigma(X.T.contiguous()).T)

        w, v = davidson(matmul, self.diagonal(), n, k=k, tol=tol, max_iter=max_iter, device=self.device)
        return w, v.T.contiguous()


def fragment_states(n_orbs, h, V, elec_counts, n_states, n_core=0, device=0, tol=1e-9):
    """{n_elec: (energies, coeffs (n_states x n_configs), configs)} of the
    lowest FCI states for each electron count: the fragment states the
    density build takes (general-XRCC/atomic_states/monomer_states.py)."""
    out = {}
    for ne in elec_counts:
        H = Hamiltonian(n_orbs, ne, h, V, n_core, device)
        k = min(int(n_states[ne] if isinstance(n_states, dict) else n_states), H.n_configs)
        w, z = H.lowest(k=k, tol=tol)
        out[ne] = (w, z, H.configs)
    return out
