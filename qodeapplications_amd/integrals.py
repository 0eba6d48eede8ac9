"""Integral transforms of the orbital loop on the device
(hermitian-XRCC/get_ints.py:62-155, StateSpaceOptimizer/orbital_solver.py:
323-437): core projection, biorthogonalisation, the half-biorthogonal V and
V_diff, the linearised orbital update, and the fragment-blocked
spin-orbital integrals the XR contractions read.

Orbitals of different fragments overlap (S != 1).  Conventions, in the
spatial fragment-MO basis of the whole supersystem:

* one-body A[i, j] = <i|a|j> (bra i, ket j); two-body V in chemist
  notation V[p, q, r, s] = (pq|rs): electron 1 has bra p and ket q,
  electron 2 bra r and ket s;
* a transformation (L, R) maps A -> L A R and V[p, q, r, s] ->
  sum L[p, p'] R[q', q] L[r, r'] R[s', s] V[p', q', r', s'] (bras through
  L, kets through R: Qode's bra_transformed / ket_transformed);
* core projection (direct_CoreProj, get_ints.py:68-112): R = 1 with
  R[core, valence] = -S_cc^-1 S_cv (valence orbitals made orthogonal to the
  core), L = R^T;
* biorthogonalisation (direct_Sinv, :57-60 and :125-126): bras through
  S^-1, kets unchanged, so the bra functions become the duals <p~| of the
  kets; the Hamiltonian H = sum h~_pq a+_p a~_q + ... with these integrals
  has, on the non-orthogonal determinant basis, a matrix whose eigenvalues
  are the energies of the orthonormalised problem (checked in
  tests/test_oracle_integrals.py);
* V_half: electron 1's bra through S^-1, electron 2's not; V_diff =
  V_half - V_bior (get_ints.py:145-146).  The reference takes V_half from
  Qode's spin_orb_integrals, absent here: its definition is this build's
  reading and is not pinned against the reference.

The transforms are chains of leading-axis products on the library's FP64
GEMM (orbitals.transform); the small inverses are host LAPACK (n <= a few
hundred orbitals).
"""
import numpy as np
import torch

from .gemm import mm


def _dev(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def _axes_transform(A, mats):
    """A[i0, i1, ...] -> sum A[j0, j1, ...] M0[j0, i0] M1[j1, i1] ...: each
    axis contracted with its matrix's rows (the leading axis moved to the
    end once per axis, as orbitals.transform does)."""
    Y = A.contiguous()
    for M in mats:
        n = M.shape[0]
        Ym = Y.reshape(n, -1)
        Y = mm(Ym, M.contiguous(), transA=1).view(*Y.shape[1:], M.shape[1])
    return Y


def transform_one_body(A, L, R):
    """L A R."""
    return mm(mm(L, A), R)


def transform_two_body(V, L, R):
    """sum L[p, p'] R[q', q] L[r, r'] R[s', s] V[p', q', r', s']."""
    return _axes_transform(V, [L.T, R, L.T, R])


def core_projection(S, core):
    """(L, R) of get_ints.direct_CoreProj for the overlap S (n x n) and the
    list of core orbital indices."""
    Sn = S.cpu().numpy() if isinstance(S, torch.Tensor) else np.asarray(S, dtype=np.float64)
    n = Sn.shape[0]
    core = list(core)
    val = [j for j in range(n) if j not in core]
    R = np.eye(n)
    if core:
        Scc = Sn[np.ix_(core, core)]
        Scv = Sn[np.ix_(core, val)]
        R[np.ix_(core, val)] = -np.linalg.solve(Scc, Scv)
    dev = S.device if isinstance(S, torch.Tensor) else torch.device("cuda", 0)
    R = _dev(R, dev)
    return R.T.contiguous(), R


def bra_inverse(S):
    """S^-1 (direct_Sinv) on the device."""
    Sn = S.cpu().numpy() if isinstance(S, torch.Tensor) else np.asarray(S, dtype=np.float64)
    dev = S.device if isinstance(S, torch.Tensor) else torch.device("cuda", 0)
    return _dev(np.linalg.inv(Sn), dev)


class SupersystemIntegrals:
    """Spatial supersystem integrals on the device: S, h (= T + sum_a U_a),
    U (per nucleus, optional) and V (chemist), with the frame_orbs of each
    fragment in order.  Methods return new objects."""

    def __init__(self, frag_orbs, S, h, V, U=None, device=0):
        dev = torch.device("cuda", device) if isinstance(device, int) else device
        self.device = dev
        self.frag_orbs = list(frag_orbs)
        self.S = _dev(S, dev)
        self.h = _dev(h, dev)
        self.V = _dev(V, dev)
        self.U = None if U is None else _dev(U, dev)

    def _new(self, S, h, V, U):
        out = SupersystemIntegrals.__new__(SupersystemIntegrals)
        out.device, out.frag_orbs = self.device, list(self.frag_orbs)
        out.S, out.h, out.V, out.U = S, h, V, U
        return out

    def transformed(self, L, R):
        """Bras through L, kets through R (S too)."""
        U = None if self.U is None else torch.stack([transform_one_body(u, L, R) for u in self.U])
        return self._new(transform_one_body(self.S, L, R), transform_one_body(self.h, L, R),
                         transform_two_body(self.V, L, R), U)

    def core_projected(self, core):
        L, R = core_projection(self.S, core)
        return self.transformed(L, R)

    def biorthogonal(self):
        """Bras through S^-1 (get_ints.py:125-126); S becomes 1."""
        Si = bra_inverse(self.S)
        I = torch.eye(self.S.shape[0], dtype=torch.float64, device=self.device)
        return self.transformed(Si, I)

    def v_half(self):
        """V with only electron 1's bra through S^-1."""
        Si = bra_inverse(self.S)
        I = torch.eye(self.S.shape[0], dtype=torch.float64, device=self.device)
        return _axes_transform(self.V, [Si.T, I, I, I])

    def v_diff(self):
        """V_half - V_bior (get_ints.py:146)."""
        return self.v_half() - self.biorthogonal().V

    def fragment_blocked(self):
        """Fragment-blocked spin-orbital integrals (T, U, V blocks in the
        convention the XR contractions read), from these spatial arrays."""
        from .synthetic import FragmentBlocked
        U = None if self.U is None else self.U.cpu().numpy()
        return FragmentBlocked(self.frag_orbs, self.h.cpu().numpy(), self.V.cpu().numpy(), U)


def linear_orbital_update(C, x, s=1.0):
    """The orbital update of orbital_solver.py:431: U = 1 + s x applied to
    the MO coefficients (columns), C' = C U."""
    C = C if isinstance(C, torch.Tensor) else torch.as_tensor(np.asarray(C, dtype=np.float64), device="cuda")
    x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, dtype=np.float64), device=C.device)
    U = torch.eye(x.shape[0], dtype=torch.float64, device=C.device) + s * x
    return mm(C.contiguous(), U.contiguous())
