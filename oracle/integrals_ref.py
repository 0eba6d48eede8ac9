"""CPU restatement of the orbital-loop integral transforms
(hermitian-XRCC/get_ints.py:57-155).  TEST INFRASTRUCTURE ONLY: the device
transforms of qodeapplications_amd/integrals.py are compared against these.

Conventions: one-body A[i, j] = <i|a|j>; two-body chemist V[p, q, r, s] =
(pq|rs) (electron 1: bra p, ket q; electron 2: bra r, ket s).  A
transformation (L, R) takes A -> L A R and sends both bras of V through L
and both kets through R (Qode's bra_transformed / ket_transformed).

How they are pinned (tests/test_oracle_integrals.py): the core projection
makes valence orbitals orthogonal to the core; the biorthogonal
Hamiltonian's Fock-space matrix has the spectrum of the orthonormalised
problem (brute-force FCI in both); V_half has no reference definition here
(Qode's spin_orb_integrals is absent): parity unpinned.
"""
import numpy as np


def core_projection(S, core):
    """get_ints.direct_CoreProj (:68-112): R = 1, R[core, valence] =
    -S_cc^-1 S_cv; L = R^T."""
    n = S.shape[0]
    core = list(core)
    val = [j for j in range(n) if j not in core]
    R = np.eye(n)
    if core:
        R[np.ix_(core, val)] = -np.linalg.inv(S[np.ix_(core, core)]) @ S[np.ix_(core, val)]
    return R.T.copy(), R


def transform(A, V, L, R):
    A2 = L @ A @ R
    V2 = np.einsum("pa,bq,rc,ds,abcd->pqrs", L, R, L, R, V, optimize=True)
    return A2, V2


def biorthogonal(S, A, V):
    """get_ints.py:125-126: bras through S^-1."""
    Si = np.linalg.inv(S)
    return transform(A, V, Si, np.eye(S.shape[0]))


def v_half(S, V):
    Si = np.linalg.inv(S)
    return np.einsum("pa,aqrs->pqrs", Si, V)
