"""The oracle's restatement of the orbital-loop integral transforms
(oracle/integrals_ref.py, hermitian-XRCC/get_ints.py:57-155) pinned on the
CPU by what the mathematics fixes:

  * the core projection leaves the core orbitals and makes every valence
    orbital orthogonal to them (projected overlap's core-valence block 0);
  * the biorthogonalised integrals (bras through S^-1) define, on the
    occupation vectors of the non-orthogonal orbitals, a Fock-space matrix
    whose spectrum is that of the Hamiltonian in an orthonormal basis
    (brute-force FCI of both, all electron counts).
"""
import numpy as np
import pytest

from oracle import brute
from oracle import integrals_ref as ir
from qodeapplications_amd import synthetic as syn


def _problem(seed, n=4, eps=0.3):
    rng = np.random.default_rng(seed)
    h_o = syn.spatial_h(n, rng)
    V_o = syn.spatial_V(n, rng)
    C = np.eye(n) + eps * rng.standard_normal((n, n))        # non-orthogonal orbitals
    S = C.T @ C
    h = C.T @ h_o @ C
    V = np.einsum("ap,bq,cr,ds,abcd->pqrs", C, C, C, C, V_o, optimize=True)
    return S, h, V, h_o, V_o


def _spin(h, V):
    """Spin-orbital h and V (the XR convention) of one 'fragment' holding
    every orbital."""
    n = h.shape[0]
    ints = syn.FragmentBlocked([n], h, V)
    return ints.T[0, 0], ints.V[0, 0, 0, 0]


@pytest.mark.parametrize("seed", [1, 2])
def test_biorthogonal_spectrum_equals_orthonormal(seed):
    S, h, V, h_o, V_o = _problem(seed)
    hb, Vb = ir.biorthogonal(S, h, V)
    # one electron's check: S^-1 h has the eigenvalues of h_o
    assert np.allclose(np.sort(np.linalg.eigvals(hb).real), np.linalg.eigvalsh(h_o), atol=1e-10)
    ho_so, Vo_so = _spin(h_o, V_o)
    hb_so, Vb_so = _spin(hb, Vb)
    ref = np.sort(np.linalg.eigvalsh(brute.hamiltonian(ho_so, Vo_so)))
    ev = np.linalg.eigvals(brute.hamiltonian(hb_so, Vb_so))
    assert float(np.abs(ev.imag).max()) < 1e-8
    np.testing.assert_allclose(np.sort(ev.real), ref, atol=1e-9)
    # the plain (non-biorthogonal) MO integrals do not give it
    hm_so, Vm_so = _spin(h, V)
    bad = np.sort(np.linalg.eigvals(brute.hamiltonian(hm_so, Vm_so)).real)
    assert float(np.abs(bad - ref).max()) > 1e-3


def test_core_projection_orthogonalises_valence_to_core():
    S, h, V, _, _ = _problem(3, n=5)
    core = [0, 3]
    L, R = ir.core_projection(S, core)
    Sp = L @ S @ R
    val = [1, 2, 4]
    assert np.abs(Sp[np.ix_(core, val)]).max() < 1e-12
    assert np.abs(Sp[np.ix_(val, core)]).max() < 1e-12
    assert np.allclose(Sp[np.ix_(core, core)], S[np.ix_(core, core)])
    assert np.allclose(R[:, core], np.eye(5)[:, core])            # core orbitals unchanged
    assert abs(np.linalg.det(R) - 1.0) < 1e-12                     # same span
