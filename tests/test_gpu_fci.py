"""Fragment FCI (fci.py): the sigma H psi from the factorised density
adjoints against the brute-force Fock-space Hamiltonian restricted to the
charge block's determinants (oracle/brute.py), with and without a frozen
core; the determinant diagonal; the lowest states by Davidson against the
dense spectrum; and states made this way through the XR dimer (the upstream
step of general-XRCC: atomic_states/LMO_frozen_Hamiltonian.c compute_HPsi)."""
import numpy as np
import pytest
import torch

from oracle import brute
from qodeapplications_amd import fci
from qodeapplications_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _ints(n_orbs, seed):
    rng = np.random.default_rng(seed)
    ints = syn.FragmentBlocked([n_orbs], syn.spatial_h(n_orbs, rng), syn.spatial_V(n_orbs, rng))
    return ints.T[0, 0], ints.V[0, 0, 0, 0]


@pytest.mark.parametrize("n_orbs,n_elec,n_core", [(3, 2, 0), (3, 3, 0), (4, 3, 1), (4, 5, 1), (5, 1, 0)])
def test_sigma_equals_brute_force(n_orbs, n_elec, n_core):
    h, V = _ints(n_orbs, 10 * n_orbs + n_elec)
    H = fci.Hamiltonian(n_orbs, n_elec, h, V, n_core)
    Hf = brute.hamiltonian(h, V)
    bits = [brute.config_bits(c) for c in H.configs]
    Hs = Hf[np.ix_(bits, bits)]
    rng = np.random.default_rng(1)
    Z = rng.standard_normal((5, H.n_configs))
    got = H.sigma(torch.as_tensor(Z, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(got, Z @ Hs.T, atol=1e-12)
    np.testing.assert_allclose(H.diagonal().cpu().numpy(), np.diag(Hs), atol=1e-12)
    w, z = H.lowest(k=3)
    np.testing.assert_allclose(w.cpu().numpy(), np.linalg.eigvalsh(Hs)[:3], atol=1e-10)


def test_davidson_states_of_a_larger_block():
    n_orbs, n_elec = 9, 4                   # 3060 determinants: the iterative path
    h, V = _ints(n_orbs, 99)
    H = fci.Hamiltonian(n_orbs, n_elec, h, V)
    assert H.n_configs == 3060
    w, z = H.lowest(k=4, tol=1e-10)
    # dense reference from the sigma of the identity, in chunks
    E = torch.eye(H.n_configs, dtype=torch.float64, device="cuda")
    Hd = H.sigma(E)
    assert float((Hd - Hd.T).abs().max()) < 1e-12
    ref = torch.linalg.eigvalsh(Hd)[:4]
    assert float((w - ref).abs().max()) < 1e-9
    r = H.sigma(z) - w[:, None] * z
    assert float(r.norm(dim=1).max()) < 1e-8 * max(1.0, float(w.abs().max()))
    assert float((z @ z.T - torch.eye(4, dtype=torch.float64, device="cuda")).abs().max()) < 1e-10


def test_fci_states_feed_the_xr_dimer():
    """Fragment states from fragment_states go through the XR dimer; with
    the complete state spaces the XR spectrum is the dimer's exact FCI."""
    import oracle.oracle as orc
    from qodeapplications_amd.xr import XRSystem
    rng = np.random.default_rng(5)
    orbs = [2, 2]
    h = syn.spatial_h(4, rng)
    Vsp = syn.spatial_V(4, rng)
    ints = syn.FragmentBlocked(orbs, h, Vsp)
    frags = []
    for m, n in enumerate(orbs):
        st = fci.fragment_states(n, ints.T[m, m], ints.V[m, m, m, m], range(0, 2 * n + 1), 16)
        states = {2 - ne: syn.ChargeStates(configs=cfg, coeffs=z.cpu().numpy()) for ne, (w, z, cfg) in st.items()}
        frags.append(syn.Fragment(n_orbs=n, n_core=0, n_elec_ref=2, states=states))
    X = XRSystem(frags, ints, np.zeros((2, 2)))
    M = X.xr_matrix(spectator_sign=False)
    xr = torch.linalg.eigvalsh(0.5 * (M + M.T)).cpu().numpy()
    hf, Vf = ints.full_spin_orbital()
    exact = np.sort(np.linalg.eigvalsh(brute.hamiltonian(hf, Vf)))
    np.testing.assert_allclose(xr, exact, atol=1e-10)
