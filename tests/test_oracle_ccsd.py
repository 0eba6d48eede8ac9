"""The oracle's brute-force XR-CCSD (oracle.xr_ccsd, general-XRCC/
excitonic.py:28-53) pinned on the CPU: with no pair coupling it gives the
sum of the reference states' monomer energies; for a dimer, singles and
doubles span every excitation, so its energy is the exact lowest XR
eigenvalue of the sector of the reference state."""
import numpy as np
import pytest

import oracle.oracle as orc
from qodeapplications_amd import synthetic as syn


def xr_blocks(orbs, seed, n_states=3):
    """Oracle monomer and pair blocks of a chain of fragments, with each
    charge block rotated to the eigenbasis of its monomer block (state 0 of
    each fragment is then its lowest neutral monomer state)."""
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, [1, 2, 3], n_states, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    # weak inter-fragment coupling: the reference product state dominates
    frag_of = np.concatenate([[m] * n for m, n in enumerate(orbs)])
    local = np.concatenate([np.arange(n) for n in orbs])
    ints = syn.FragmentBlocked(list(orbs), syn.spatial_h(n_sp, rng, frag_of, 0.25),
                               syn.spatial_V(n_sp, rng, local, frag_of, 0.25))
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    H1, H2, _ = orc.compute_H_blocks(frags, ints, np.zeros((len(orbs),) * 2))
    U = []
    for m, f in enumerate(frags):
        d = H1[m].shape[0]
        u = np.zeros((d, d))
        for c in f.states:
            idx = [k for k, (cc, _) in enumerate(f.state_indices) if cc == c]
            w, v = np.linalg.eigh(H1[m][np.ix_(idx, idx)])
            u[np.ix_(idx, idx)] = v
        U.append(u)
    H1 = [U[m].T @ H1[m] @ U[m] for m in range(len(orbs))]
    for m in range(len(orbs)):
        for k in range(m + 1, len(orbs)):
            W = np.kron(U[m], U[k])
            H2[m][k] = W.T @ H2[m][k] @ W
    charges = [np.array([c for c, _ in f.state_indices]) for f in frags]
    return H1, H2, charges


def test_uncoupled_fragments_give_the_monomer_sum():
    H1, H2, _ = xr_blocks((2, 2, 2), 1)
    Z = [[None if x is None else np.zeros_like(x) for x in row] for row in H2]
    E, t1, t2 = orc.xr_ccsd(H1, Z)
    assert abs(E - sum(h[0, 0] for h in H1)) < 1e-12
    assert max(np.abs(t).max() for t in t1) < 1e-12


@pytest.mark.parametrize("seed", [3, 4])
def test_dimer_ccsd_is_the_exact_sector_ground_state(seed):
    H1, H2, ch = xr_blocks((2, 2), seed)
    E, _, _ = orc.xr_ccsd(H1, H2)
    spf = H1[0].shape[0]
    Hm = np.zeros((spf ** 2, spf ** 2))
    orc.braket_loops(Hm, 2, spf, (H1, H2, orc._NoTrimers()))
    q = np.add.outer(ch[0], ch[1]).reshape(-1)
    sel = np.nonzero(q == 0)[0]
    ev, vec = np.linalg.eigh(Hm[np.ix_(sel, sel)])
    assert abs(vec[0, 0]) > 0.7           # the reference dominates the sector ground state
    assert abs(E - ev[0]) < 1e-9, (E, ev[:3])
