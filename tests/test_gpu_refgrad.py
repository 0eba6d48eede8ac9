"""The adjoint's gradients against the reference's own gradient definitions,
restated on the CPU in oracle/refgrad.py (pinned there by brute force and
finite differences, tests/test_oracle_grad.py), on small dimers whose
ground state lies in the total-charge-0 sector (the reference's
assumption, StateSpaceOptimizer/state_gradients.py:30-32).

Mapping (also in grad.herm_state_gradient and orbitals.energy_and_gradients):
  * state gradient, "herm" level (state_gradients.py:121-244):
        g_ref[c] = (1/2) dE/dz[c] (1 - z[c]^T z[c])
    for each fragment and charge c (factor 1/2: the reference's bra-side
    half; the projector removes the in-span part, where the variational
    energy the reference differentiates is flat and the assembled XR matrix
    is not; the -2 E D z term of lagrangian=True lies in the span);
  * orbital gradient (orb_grads.py:867-999, over spin orbitals):
        dE/dkappa_m = - (g_ref[m alpha, m alpha] + g_ref[m beta, m beta])
    for the rotation exp(K), K = blockdiag(X_m - X_m^T), applied to both
    spins of fragment m (sign: orb_grads returns grad - grad^T of the
    t/u/v diagrams, the negative of dE/dX; spin sum: one spatial rotation
    moves both spin blocks).
"""
import numpy as np
import pytest
import torch

import oracle.oracle as orc
from oracle import refgrad as rg
from qodeapplications_amd import synthetic as syn
from qodeapplications_amd.grad import energy_and_gradient, herm_state_gradient
from qodeapplications_amd.orbitals import OrbitalRotation, energy_and_gradients
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu


def _dimer(seed, orbs=(2, 2), n_states=2):
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, [1, 2, 3], n_states, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    return frags, syn.FragmentBlocked(list(orbs), h, V, None), np.zeros((2, 2))


@pytest.mark.parametrize("seed,orbs", [(0, (2, 2)), (5, (2, 2)), (2, (2, 3))])
@pytest.mark.parametrize("lagrangian", [False, True])
def test_state_gradient_equals_reference_herm_definition(seed, orbs, lagrangian):
    frags, ints, nuc = _dimer(seed, orbs)
    X = XRSystem(frags, ints, nuc)
    E, v, zbar, _ = energy_and_gradient(X, lagrangian=lagrangian)
    g = herm_state_gradient(zbar, X)
    for m in (0, 1):
        Er, gr, d = rg.herm_state_gradient(frags, ints, nuc, frag_ind=m)
        assert abs(E - Er) < 1e-12
        for c in frags[m].states:
            ref = gr[c]
            got = g[(m, c)].cpu().numpy()
            scale = max(1.0, float(np.abs(ref).max()))
            assert float(np.abs(got - ref).max()) <= 1e-11 * scale, (m, c)


@pytest.mark.parametrize("seed,orbs", [(5, (2, 2)), (2, (2, 3))])
def test_orbital_gradient_equals_reference_definition(seed, orbs):
    frags, ints, nuc = _dimer(seed, orbs)
    X = XRSystem(frags, ints, nuc)
    rot = OrbitalRotation(ints.frag_orbs, ints.h_sp, ints.V_sp, None)
    E, gz, gk = energy_and_gradients(X, rot, lagrangian=True)
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    Er, d, _ = rg.dimer_root(frags, ints, nuc)
    assert abs(E - Er) < 1e-12
    gamma, Gamma = rg.xr_densities(frags, d, d, sum(f.n_elec_ref for f in frags))
    h, V = ints.full_spin_orbital()
    blocks = rg.spin_summed_fragment_blocks(rg.orbital_gradient(gamma, Gamma, h, V), list(orbs))
    for m in range(2):
        got = gk[m].detach().cpu().numpy()
        ref = -blocks[m]
        assert float(np.abs(got - ref).max()) <= 1e-11 * max(1.0, float(np.abs(ref).max())), m
