"""The oracle's restatement of the reference gradient definitions
(oracle/refgrad.py: StateSpaceOptimizer/state_gradients.py:121-244 at the
"herm" level, orb_grads.py:867-999) pinned on tiny dimers, on the CPU:

  * the XR one- and two-particle densities assembled from fragment
    densities per the orb_grads catalogs equal the brute-force Fock-space
    expectation values of the embedded XR state, and give back E;
  * the "herm" state gradient is half the derivative of the variational
    XR energy: no component inside each charge block's state span, and
    half the central finite difference of the XR eigenvalue along
    directions orthogonal to that span;
  * the orbital gradient, spin-summed over each fragment's diagonal block,
    is minus the central finite difference of the XR energy under the
    rotation exp(K), K = blockdiag(X_m - X_m^T), of the spatial orbitals.
"""
import numpy as np
import pytest

import oracle.oracle as orc
from oracle import refgrad as rg
from qodeapplications_amd import synthetic as syn


def _dimer(seed, orbs=(2, 2), n_states=2, U=False):
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, [1, 2, 3], n_states, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    Usp = np.stack([0.3 * syn.spatial_h(n_sp, rng) for _ in orbs]) if U else None
    return frags, syn.FragmentBlocked(list(orbs), h, V, Usp), np.zeros((2, 2))


def _xr_energy(frags, ints, nuc):
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    E, d, _ = rg.dimer_root(frags, ints, nuc)
    return E, d


def _generalised_energy(frags, ints, nuc):
    """Lowest eigenvalue of H c = E S c over the product states of the
    (possibly non-orthonormal) fragment states."""
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    H1, H2, _ = orc.compute_H_blocks(frags, ints, nuc)
    _, H = orc.fci((H1, H2, None))
    S = None
    for f in frags:
        si = f.state_indices
        Sm = np.zeros((len(si), len(si)))
        for a, (ca, ia) in enumerate(si):
            for b, (cb, ib) in enumerate(si):
                if ca == cb:
                    Sm[a, b] = f.states[ca].coeffs[ia] @ f.states[cb].coeffs[ib]
        S = Sm if S is None else np.kron(S, Sm)
    L = np.linalg.cholesky(S)
    Li = np.linalg.inv(L)
    return float(np.linalg.eigvalsh(Li @ (0.5 * (H + H.T)) @ Li.T)[0])


def _ground_in_neutral_sector(frags, d):
    ch = [[c for c in f.states for _ in range(f.states[c].coeffs.shape[0])] for f in frags]
    q = np.add.outer(np.array(ch[0]), np.array(ch[1]))
    return float(np.abs(d[q != 0]).max()) < 1e-12


@pytest.mark.parametrize("seed,orbs", [(0, (2, 2)), (2, (2, 3))])
def test_xr_densities_match_brute_force(seed, orbs):
    frags, ints, nuc = _dimer(seed, orbs)
    E, d = _xr_energy(frags, ints, nuc)
    assert _ground_in_neutral_sector(frags, d)
    n_occ = sum(f.n_elec_ref for f in frags)
    gamma, Gamma = rg.xr_densities(frags, d, d, n_occ)
    dim = sum(2 * f.n_orbs for f in frags)
    vec = rg.embed_dimer(frags, d)
    assert abs(float(vec @ vec) - 1.0) < 1e-12
    gb, Gb = rg.brute_densities(vec, dim)
    np.testing.assert_allclose(gamma, gb, atol=1e-12)
    np.testing.assert_allclose(Gamma, Gb, atol=1e-12)
    h, V = ints.full_spin_orbital()
    assert abs(rg.rdm_energy(gamma, Gamma, h, V) - E) < 1e-11


@pytest.mark.parametrize("seed,frag_ind", [(0, 0), (0, 1), (5, 0)])
def test_herm_state_gradient_is_half_the_variational_derivative(seed, frag_ind):
    """The "herm" gradient is half the derivative of the variational XR
    energy, in which every product state carries its overlaps: it has no
    component inside the span of each charge block's states (that energy
    is invariant to linear transformations within the span: g z^T = 0), and
    along directions orthogonal to the span -- where the XR matrix as
    assembled (identity on the spectator fragment) needs no overlap
    factors -- it is half the central difference of the XR eigenvalue."""
    frags, ints, nuc = _dimer(seed)
    E, g, d = rg.herm_state_gradient(frags, ints, nuc, frag_ind=frag_ind)
    assert _ground_in_neutral_sector(frags, d if frag_ind == 0 else d.T)
    f = frags[frag_ind]
    rng = np.random.default_rng(100 + seed)
    base = {c: f.states[c].coeffs.copy() for c in f.states}
    for c, z in base.items():
        assert float(np.abs(g[c] @ z.T).max()) < 1e-12 * max(1.0, float(np.abs(g[c]).max()))
        assert float(np.abs(g[c]).max()) > 1e-3
    dirs = {}
    for c, z in base.items():
        u = rng.standard_normal(z.shape)
        dirs[c] = u - (u @ z.T) @ z            # orthogonal to the span of the rows of z
    eps = 1e-5

    def e_at(t):
        for c in f.states:
            f.states[c].coeffs = base[c] + t * dirs[c]
        return _generalised_energy(frags, ints, nuc)

    fd = (e_at(eps) - e_at(-eps)) / (2 * eps)
    for c in f.states:
        f.states[c].coeffs = base[c]
    an = sum(float(np.sum(g[c] * dirs[c])) for c in f.states)
    assert abs(fd - 2.0 * an) < 1e-7 * max(1.0, abs(fd)), (fd, 2 * an)


def _rotated_energy(frags, ints, Xs):
    from scipy.linalg import expm
    n = sum(ints.frag_orbs)
    K = np.zeros((n, n))
    off = 0
    for X in Xs:
        k = X.shape[0]
        K[off:off + k, off:off + k] = X - X.T
        off += k
    R = expm(K)
    h = R.T @ ints.h_sp @ R
    V = np.einsum("abcd,ap,bq,cr,ds->pqrs", ints.V_sp, R, R, R, R, optimize=True)
    rot = syn.FragmentBlocked(ints.frag_orbs, h, V, None)
    return _xr_energy(frags, rot, np.zeros((2, 2)))[0]


@pytest.mark.parametrize("seed,orbs", [(5, (2, 2)), (2, (2, 3))])
def test_orbital_gradient_is_minus_the_finite_difference(seed, orbs):
    frags, ints, nuc = _dimer(seed, orbs)
    E, d = _xr_energy(frags, ints, nuc)
    gamma, Gamma = rg.xr_densities(frags, d, d, sum(f.n_elec_ref for f in frags))
    h, V = ints.full_spin_orbital()
    g = rg.orbital_gradient(gamma, Gamma, h, V)
    np.testing.assert_allclose(g, -g.T, atol=1e-14)
    blocks = rg.spin_summed_fragment_blocks(g, orbs)
    rng = np.random.default_rng(7 + seed)
    dirs = [rng.standard_normal((n, n)) for n in orbs]
    eps = 1e-5
    fd = (_rotated_energy(frags, ints, [eps * x for x in dirs])
          - _rotated_energy(frags, ints, [-eps * x for x in dirs])) / (2 * eps)
    an = -sum(float(np.sum(b * x)) for b, x in zip(blocks, dirs))
    assert abs(fd - an) < 1e-7 * max(1.0, abs(fd)), (fd, an)


def test_xr_densities_refuse_a_charged_sector():
    """The catalog sign rule holds on the total-charge-0 sector only."""
    frags, ints, nuc = _dimer(3, (2, 2))
    E, d = _xr_energy(frags, ints, nuc)
    assert not _ground_in_neutral_sector(frags, d)
    with pytest.raises(ValueError):
        rg.xr_densities(frags, d, d, 4)
