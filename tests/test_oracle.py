"""Pins the CPU oracle (oracle/) before anything is compared against it.

* lexical rank == position in itertools.combinations (closed form)
* densities == brute-force Fock-space operator matrices (oracle/brute.py,
  an independent bit-formula implementation of the sign convention)
* density invariants (traces, partial traces, adjoint symmetry)
* excitation lists rebuild the densities and keep the reference loop order
* the whole chain build_density_tensors -> build_matrix_elements ->
  braket_loops -> fci equals exact FCI for complete fragment state spaces
  (dimers and trimers, with U and nuclear terms)
"""
import itertools
import math

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as orc
from qodeapplications_amd import synthetic as syn

ALL_OPS = ("a", "c", "ca", "aa", "cc", "caa", "cca", "ccaa", "caaa", "ccca")
FIVE_OPS = ("ccaaa", "cccaa")


def test_combo_table_is_binomials():
    t = orc.combo_table(5, 18)
    for i in range(5):
        for j in range(18):
            assert t[i, j] == (math.comb(j, i) if j >= i else 0)


@pytest.mark.parametrize("n_orbs,n_core,n_elec", [(3, 0, 2), (4, 0, 3), (4, 1, 4), (5, 1, 5), (6, 0, 0), (6, 2, 5)])
def test_rank_is_lexical_position(n_orbs, n_core, n_elec):
    cfgs = orc.lex_configs(n_orbs, n_elec, n_core)
    assert cfgs.shape[0] == math.comb(2 * (n_orbs - n_core), n_elec - 2 * n_core)
    for i, c in enumerate(cfgs):
        assert orc.config_rank(c, n_elec, n_orbs, n_core) == i
    # the package's generator enumerates the same list
    assert np.array_equal(cfgs, syn.lex_configs(n_orbs, n_elec, n_core))


def _states(rng, n_orbs, n_elec, n_states, n_core):
    return syn.random_states(n_orbs, n_elec, n_states, rng, n_core)


@pytest.mark.parametrize("op", ALL_OPS)
@pytest.mark.parametrize("n_orbs,n_core,n_ket", [(3, 0, 3), (4, 1, 4), (4, 0, 2)])
def test_density_matches_bruteforce(op, n_orbs, n_core, n_ket):
    dim = 2 * n_orbs
    n_bra = n_ket + orc.n_elec_change(op)
    if n_bra < 2 * n_core or n_bra > dim:
        pytest.skip("empty bra space")
    rng = np.random.default_rng(7)
    sk = _states(rng, n_orbs, n_ket, 3, n_core)
    sb = _states(rng, n_orbs, n_bra, 2, n_core)
    got = orc.density(op, n_orbs, n_core, sb.coeffs, sk.coeffs, sk.configs)
    cache = brute.operators(dim)
    zb = brute.embed(sb.coeffs, sb.configs, dim)
    zk = brute.embed(sk.coeffs, sk.configs, dim)
    ref = brute.density(op, zb, zk, dim, cache)
    # with a frozen core the reference drops results with a core hole
    # (density_tensors.c:125-134); the brute-force bra vectors live on
    # full-core determinants only, which projects the same way.
    np.testing.assert_allclose(got, ref, atol=1e-13, rtol=0)


@pytest.mark.parametrize("op", ("ca", "ccaa", "caaa", "ccca"))
def test_fast_bruteforce_equals_matrix_products(op):
    """brute.density_fast (index maps) against brute.density (explicit
    operator matrices) where both are cheap."""
    rng = np.random.default_rng(2)
    n_orbs, n_ket = 3, 3
    dim = 2 * n_orbs
    n_bra = n_ket + orc.n_elec_change(op)
    sk = _states(rng, n_orbs, n_ket, 3, 0)
    sb = _states(rng, n_orbs, n_bra, 2, 0)
    zb = brute.embed(sb.coeffs, sb.configs, dim)
    zk = brute.embed(sk.coeffs, sk.configs, dim)
    np.testing.assert_allclose(brute.density_fast(op, zb, zk, dim),
                               brute.density(op, zb, zk, dim, brute.operators(dim)), atol=1e-13, rtol=0)


@pytest.mark.parametrize("op", FIVE_OPS)
@pytest.mark.parametrize("n_orbs,n_core,n_ket", [(3, 0, 3), (4, 1, 4), (4, 0, 3), (4, 0, 5)])
def test_five_operator_density_matches_bruteforce(op, n_orbs, n_core, n_ket):
    """The XR order-1 strings (deprecated density_tensors.c:690-836) through
    the oracle's generic walker, against the index-map brute force."""
    dim = 2 * n_orbs
    n_bra = n_ket + orc.n_elec_change(op)
    if n_bra < 2 * n_core or n_bra > dim:
        pytest.skip("empty bra space")
    rng = np.random.default_rng(11)
    sk = _states(rng, n_orbs, n_ket, 3, n_core)
    sb = _states(rng, n_orbs, n_bra, 2, n_core)
    got = orc.density(op, n_orbs, n_core, sb.coeffs, sk.coeffs, sk.configs)
    zb = brute.embed(sb.coeffs, sb.configs, dim)
    zk = brute.embed(sk.coeffs, sk.configs, dim)
    ref = brute.density_fast(op, zb, zk, dim)
    assert float(np.abs(ref).max()) > 1e-3
    np.testing.assert_allclose(got, ref, atol=1e-13, rtol=0)


def test_density_invariants():
    rng = np.random.default_rng(3)
    n_orbs, N = 6, 4
    s = _states(rng, n_orbs, N, 5, 0)
    sm = _states(rng, n_orbs, N - 1, 4, 0)
    ca = orc.density("ca", n_orbs, 0, s.coeffs, s.coeffs, s.configs)
    ccaa = orc.density("ccaa", n_orbs, 0, s.coeffs, s.coeffs, s.configs)
    # trace of the one-particle density is N * delta_IJ (orthonormal states)
    np.testing.assert_allclose(np.einsum("ijpp->ij", ca), N * np.eye(5), atol=1e-12)
    # sum_q <p+ q+ q s> = (N-1) <p+ s>
    np.testing.assert_allclose(np.einsum("ijpqqs->ijps", ccaa), (N - 1) * ca, atol=1e-12)
    # adjoint symmetry <I|p+q|J> = <J|q+p|I>
    np.testing.assert_allclose(ca, ca.transpose(1, 0, 3, 2), atol=1e-13)
    # <I|p+|J> = <J|p|I>
    c = orc.density("c", n_orbs, 0, s.coeffs, sm.coeffs, sm.configs)
    a = orc.density("a", n_orbs, 0, sm.coeffs, s.coeffs, s.configs)
    np.testing.assert_allclose(c, a.transpose(1, 0, 2), atol=1e-13)
    # antisymmetry of creator / annihilator pairs
    np.testing.assert_allclose(ccaa, -ccaa.transpose(0, 1, 3, 2, 4, 5), atol=1e-13)
    np.testing.assert_allclose(ccaa, -ccaa.transpose(0, 1, 2, 3, 5, 4), atol=1e-13)


def test_excitation_list_order_and_rebuild():
    rng = np.random.default_rng(5)
    n_orbs, n_core = 4, 1
    sk = _states(rng, n_orbs, 5, 3, n_core)
    sb = _states(rng, n_orbs, 4, 2, n_core)
    for op in ("a", "caa", "ccaa"):
        n_bra = 5 + orc.n_elec_change(op)
        sb = _states(rng, n_orbs, n_bra, 2, n_core)
        q, idx, p, s = orc.excitations(op, n_orbs, n_core, sk.configs)
        dim = 2 * n_orbs
        k = len(op)
        rho = np.zeros((2, 3, dim ** k))
        for t in range(len(q)):
            rho[:, :, idx[t]] += s[t] * np.outer(sb.coeffs[:, p[t]], sk.coeffs[:, q[t]])
        ref = orc.density(op, n_orbs, n_core, sb.coeffs, sk.coeffs, sk.configs)
        np.testing.assert_allclose(rho.reshape(ref.shape), ref, atol=1e-14)
        # loop order: Q ascending; inside Q, the rightmost operator index is
        # the outermost loop, the leftmost the innermost
        assert np.all(np.diff(q) >= 0)
        digits = np.stack([(idx // dim ** (k - 1 - l)) % dim for l in range(k)], axis=1)
        key = digits[:, ::-1]
        for Q in np.unique(q):
            sel = key[q == Q]
            assert all(tuple(sel[i]) < tuple(sel[i + 1]) for i in range(len(sel) - 1))
        # the count mode agrees with a sub-range emission
        q2, _, _, _ = orc.excitations(op, n_orbs, n_core, sk.configs, 1, 3)
        assert len(q2) == int(np.sum((q >= 1) & (q < 3)))


# ---------------------------------------------------------------------------
# whole chain vs exact FCI
# ---------------------------------------------------------------------------

def _complete_fragment(n_orbs, n_elec_ref, rng):
    states = {}
    for ne in range(0, 2 * n_orbs + 1):
        cfg = syn.lex_configs(n_orbs, ne, 0)
        n = cfg.shape[0]
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        states[n_elec_ref - ne] = syn.ChargeStates(configs=cfg, coeffs=Q.T.copy())
    return syn.Fragment(n_orbs=n_orbs, n_core=0, n_elec_ref=n_elec_ref, states=states)


def _xr_vs_fci(frag_orbs, n_elec_ref, with_trimers, seed):
    rng = np.random.default_rng(seed)
    frags = [_complete_fragment(n, n_elec_ref, rng) for n in frag_orbs]
    n_sp = sum(frag_orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    U = np.stack([0.3 * syn.spatial_h(n_sp, rng) for _ in frag_orbs])
    ints = syn.FragmentBlocked(frag_orbs, h, V, U)
    nf = len(frag_orbs)
    nuc = rng.standard_normal((nf, nf))
    nuc = nuc + nuc.T
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    H1, H2, H3 = orc.compute_H_blocks(frags, ints, nuc, with_trimers=with_trimers)
    par = orc.state_parities(frags)
    E_ref, H_ref = orc.fci((H1, H2, H3))
    E0, Hmat = orc.fci((H1, H2, H3), par=par)
    xr = np.sort(np.linalg.eigvalsh(0.5 * (Hmat + Hmat.T)))
    np.testing.assert_allclose(Hmat, Hmat.T, atol=1e-12)
    hf, Vf = ints.full_spin_orbital()
    Hf = brute.hamiltonian(hf, Vf)
    const = sum(nuc[m, m] for m in range(nf)) + sum(nuc[m, n] for m in range(nf) for n in range(m + 1, nf))
    exact = np.sort(np.linalg.eigvalsh(Hf)) + const
    np.testing.assert_allclose(xr, exact, atol=1e-10)
    assert abs(E0 - exact[0]) < 1e-10
    # The reference assembly (general-XRCC/hamiltonian.py:45-47) adds the
    # pair coupling H2[M][N] without the sign of the spectator fragments
    # between M and N.  It differs from the exact matrix only there: one-
    # electron transfers across a spectator with an odd electron count.
    D = np.abs(H_ref - Hmat) > 1e-12
    if nf == 2:
        assert not D.any()
    else:
        spf = H1[0].shape[0]
        si = [f.state_indices for f in frags]
        for i, j in np.argwhere(D):
            bi = [(i // spf ** (nf - 1 - m)) % spf for m in range(nf)]
            bj = [(j // spf ** (nf - 1 - m)) % spf for m in range(nf)]
            tr = [m for m in range(nf) if bi[m] != bj[m]]
            assert len(tr) == 2 and tr[1] - tr[0] > 1
            M, N = tr
            assert abs(si[M][bi[M]][0] - si[M][bj[M]][0]) == 1
            assert any(par[O][bi[O]] for O in range(M + 1, N))


def test_dimer_xr_equals_fci():
    _xr_vs_fci([2, 2], 2, False, 11)


def test_trimer_xr_equals_fci():
    _xr_vs_fci([1, 1, 1], 1, True, 12)


@pytest.mark.parametrize("ops", ["ca", "ccaa", "caaa", "ccca", "caa", "cca", "aa", "cc", "a", "c", "ccaaa"])
def test_density_elements_match_brute_force(ops):
    """oracle.brute.density_elements (bit-mask walk over the ket
    determinants, used for sampled elements at full size) against the
    explicit-operator densities on a 7-orbital space."""
    import itertools
    from oracle import brute
    dim = 7
    rng = np.random.default_rng(len(ops) * 31 + ops.count("c"))
    nk = 3
    nb = nk + ops.count("c") - ops.count("a")
    kc = list(itertools.combinations(range(dim), nk))
    bc = list(itertools.combinations(range(dim), nb))
    zk = rng.standard_normal(len(kc))
    zb = rng.standard_normal(len(bc))
    full = brute.density_fast(ops, brute.embed(zb[None], bc, dim), brute.embed(zk[None], kc, dim), dim)[0, 0]
    tuples = list(itertools.product(range(dim), repeat=len(ops)))
    got = brute.density_elements(ops, tuples, zb, zk, bc, kc)
    ref = np.array([full[t] for t in tuples])
    assert np.abs(got - ref).max() < 1e-13
    assert np.abs(ref).max() > 0.1
