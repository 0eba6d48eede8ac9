"""The oracle's restatement of the hermitian-XRCC dimer diagrams
(oracle/hermitian_ref.py) pinned by brute force: between product states of
two fragments whose orbitals overlap (S = 1 + lam S_off), the matrix
elements of H = sum T a~+ a~ + sum_a U^a a~+ a~ + sum V a~+ a~+ a~ a~ (the
biorthogonal operators a~ on both sides) and the overlaps, expanded in lam
(oracle/nonorth.py, exact dual-number propagation), equal

  * lam^0:  the order-0 diagrams (t01, u100, u001, u101, v0101, v0001,
            v0100, v0011 and the monomer terms t00, u000, v0000 on each
            fragment) -- and the identity for the overlap;
  * lam^1:  the order-1 diagrams (the 4 s01-T, 8 s01-U and 8 s01-V of
            diagram_lists.py) -- and s01 for the overlap,

element by element, for random fragment states of every charge block, a
general (non-symmetric) T and U, and a V that is only antisymmetric in its
creator and in its annihilator pair (as the biorthogonal integrals are).
The product-state ordering of the diagrams is fixed by the lam^0 check."""
import math

import numpy as np
import pytest

from oracle import hermitian_ref as hr
from oracle import nonorth
from oracle import oracle as orc
from qodeapplications_amd import synthetic as syn

OPS = ("c", "a", "ca", "cc", "aa", "cca", "caa", "ccaa", "caaa", "ccca", "ccaaa", "cccaa")


class _Blocked:
    """Fragment blocks of global spin-orbital integrals."""

    def __init__(self, slices, S, T, U, V):
        sl = slices
        nf = len(sl)
        self.S = [[S[sl[m], sl[n]] for n in range(nf)] for m in range(nf)]
        self.T = [[T[sl[m], sl[n]] for n in range(nf)] for m in range(nf)]
        self.U = [[[U[a][sl[m], sl[n]] for n in range(nf)] for m in range(nf)] for a in range(nf)]
        self.V = [[[[V[sl[m], sl[n], sl[o], sl[p]] for p in range(nf)] for o in range(nf)] for n in range(nf)]
                  for m in range(nf)]


def _system(seed, n_orbs=(2, 2), n_ref=(2, 2), counts=(1, 2, 3)):
    rng = np.random.default_rng(seed)
    d = [2 * n for n in n_orbs]
    D = sum(d)
    slices = [slice(0, d[0]), slice(d[0], D)]
    T = rng.standard_normal((D, D))
    U = [rng.standard_normal((D, D)) for _ in range(2)]
    V = rng.standard_normal((D,) * 4)
    V = V - V.transpose(1, 0, 2, 3)
    V = V - V.transpose(0, 1, 3, 2)
    s = rng.standard_normal((n_orbs[0], n_orbs[1])) * 0.5
    S_off = np.zeros((D, D))
    for sp in range(2):                 # alpha-alpha and beta-beta blocks
        r0 = slice(sp * n_orbs[0], (sp + 1) * n_orbs[0])
        r1 = slice(d[0] + sp * n_orbs[1], d[0] + (sp + 1) * n_orbs[1])
        S_off[r0, r1] = s
    S_off = S_off + S_off.T
    frags = []
    for m in range(2):
        states = {}
        for N in counts:
            ns = min(3, math.comb(d[m], N))
            states[n_ref[m] - N] = syn.random_states(n_orbs[m], N, ns, rng)
        frags.append(states)
    rho = []
    for m in range(2):
        r = {}
        for op in OPS:
            dn = sum(1 if o == "c" else -1 for o in op)
            for cb, sb in frags[m].items():
                for ck, sk in frags[m].items():
                    if cb != ck - dn:
                        continue
                    r.setdefault(op, {})[(cb, ck)] = orc.density(op, n_orbs[m], 0, sb.coeffs, sk.coeffs, sk.configs)
        rho.append(r)
    ints = _Blocked(slices, S_off, T, U, V)
    return dict(n_orbs=n_orbs, d=d, n_ref=n_ref, frags=frags, rho=rho, ints=ints, T=T, U=U, V=V, S_off=S_off)


def _brute(sysd, inner):
    """{(bra charges, ket charges): (S0, S1, H0, H1)} in the state basis
    [i0, i1, j0, j1]."""
    d, n_ref, frags = sysd["d"], sysd["n_ref"], sysd["frags"]
    E = nonorth.Elements(d[0], d[1], sysd["S_off"], inner=inner)
    T, U, V = sysd["T"], sysd["U"], sysd["V"]
    Ug = U[0] + U[1]
    out = {}
    for cb0, sb0 in frags[0].items():
        for cb1, sb1 in frags[1].items():
            for ck0, sk0 in frags[0].items():
                for ck1, sk1 in frags[1].items():
                    if cb0 + cb1 != ck0 + ck1:
                        continue
                    bra = (n_ref[0] - cb0, n_ref[1] - cb1)
                    ket = (n_ref[0] - ck0, n_ref[1] - ck1)
                    ov = E.overlap(bra, ket)
                    one = E.one_body(bra, ket)
                    two = E.two_body(bra, ket)
                    h = [np.einsum("PQpq,pq->PQ", one[k], T + Ug) + np.einsum("PQpqsr,pqrs->PQ", two[k], V)
                         for k in (0, 1)]
                    tr = lambda M: nonorth.state_transform(M, sb0.coeffs, sb1.coeffs, sk0.coeffs, sk1.coeffs)
                    out[((cb0, ck0), (cb1, ck1))] = (tr(ov[0]), tr(ov[1]), tr(h[0]), tr(h[1]))
    return out


ORDER0 = hr.ST2[0] + hr.SU2[0] + hr.SV2[0]
ORDER1 = hr.ST2[1] + hr.SU2[1] + hr.SV2[1]


def _check(sysd, inner):
    n_elec0 = {c: sysd["n_ref"][0] - c for c in sysd["frags"][0]}
    bad0 = bad1 = bads = 0.0
    scale = 1.0
    for charges, (S0, S1, H0, H1) in _brute(sysd, inner).items():
        g0 = hr.dimer_block(sysd["rho"], charges, n_elec0, sysd["ints"], ORDER0,
                            monomer_terms=("t00", "u000", "v0000"))
        g1 = hr.dimer_block(sysd["rho"], charges, n_elec0, sysd["ints"], ORDER1)
        gs = hr.dimer_block(sysd["rho"], charges, n_elec0, sysd["ints"], hr.S2[1])
        scale = max(scale, float(np.abs(H0).max()), float(np.abs(H1).max()))
        bad0 = max(bad0, float(np.abs(g0 - H0).max()))
        bad1 = max(bad1, float(np.abs(g1 - H1).max()))
        bads = max(bads, float(np.abs(gs - S1).max()))
        eye = np.einsum("ij,kl->ikjl", np.eye(S0.shape[0]), np.eye(S0.shape[1])) if charges[0][0] == charges[0][1] \
            and charges[1][0] == charges[1][1] else 0.0
        assert np.abs(S0 - eye).max() < 1e-12
    return bad0 / scale, bad1 / scale, bads


@pytest.mark.parametrize("seed,n_orbs", [(1, (2, 2)), (2, (2, 3))])
def test_diagrams_are_the_overlap_expansion(seed, n_orbs):
    """inner='f1': fragment 0's creation string leftmost (the other ordering
    differs by (-1)^(N0 N1) and fails the lam^0 check)."""
    sysd = _system(seed, n_orbs=n_orbs, counts=(1, 2, 3) if n_orbs == (2, 2) else (2, 3))
    b0, b1, bs = _check(sysd, "f1")
    assert b0 < 1e-12 and bs < 1e-12 and b1 < 1e-12, (b0, bs, b1)


def test_other_product_ordering_fails_order_zero():
    sysd = _system(3, n_orbs=(2, 2), counts=(1, 2, 3))
    assert _check(sysd, "f0")[0] > 1e-3


def _one_electron_dimer(seed):
    """Fragments of 2 and 3 spatial orbitals with 0 or 1 electrons (every
    state), symmetric h split into T and U, a plain (symmetric) V, and the
    biorthogonal integrals of get_ints.py:125-146 in spin-orbital form."""
    rng = np.random.default_rng(seed)
    n_orbs = (2, 3)
    d = [2 * n for n in n_orbs]
    D = sum(d)
    slices = [slice(0, d[0]), slice(d[0], D)]
    # fragment orbitals orthonormal within each fragment, overlapping across
    S_sp = np.eye(sum(n_orbs))
    s01 = 0.3 * rng.standard_normal(n_orbs)
    S_sp[:n_orbs[0], n_orbs[0]:] = s01
    S_sp[n_orbs[0]:, :n_orbs[0]] = s01.T
    h_sp = syn.spatial_h(sum(n_orbs), rng)
    frag_of = np.array([0] * n_orbs[0] + [1] * n_orbs[1])
    # spatial -> spin-orbital (fragment blocks alpha, beta)
    glob = []
    for m in range(2):
        base = sum(n_orbs[:m])
        glob += [base + k for k in range(n_orbs[m])] * 2
    spin = [0] * n_orbs[0] + [1] * n_orbs[0] + [0] * n_orbs[1] + [1] * n_orbs[1]
    same = np.equal.outer(spin, spin)
    S = S_sp[np.ix_(glob, glob)] * same
    h = h_sp[np.ix_(glob, glob)] * same
    U = [0.3 * h * np.equal.outer(frag_of[glob] == a, frag_of[glob] == a) for a in (0, 1)]
    T = h - U[0] - U[1]
    V = rng.standard_normal((D,) * 4) * 0.1
    V = V + V.transpose(2, 3, 0, 1)
    V = V - V.transpose(1, 0, 2, 3)
    V = V - V.transpose(0, 1, 3, 2)
    Si = np.linalg.inv(S)
    symm = _Blocked(slices, S - np.eye(D), T, U, V)
    bior = _Blocked(slices, S - np.eye(D), Si @ T, [Si @ u for u in U],
                    np.einsum("pa,qb,abrs->pqrs", Si, Si, V))
    Vd = np.einsum("pa,aqrs->pqrs", Si, V) - np.einsum("pa,qb,abrs->pqrs", Si, Si, V)
    bior.Vd = _Blocked(slices, S, T, U, Vd).V
    frags, rho = [], []
    for m in range(2):
        states = {1 - N: syn.random_states(n_orbs[m], N, math.comb(d[m], N), rng) for N in (0, 1)}
        frags.append(states)
        r = {}
        for op in OPS:
            dn = sum(1 if o == "c" else -1 for o in op)
            for cb, sb in states.items():
                for ck, sk in states.items():
                    if cb == ck - dn:
                        r.setdefault(op, {})[(cb, ck)] = orc.density(op, n_orbs[m], 0, sb.coeffs, sk.coeffs,
                                                                     sk.configs)
        rho.append(r)
    return dict(h_sp=h_sp, S_sp=S_sp, symm=symm, bior=bior, frags=frags, rho=rho)


@pytest.mark.parametrize("order", [0, 1])
def test_one_electron_xr_is_exact(order):
    """With one electron in the dimer the order-1 overlap and Hamiltonian
    expansions terminate, so the XR matrix of get_xr_H (order 1: S2^-1 S2H2
    with the monomers added back; order 0: the biorthogonal blocks) has the
    eigenvalues of h in the non-orthogonal orbitals (each once per spin)."""
    from scipy.linalg import eigh
    sysd = _one_electron_dimer(4)
    charges = [[1, 0], [1, 0]]                        # 0 then 1 electron
    n_elec0 = {1: 0, 0: 1}
    H1, H2 = hr.get_xr_H(sysd["symm"], sysd["bior"], sysd["rho"], n_elec0, charges, order)
    ns = [[sysd["frags"][m][c].coeffs.shape[0] for c in charges[m]] for m in range(2)]
    M = hr.xr_matrix(H1, H2, ns[0], ns[1])
    # the one-electron sector: charge blocks (1, 0) and (0, 1)
    cb = [(a, b) for a in charges[0] for b in charges[1]]
    sizes = [ns[0][charges[0].index(a)] * ns[1][charges[1].index(b)] for a, b in cb]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    idx = np.concatenate([np.arange(off[k], off[k + 1]) for k, (a, b) in enumerate(cb) if a + b == 1])
    ev = np.sort(np.linalg.eigvals(M[np.ix_(idx, idx)]).real)
    ref = np.sort(np.repeat(eigh(sysd["h_sp"], sysd["S_sp"], eigvals_only=True), 2))
    np.testing.assert_allclose(ev, ref, atol=1e-10)
