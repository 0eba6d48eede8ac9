"""CPU restatement of the reference's gradient definitions for a dimer at XR
order 0.  TEST INFRASTRUCTURE ONLY: the GPU adjoint (qodeapplications_amd/
grad.py, orbitals.py) is compared against these; nothing in the product
package imports them.

State gradient, "herm" level (StateSpaceOptimizer/state_gradients.py:121-244):
  d = the target eigenvector of the XR dimer matrix, reshaped (A states x B
  states) and normalised (:145-158; "herm" takes dl = dr = the right
  eigenvector, :114-116).  For fragment A = frag_ind and each charge c of A:
      g[c][k, p] =   sum_i H1_A[p, i] O[c][k, i]                  (:211)
                   - E sum_i z_A[c][i, p] O[c][k, i]              (:212, E_prefac = -1)
                   + sum_{l, i, l'} d[k, l] H2[(p, l), (i, l')] d[i, l']   (:214/216)
                   + sum_i z_A[c][i, p] sum_{l, l'} d[k, l] H1_B[l, l'] d[i, l']   (:218)
  with p running over A's determinants of charge c (the 'bra_det'
  densities of :170: bra coefficients = the identity over determinants),
  O[c][k, i] = sum_l d[k, l] d[i, l] over the B states of charge -c
  (get_adapted_overlaps, :29-37) and l restricted to B charge -c
  (contract_dens_with_d, :39-71).  Like the reference it requires the
  fragment charges to be symmetric about 0 (:30-32) and, through that
  slicing, reads d on the total-charge-0 sector only.  The matrix elements
  come from the oracle's build_matrix_elements restatement (general-XRCC/
  build_H.py:37-188) on the oracle's densities.

Orbital gradient (StateSpaceOptimizer/orb_grads.py:867-999):
  the XR one- and two-particle densities over the dimer's spin orbitals,
      gamma[P, Q] = <Psi| P+ Q |Psi>,  Gamma[P, Q, R, S] = <Psi| P+ Q+ S R |Psi>,
  assembled from fragment densities contracted with dl, dr per the
  catalogs one_p_dens_catalog / two_p_dens_catalog (:281-298, diagrams
  one_p_diagonal / one_p_off_diagonal :36-58, two_p_0000 ... two_p_0011
  :260-279, loops one_p_dens :867-921 and two_p_dens :923-966, the sign
  (-1)^(n_i1 + perm) with n_i1 = get_prefactor_exponent, :29-30), then the
  t / u / v diagrams (t_diagram, t_diagram2 :339-370, u_diagram,
  u_diagram2 :413-447, v_diagram, v_diagram2 :649-695) and
  grad - grad^T (orb_grads :985-999).  With every block summed the
  diagrams are
      grad  = gamma^T h - h gamma^T
            + sum V[a, b, c, Q] Gamma[a, b, c, P] + V[a, b, Q, d] Gamma[a, b, P, d]
            - V[P, b, c, d] Gamma[Q, b, c, d] - V[a, P, c, d] Gamma[a, Q, c, d]
  (h = T + sum_n U_n, the biorthogonalised integrals, which equal the plain
  ones for orthonormal orbitals, the XR-order-0 case).

How these are pinned (tests/test_oracle_grad.py): the densities gamma and
Gamma against the brute-force Fock-space expectation values of the
embedded XR state; E recomputed from them; the state gradient against
central finite differences of the generalised XR eigenvalue; the orbital
gradient against central finite differences of the XR energy under
orbital rotations.
"""
import itertools

import numpy as np

from . import oracle as orc


# ---------------------------------------------------------------------------
# the XR dimer and its eigenvector in the reference's charge-block layout
# ---------------------------------------------------------------------------

def charge_slices(frag, charges):
    """d_slices of state_gradients.get_slices (:73-93): per charge, the slice
    of its states in the charge-block order."""
    out, off = {}, 0
    for c in charges:
        n = frag.states[c].coeffs.shape[0]
        out[c] = slice(off, off + n)
        off += n
    return out


def block_order(frag, charges):
    """Position in the fragment's state_indices order of every state in the
    charge-block order."""
    where = {si: k for k, si in enumerate(frag.state_indices)}
    return np.array([where[(c, i)] for c in charges for i in range(frag.states[c].coeffs.shape[0])])


def dimer_root(frags, ints, nuc, target=0):
    """(E, d) of the XR dimer (d: A states x B states in the charge-block
    order of each fragment's charges, normalised) and the XR matrix in the
    state_indices product order."""
    H1, H2, _ = orc.compute_H_blocks(frags, ints, nuc)
    _, Hmat = orc.fci((H1, H2, None))
    w, v = np.linalg.eigh(0.5 * (Hmat + Hmat.T))
    E = float(w[target])
    dA, dB = H1[0].shape[0], H1[1].shape[0]
    d_prod = v[:, target].reshape(dA, dB)
    oA = block_order(frags[0], list(frags[0].states))
    oB = block_order(frags[1], list(frags[1].states))
    d = d_prod[oA][:, oB]
    return E, d / np.linalg.norm(d), Hmat


# ---------------------------------------------------------------------------
# state gradient, "herm" level
# ---------------------------------------------------------------------------

class _DetBraFragment:
    """A fragment whose densities have determinant bras (the 'bra_det'
    option): rho[op][(bc, kc)][p][j] = <p| op |j> for determinant p of
    charge bc; the ccaa entries contracted with V_AAAA per (p, j), as
    build_density_tensors does (general-XRCC/build_density_tensors.py:128-133)."""

    def __init__(self, frag, V):
        self.n_elec_ref = frag.n_elec_ref
        self.rho = {op: {} for op in ("aa", "a", "caa", "ca", "ccaa", "c", "cca", "cc")}
        n_so = 2 * frag.n_orbs
        for bc in frag.states:
            eye = np.eye(frag.states[bc].configs.shape[0])
            for kc, sk in frag.states.items():
                for op in orc.OPS_BY_CHG_DIFF.get(bc - kc, ()):
                    R = orc.density(op, frag.n_orbs, frag.n_core, eye, sk.coeffs, sk.configs)
                    if op == "ccaa":
                        self.rho[op][(bc, kc)] = [[orc.monomer_2e(n_so, R[p, j], V) for j in range(R.shape[1])]
                                                  for p in range(R.shape[0])]
                    else:
                        self.rho[op][(bc, kc)] = [[R[p, j] for j in range(R.shape[1])] for p in range(R.shape[0])]


def herm_state_gradient(frags, ints, nuc, frag_ind=0, target=0):
    """g[c] (A states of charge c x A determinants of charge c), the "herm"
    state gradient of state_gradients.py:121-244 at xr_order 0, for
    fragment A = frags[frag_ind].  Returns (E, g, d)."""
    charges = [list(f.states) for f in frags]
    for ch in charges:
        if sum(ch) != 0:
            raise ValueError("charges must be symmetric about zero (state_gradients.py:30-32)")
    # the determinant-bra elements below reuse build_matrix_elements, whose
    # nuclear and extPot terms carry a Kronecker delta between the bra and
    # ket state indices; with a determinant bra that delta is an overlap
    # <p|i>, so those terms are left out and required to vanish
    if np.any(np.asarray(nuc) != 0.0) or (ints.U_sp is not None and np.any(ints.U_sp != 0.0)):
        raise ValueError("herm_state_gradient supports U = 0 and zero nuclear repulsion only")
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], 0)
    E, d, _ = dimer_root(frags, ints, nuc, target)
    if frag_ind == 1:
        d = d.T
    A, B = frag_ind, 1 - frag_ind
    fA, fB = frags[A], frags[B]
    slA, slB = charge_slices(fA, charges[A]), charge_slices(fB, charges[B])
    # monomer and pair matrix elements with A's bra a determinant
    det = _DetBraFragment(fA, np.ascontiguousarray(ints.V[A, A, A, A]))
    sys_det = [det, frags[B]] if A == 0 else [frags[B], det]
    bme_det = orc.build_matrix_elements(sys_det, ints, nuc)
    bme = orc.build_matrix_elements(frags, ints, nuc)
    statesA = [(c, i) for c in charges[A] for i in range(fA.states[c].coeffs.shape[0])]
    statesB = [(c, i) for c in charges[B] for i in range(fB.states[c].coeffs.shape[0])]
    H1B = np.array([[bme.monomer(B, I, J) for J in statesB] for I in statesB])
    g = {}
    for c in charges[A]:
        z = fA.states[c].coeffs
        nP = z.shape[1]
        kA = slA[c]
        lB = slB[-c] if -c in slB else slice(0, 0)
        dk = d[kA][:, lB]                              # d[k, l], l in B(-c)
        O = dk @ d[kA][:, lB].T                        # O[c][k, i], get_adapted_overlaps
        dets = [(c, p) for p in range(nP)]
        # A monomer block between determinant bras and state kets of charge c
        H1A = np.array([[bme_det.monomer(A, Pd, (c, i)) for i in range(z.shape[0])] for Pd in dets])
        gc = O @ H1A.T                                  # sum_i H1A[p, i] O[k, i]
        gc -= E * (O @ z)                               # E term
        # pair term: sum_{l, i, l'} d[k, l] <p l| H_AB |i l'> d[i, l']
        Hd = np.zeros((nP, len(statesB)))                # sum_{i, l'} <p l|H2|i l'> d[i, l']
        for pi, Pd in enumerate(dets):
            for li, L in enumerate(statesB):
                if L[0] != -c:
                    continue
                acc = 0.0
                for ii, Ii in enumerate(statesA):
                    for lj, Lj in enumerate(statesB):
                        if d[ii, lj] == 0.0:
                            continue
                        if A == 0:
                            el = bme_det.dimer((0, 1), (Pd, L), (Ii, Lj))
                        else:
                            el = bme_det.dimer((0, 1), (L, Pd), (Lj, Ii))
                        acc += el * d[ii, lj]
                Hd[pi, li] = acc
        gc += dk @ Hd[:, lB].T
        # B monomer term: sum_i z[i, p] sum_{l, l'} d[k, l] H1_B[l, l'] d[i, l']
        W = dk @ H1B[lB] @ d[kA].T                      # (k, i) with i over A(c)
        gc += W @ z
        g[c] = gc
    return E, g, d


# ---------------------------------------------------------------------------
# XR one- and two-particle densities, orbital gradient
# ---------------------------------------------------------------------------

def _prefactor_exponent(n_occ_total, bra_chg_second):
    """get_prefactor_exponent (orb_grads.py:29-30): (n_occ - bra charge of
    the second fragment of the diagram) mod 2."""
    return (n_occ_total - bra_chg_second) % 2


def xr_densities(frags, dl_, dr_, n_occ_total):
    """(gamma, Gamma) over the dimer's spin orbitals (fragment 0 first, each
    fragment alpha then beta), per the catalogs of orb_grads.py:281-298 and
    the loops of one_p_dens / two_p_dens (:867-966).

    The sign (-1)^(n_i1 + perm) with n_i1 = (n_occ - bra charge of the
    second fragment) holds when dl, dr lie in the total-charge-0 sector,
    the sector state_gradients.py:30-32 assumes; a state of odd total
    charge Q would need another factor (-1)^Q on the perm = 1 terms, so
    other sectors are refused here."""
    ch = [[c for c in f.states for _ in range(f.states[c].coeffs.shape[0])] for f in frags]
    q = np.add.outer(np.array(ch[0]), np.array(ch[1]))
    if max(float(np.abs(dl_[q != 0]).max(initial=0.0)), float(np.abs(dr_[q != 0]).max(initial=0.0))) > 1e-12:
        raise ValueError("dl, dr must lie in the total-charge-0 sector")
    dims = [2 * f.n_orbs for f in frags]
    off = [0, dims[0], dims[0] + dims[1]]
    ms = [slice(off[0], off[1]), slice(off[1], off[2])]
    n = off[2]
    charges = [list(f.states) for f in frags]
    dsl = [charge_slices(f, ch) for f, ch in zip(frags, charges)]
    dens = [orc.build_density_arrays(f.states, f.n_orbs, f.n_core) for f in frags]
    gamma = np.zeros((n, n))
    Gamma = np.zeros((n, n, n, n))
    one_cat = {(0, 0): ((0, 0), ("ca",)), (0, 1): ((-1, +1), ("c", "a"))}
    for (i, j), (Dchgs, ops) in one_cat.items():
        for perm in (0, 1):
            dl, dr = (dl_.T, dr_.T) if perm else (dl_, dr_)
            ii, jj = (1 - i, 1 - j) if perm else (i, j)
            if ii == jj:
                for c in dsl[ii]:
                    R = dens[ii]["ca"][(c, c)]
                    int1 = dl[dsl[ii][c], :] @ dr[dsl[ii][c], :].T          # (i0, j0)
                    gamma[ms[ii], ms[ii]] += np.einsum("ij,ijpq->pq", int1, R)
                continue
            f0, f1 = 0 + perm, 1 - perm
            for bi in dsl[f0]:
                for ki in dsl[f0]:
                    if bi - ki != Dchgs[0]:
                        continue
                    for bj in dsl[f1]:
                        for kj in dsl[f1]:
                            if bj - kj != Dchgs[1]:
                                continue
                            s = (-1) ** (_prefactor_exponent(n_occ_total, bj) + perm)
                            R0 = dens[f0][ops[0]][(bi, ki)]
                            R1 = dens[f1][ops[1]][(bj, kj)]
                            L = dl[dsl[ii][bi], dsl[jj][bj]]
                            Rr = dr[dsl[ii][ki], dsl[jj][kj]]
                            int1 = np.einsum("ab,acp,cd->pbd", L, R0, Rr)      # (p, i1, j1)
                            gamma[ms[ii], ms[jj]] += s * np.einsum("pbd,bdq->pq", int1, R1)
    two_cat = [((0, 0, 0, 0), (0, 0), ("ccaa",), "0000", (0, 1, 2, 3)),
               ((0, 1, 0, 1), (0, 0), ("ca", "ca"), "0101", (0, 1, 2, 3)),
               ((0, 1, 1, 0), (0, 0), ("ca", "ca"), "0101", (0, 1, 3, 2)),
               ((0, 0, 1, 0), (-1, +1), ("cca", "a"), "0010", (0, 1, 2, 3)),
               ((0, 0, 0, 1), (-1, +1), ("cca", "a"), "0010", (0, 1, 3, 2)),
               ((0, 1, 1, 1), (-1, +1), ("c", "caa"), "0111", (0, 1, 2, 3)),
               ((1, 0, 1, 1), (-1, +1), ("c", "caa"), "0111", (1, 0, 2, 3)),
               ((0, 0, 1, 1), (-2, +2), ("cc", "aa"), "0011", (0, 1, 2, 3))]
    for key, Dchgs, ops, diag, tr in two_cat:
        pref = 1 if tr == (0, 1, 2, 3) else -1
        for perm in (0, 1):
            dl, dr = (dl_.T, dr_.T) if perm else (dl_, dr_)
            kk = tuple(1 - x for x in key) if perm else key
            blk = tuple(ms[x] for x in kk)
            if len(ops) == 1:
                f = kk[0]
                for c in dsl[f]:
                    R = dens[f]["ccaa"][(c, c)]
                    int1 = dl[dsl[f][c], :] @ dr[dsl[f][c], :].T
                    G = np.einsum("ijpqsr,ij->pqrs", R, int1)
                    Gamma[blk] += pref * np.transpose(G, tr)
                continue
            f0, f1 = 0 + perm, 1 - perm
            for b0 in dsl[f0]:
                for k0 in dsl[f0]:
                    if b0 - k0 != Dchgs[0]:
                        continue
                    for b1 in dsl[f1]:
                        for k1 in dsl[f1]:
                            if b1 - k1 != Dchgs[1]:
                                continue
                            sgn = (-1) ** (_prefactor_exponent(n_occ_total, b1) + perm)
                            R0 = dens[f0][ops[0]][(b0, k0)]
                            R1 = dens[f1][ops[1]][(b1, k1)]
                            L = dl[dsl[f0][b0], dsl[f1][b1]]
                            Rr = dr[dsl[f0][k0], dsl[f1][k1]]
                            if diag == "0101":
                                i1 = np.einsum("ab,bdqs,cd->acqs", L, R1, Rr)
                                G = np.einsum("acqs,acpr->pqrs", i1, R0)
                            elif diag == "0010":
                                i1 = np.einsum("ab,bdr,cd->acr", L, R1, Rr)
                                G = sgn * np.einsum("acpqs,acr->pqrs", R0, i1)
                            elif diag == "0111":
                                i1 = np.einsum("ab,acp,cd->bdp", L, R0, Rr)
                                G = sgn * np.einsum("bdp,bdqsr->pqrs", i1, R1)
                            else:      # 0011
                                i1 = np.einsum("ab,acpq,cd->bdpq", L, R0, Rr)
                                G = np.einsum("bdpq,bdsr->pqrs", i1, R1)
                            Gamma[blk] += pref * np.transpose(G, tr)
    return gamma, Gamma


def orbital_gradient(gamma, Gamma, h, V):
    """t + u + v diagrams and grad - grad^T (orb_grads.py:339-447, 649-695,
    985-999) with every block summed; h = T + sum_n U_n."""
    grad = gamma.T @ h - h @ gamma.T
    grad += np.einsum("abcq,abcp->pq", V, Gamma) + np.einsum("abqd,abpd->pq", V, Gamma)
    grad -= np.einsum("pbcd,qbcd->pq", V, Gamma) + np.einsum("apcd,aqcd->pq", V, Gamma)
    return grad - grad.T


def rdm_energy(gamma, Gamma, h, V):
    return float(np.sum(h * gamma) + np.sum(V * Gamma))


def spin_summed_fragment_blocks(g, frag_orbs):
    """Per fragment m, sum over spin of the (m, m) spin-orbital block of g:
    the gradient with respect to a spatial rotation applied to both spins."""
    out, off = [], 0
    for n in frag_orbs:
        a = slice(off, off + n)
        b = slice(off + n, off + 2 * n)
        out.append(g[a, a] + g[b, b])
        off += 2 * n
    return out


def embed_dimer(frags, d):
    """The XR dimer state sum_{i,l} d[i, l] |i>|l> (d in the charge-block
    order) as a Fock-space vector over the bit strings of the dimer's spin
    orbitals, fragment 0's bits lowest.  The product of a determinant of
    each fragment is the determinant of the union of their occupations with
    no extra sign: the convention in which the pair terms of
    general-XRCC/H_contractions.c carry sign2 = (-1)^(electrons of the
    second fragment's bra) (build_H.py:63-64).  A state whose product
    basis were ordered the other way would differ by (-1)^(n_A n_B) per
    product state, and the brute-force densities would then miss E."""
    f0, f1 = frags
    d0 = 2 * f0.n_orbs
    from . import brute
    vec = np.zeros(1 << (d0 + 2 * f1.n_orbs))
    ch = [list(f.states) for f in frags]
    A = [(c, i) for c in ch[0] for i in range(f0.states[c].coeffs.shape[0])]
    Bs = [(c, i) for c in ch[1] for i in range(f1.states[c].coeffs.shape[0])]
    for ia, (ca, i) in enumerate(A):
        za, cfa = f0.states[ca].coeffs[i], f0.states[ca].configs
        for lb, (cb, l) in enumerate(Bs):
            if d[ia, lb] == 0.0:
                continue
            zb, cfb = f1.states[cb].coeffs[l], f1.states[cb].configs
            for x, ca_cfg in enumerate(cfa):
                bits_a = brute.config_bits(ca_cfg)
                for y, cb_cfg in enumerate(cfb):
                    bits = bits_a | (brute.config_bits(cb_cfg) << d0)
                    vec[bits] += d[ia, lb] * za[x] * zb[y]
    return vec


def _annihilate(p, vec, dim):
    """a_p applied to a Fock-space vector over bit strings, with the sign
    convention of oracle/brute.py: a_p|n> = (-1)^popcount(n >> (p+1)) |n - p>."""
    idx = np.arange(1 << dim)
    occ = ((idx >> p) & 1).astype(bool)
    above = idx >> (p + 1)
    par = np.zeros(idx.size, dtype=np.int64)
    x = above.copy()
    while np.any(x):
        par ^= x & 1
        x >>= 1
    out = np.zeros_like(vec)
    src = idx[occ]
    out[src & ~(1 << p)] = np.where(par[occ] == 1, -1.0, 1.0) * vec[occ]
    return out


def brute_densities(vec, dim):
    """gamma[P, Q] = <v|P+ Q|v>, Gamma[P, Q, R, S] = <v|P+ Q+ S R|v> by explicit
    Fock-space annihilators on the bit-string vector (the sign rule of
    oracle/brute.py)."""
    av = [_annihilate(r, vec, dim) for r in range(dim)]                       # R|v>
    aav = [[_annihilate(s, av[r], dim) for r in range(dim)] for s in range(dim)]   # S R|v>
    gamma = np.array([[av[p] @ av[q] for q in range(dim)] for p in range(dim)])   # <v|P+ Q|v> = (P v).(Q v)
    Gamma = np.zeros((dim,) * 4)
    for p, q, r, s in itertools.product(range(dim), repeat=4):
        # <v|P+ Q+ S R|v> = (Q P v) . (S R v)
        Gamma[p, q, r, s] = aav[q][p] @ aav[s][r]
    return gamma, Gamma
