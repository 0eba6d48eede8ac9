"""CPU restatement of the hermitian-XRCC dimer diagrams through XR order 1
and of get_xr_H at orders 0 and 1.  TEST INFRASTRUCTURE ONLY.

Diagrams (hermitian-XRCC/diagrams/S_0mer_0.py, S_2mer_1.py,
S{T,U,V}_1mer_0.py, S{T,U,V}_2mer_0.py, S{T,U,V}_2mer_1.py; the commented
einsum form above each precontracted product), their charge conditions
and fragment permutations (S{,T,U,V}_diagrams.py catalogs), the sign
(-1)^n_j0 with n_j0 the ket electron parity of the unpermuted fragment 0
(diagrams/build_diagram.py:79-84), and the block assembly
(XR_term.py:22-165, get_xr_result.py:86-198, the reorder :300-353).

Each diagram is one numpy einsum over the two fragments' densities and the
integral blocks (the reference factors it through cached precontractions,
precontract.py:26-94; the sum is the same).  Pinned in
tests/test_oracle_hermitian.py: the order-0 diagrams sum to the lam^0 and
the order-1 (s01) diagrams to the lam^1 coefficient of the matrix elements
between product states of non-orthogonal orbitals (oracle/nonorth.py).

Densities: rho[op][(bra_chg, ket_chg)] = (n_i, n_j, d, ..., d); integrals
are fragment-blocked spin-orbital blocks: S[m][n], T[m][n], U[a][m][n],
V[m][n][o][p] (numpy arrays; S only inter-fragment blocks are read).
"""
import numpy as np

# name: (prefactor, j-sign exponent or None, (op, idx) of role 0, (op, idx) of
#        role 1 (None: delta on role 1), integral factors, Dchgs, permutations)
# integral factors: (kind, roles, idx) with kind S / T roles (x, y);
# U roles (nucleus, x, y); V roles (w, x, y, z).
_P2 = [(+1, (0, 1)), (+1, (1, 0))]
_PM = [(+1, (0, 1)), (-1, (1, 0))]
DIMER = {
    # order 0 (S{T,U,V}_2mer_0.py)
    "t01":   (1, 0, ("c", "p"), ("a", "q"), [("T", (0, 1), "pq")], (-1, 1), _PM),
    "u100":  (1, None, ("ca", "pq"), None, [("U", (1, 0, 0), "pq")], (0, 0), _P2),
    "u001":  (1, 0, ("c", "p"), ("a", "q"), [("U", (0, 0, 1), "pq")], (-1, 1), _PM),
    "u101":  (1, 0, ("c", "p"), ("a", "q"), [("U", (1, 0, 1), "pq")], (-1, 1), _PM),
    "v0101": (4, None, ("ca", "pr"), ("ca", "qs"), [("V", (0, 1, 0, 1), "pqrs")], (0, 0), [(+1, (0, 1))]),
    "v0001": (2, 1, ("cca", "pqr"), ("a", "s"), [("V", (0, 0, 0, 1), "pqrs")], (-1, 1), _PM),
    "v0100": (2, 0, ("caa", "psr"), ("c", "q"), [("V", (0, 1, 0, 0), "pqrs")], (1, -1), _PM),
    "v0011": (1, None, ("cc", "pq"), ("aa", "sr"), [("V", (0, 0, 1, 1), "pqrs")], (-2, 2), _P2),
    # order 1 (S_2mer_1.py, S{T,U,V}_2mer_1.py)
    "s01":   (1, 0, ("c", "p"), ("a", "q"), [("S", (0, 1), "pq")], (-1, 1), _PM),
    "s01t10": (-1, None, ("ca", "tq"), ("ca", "pu"), [("S", (0, 1), "tu"), ("T", (1, 0), "pq")], (0, 0), _P2),
    "s01t00": (1, 1, ("cca", "ptq"), ("a", "u"), [("S", (0, 1), "tu"), ("T", (0, 0), "pq")], (-1, 1), _PM),
    "s01t11": (1, 1, ("c", "t"), ("caa", "puq"), [("S", (0, 1), "tu"), ("T", (1, 1), "pq")], (-1, 1), _PM),
    "s01t01": (1, None, ("cc", "pt"), ("aa", "uq"), [("S", (0, 1), "tu"), ("T", (0, 1), "pq")], (-2, 2), _P2),
    "s01v0100": (-2, None, ("ccaa", "ptsr"), ("ca", "qu"), [("S", (0, 1), "tu"), ("V", (0, 1, 0, 0), "pqrs")],
                 (0, 0), _P2),
    "s01v1101": (2, None, ("ca", "tr"), ("ccaa", "pqus"), [("S", (0, 1), "tu"), ("V", (1, 1, 0, 1), "pqrs")],
                 (0, 0), _P2),
    "s01v0000": (1, 0, ("cccaa", "pqtsr"), ("a", "u"), [("S", (0, 1), "tu"), ("V", (0, 0, 0, 0), "pqrs")],
                 (-1, 1), _PM),
    "s01v0101": (4, 0, ("cca", "ptr"), ("caa", "qus"), [("S", (0, 1), "tu"), ("V", (0, 1, 0, 1), "pqrs")],
                 (-1, 1), _PM),
    "s01v1100": (1, 0, ("caa", "tsr"), ("cca", "pqu"), [("S", (0, 1), "tu"), ("V", (1, 1, 0, 0), "pqrs")],
                 (1, -1), _PM),
    "s01v1111": (1, 0, ("c", "t"), ("ccaaa", "pqusr"), [("S", (0, 1), "tu"), ("V", (1, 1, 1, 1), "pqrs")],
                 (-1, 1), _PM),
    "s01v0001": (2, None, ("ccca", "pqtr"), ("aa", "us"), [("S", (0, 1), "tu"), ("V", (0, 0, 0, 1), "pqrs")],
                 (-2, 2), _P2),
    "s01v0111": (-2, None, ("cc", "pt"), ("caaa", "qusr"), [("S", (0, 1), "tu"), ("V", (0, 1, 1, 1), "pqrs")],
                 (-2, 2), _P2),
}
# the s01-U diagrams: the s01-T ones with U of nucleus 0 or 1 (SU_2mer_1.py)
for _nuc in (0, 1):
    for _t, _u in (("s01t10", f"s01u{_nuc}10"), ("s01t00", f"s01u{_nuc}00"), ("s01t11", f"s01u{_nuc}11"),
                   ("s01t01", f"s01u{_nuc}01")):
        f, js, A, B, ints, dch, perms = DIMER[_t]
        (_, roles, idx) = ints[1]
        DIMER[_u] = (f, js, A, B, [ints[0], ("U", (_nuc,) + roles, idx)], dch, perms)

# diagram lists (hermitian-XRCC/diagram_lists.py:11-73)
ST2 = {0: ["t01"], 1: ["s01t10", "s01t00", "s01t11", "s01t01"]}
SU2 = {0: ["u100", "u001", "u101"],
       1: ["s01u010", "s01u000", "s01u011", "s01u001", "s01u110", "s01u100", "s01u111", "s01u101"]}
SV2 = {0: ["v0101", "v0001", "v0100", "v0011"],
       1: ["s01v0100", "s01v1101", "s01v0000", "s01v0101", "s01v1100", "s01v1111", "s01v0001", "s01v0111"]}
S2 = {1: ["s01"]}


def _integral(ints, kind, frags):
    if kind == "S":
        return ints.S[frags[0]][frags[1]]
    if kind == "T":
        return ints.T[frags[0]][frags[1]]
    if kind == "U":
        return ints.U[frags[0]][frags[1]][frags[2]]
    return ints.V[frags[0]][frags[1]][frags[2]][frags[3]]


def diagram(name, rho, charges, n_j0, ints, frags=(0, 1)):
    """Block [i0, i1, j0, j1] of one diagram summed over its permutations
    whose charge condition holds (None if none does).  rho = (rho of
    fragment 0, rho of fragment 1), charges = ((ci0, cj0), (ci1, cj1))."""
    f, js, (opA, ixA), B, terms, dchgs, perms = DIMER[name]
    dchg = [charges[0][0] - charges[0][1], charges[1][0] - charges[1][1]]
    out = None
    for phase, perm in perms:
        if (dchg[perm[0]], dchg[perm[1]]) != tuple(dchgs):
            continue
        a, b = perm
        RA = rho[a][opA][tuple(charges[a])]
        ops = [RA]
        spec = ["ij" + ixA]
        if B is not None:
            RB = rho[b][B[0]][tuple(charges[b])]
            ops.append(RB)
            spec.append("kl" + B[1])
        for kind, roles, idx in terms:
            role_frag = [frags[(a, b)[r]] for r in roles]
            ops.append(np.asarray(_integral(ints, kind, role_frag)))
            spec.append(idx)
        sign = f * ((-1) ** (n_j0 + js) if js is not None else 1)
        if B is None:
            R = np.einsum(",".join(spec) + "->ij", *ops, optimize=True)
            nb = rho[b]["ca"][tuple(charges[b])].shape[0]
            blk = np.einsum("ij,kl->ikjl", R, np.eye(nb))
        else:
            blk = np.einsum(",".join(spec) + "->ikjl", *ops, optimize=True)
        if perm == (1, 0):
            blk = blk.transpose(1, 0, 3, 2)
        blk = phase * sign * blk
        out = blk if out is None else out + blk
    return out


def monomer(rho_m, chg, ints, m, terms=("t00", "u000", "v0000")):
    """t00 + u000 + v0000 of fragment m (ST/SU/SV_1mer_0.py): (n x n)."""
    R = rho_m["ca"][(chg, chg)]
    out = np.zeros(R.shape[:2])
    if "t00" in terms:
        out += np.einsum("ijpq,pq->ij", R, ints.T[m][m])
    if "u000" in terms:
        out += np.einsum("ijpq,pq->ij", R, ints.U[m][m][m])
    if "v0000" in terms and "ccaa" in rho_m and (chg, chg) in rho_m["ccaa"]:
        out += np.einsum("ijpqsr,pqrs->ij", rho_m["ccaa"][(chg, chg)], ints.V[m][m][m][m])
    return out


def dimer_block(rho, charges, n_elec0, ints, names, monomer_terms=()):
    """Sum of the named dimer diagrams (and of the fragment-order-1 monomer
    terms on each fragment, where the other keeps its state:
    XR_term.py:22-77) for one charge combination: [i0, i1, j0, j1]."""
    n_j0 = n_elec0[charges[0][1]] % 2
    ni0, nj0 = rho[0]["ca"][(charges[0][0], charges[0][0])].shape[0], rho[0]["ca"][(charges[0][1], charges[0][1])].shape[0]
    ni1, nj1 = rho[1]["ca"][(charges[1][0], charges[1][0])].shape[0], rho[1]["ca"][(charges[1][1], charges[1][1])].shape[0]
    out = np.zeros((ni0, ni1, nj0, nj1))
    for name in names:
        blk = diagram(name, rho, charges, n_j0, ints)
        if blk is not None:
            out += blk
    if monomer_terms:
        (ci0, cj0), (ci1, cj1) = charges
        if ci1 == cj1 and ci0 == cj0:
            out += np.einsum("ij,kl->ikjl", monomer(rho[0], ci0, ints, 0, monomer_terms), np.eye(ni1))
        if ci0 == cj0 and ci1 == cj1:
            out += np.einsum("ij,kl->ikjl", np.eye(ni0), monomer(rho[1], ci1, ints, 1, monomer_terms))
    return out


def dimer_matrix(rho, n_elec0, ints, names, charge_blocks, monomer_terms=(), identity=False):
    """XR_term.dimer_matrix: rows / columns over the charge blocks (c0, c1)
    in order, each (i0, i1) row-major."""
    ns = [(rho[0]["ca"][(a, a)].shape[0], rho[1]["ca"][(b, b)].shape[0]) for a, b in charge_blocks]
    sizes = [x * y for x, y in ns]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    M = np.zeros((off[-1], off[-1]))
    for x, (ci0, ci1) in enumerate(charge_blocks):
        for y, (cj0, cj1) in enumerate(charge_blocks):
            blk = dimer_block(rho, ((ci0, cj0), (ci1, cj1)), n_elec0, ints, names, monomer_terms)
            M[off[x]:off[x + 1], off[y]:off[y + 1]] = blk.reshape(sizes[x], sizes[y])
    if identity:
        M += np.eye(off[-1])
    return M


def get_xr_H(symm, bior, rho, n_elec0, charges, xr_order):
    """(H1, H2blocked) of get_xr_result.py:86-198 at order 0 or 1 in the
    charge-block order (XR_term layout).  symm / bior: integral objects
    (.S, .T, .U, .V; bior also .V_diff as .Vd)."""
    cb = [(a, b) for a in charges[0] for b in charges[1]]
    if xr_order == 0:
        ints = bior
        H1 = [_monomer_matrix(rho[m], charges[m], ints, m) for m in (0, 1)]
        names = ST2[0] + SU2[0] + SV2[0]
        return H1, dimer_matrix(rho, n_elec0, ints, names, cb)
    if xr_order != 1:
        raise NotImplementedError
    H1 = [_monomer_matrix(rho[m], charges[m], symm, m) for m in (0, 1)]
    S = dimer_matrix(rho, n_elec0, symm, S2[1], cb, identity=True)
    SH = dimer_matrix(rho, n_elec0, symm, ST2[0] + SU2[0], cb, monomer_terms=("t00", "u000"))
    SH += dimer_matrix(rho, n_elec0, _with_S(bior, symm), ST2[1] + SU2[1], cb)
    diff = _Ints(symm.S, bior.T, bior.U, bior.Vd)
    SH += dimer_matrix(rho, n_elec0, diff, SV2[0], cb, monomer_terms=("v0000",))
    SH += dimer_matrix(rho, n_elec0, _with_S(bior, symm), SV2[1], cb)
    H2 = np.linalg.solve(S, SH)
    H2 -= dimer_matrix(rho, n_elec0, symm, [], cb, monomer_terms=("t00", "u000", "v0000"))
    return H1, H2


def xr_matrix(H1, H2, ns0, ns1):
    """The dimer XR matrix in the charge-block layout: H2 plus each monomer
    matrix embedded with the identity on the other fragment (charge blocks
    (c0, c1) in order; ns0 / ns1: states per charge of each fragment)."""
    M = H2.copy()
    off0 = np.concatenate([[0], np.cumsum(ns0)]).astype(int)
    off1 = np.concatenate([[0], np.cumsum(ns1)]).astype(int)
    o = 0
    for i, n0 in enumerate(ns0):
        for j, n1 in enumerate(ns1):
            h0 = H1[0][off0[i]:off0[i + 1], off0[i]:off0[i + 1]]
            h1 = H1[1][off1[j]:off1[j + 1], off1[j]:off1[j + 1]]
            k = n0 * n1
            M[o:o + k, o:o + k] += np.kron(h0, np.eye(n1)) + np.kron(np.eye(n0), h1)
            o += k
    return M


class _Ints:
    def __init__(self, S, T, U, V):
        self.S, self.T, self.U, self.V = S, T, U, V


def _with_S(ints, symm):
    return _Ints(symm.S, ints.T, ints.U, ints.V)


def _monomer_matrix(rho_m, charges, ints, m):
    ns = [rho_m["ca"][(c, c)].shape[0] for c in charges]
    off = np.concatenate([[0], np.cumsum(ns)]).astype(int)
    M = np.zeros((off[-1], off[-1]))
    for a, c in enumerate(charges):
        M[off[a]:off[a + 1], off[a]:off[a + 1]] = monomer(rho_m, c, ints, m)
    return M
