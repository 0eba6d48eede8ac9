"""Fragment densities in the layout of the hermitian-XRCC path.

Be-states/densities.py:_build_tensors (:80-231) keeps one tensor per
operator string and charge pair, rho[op][(bra_chg, ket_chg)] with axes
(i, j, p, q, ...), builds the strings of charge differences 2, 1, 0 (aa;
a, caa; ca, ccaa at xr_order 0) and derives the rest by conjugation:
for bra_chg != ket_chg, <j|op^+|i> = <i|op|j> with op^+ the reversed string
with c <-> a swapped and the index axes reversed (:192-206); for equal
charges it computes i >= j and fills j > i the same way (:179, 196-198).

Here every computed tensor comes from xr_density on the GPU (both
triangles directly: the kernels are exact for every (i, j)), and the
conjugate strings are transposed copies of them.  The reference's optional
compression (thresh, natural orbitals, Qode field_op) is not applied.
"""
import numpy as np
import torch

from .density import ChargeBlock, conjugate, conjugate_string, density

OPS_XR0 = {2: ("aa",), 1: ("a", "caa"), 0: ("ca", "ccaa")}


def build_tensors(states, n_orbs, n_elec_0, n_core=0, device=0, op_strings=OPS_XR0):
    """states: {chg: obj with .configs (lexical, core included) and .coeffs
    (n_states x n_configs)}; n_elec_0: electrons of the neutral fragment.
    Returns the densities dict of the hermitian path, device tensors."""
    blocks = {chg: ChargeBlock.from_states(st, n_orbs, n_core, device) for chg, st in states.items()}
    densities, conj = {}, {}
    for bra_chg, bb in blocks.items():
        for ket_chg, kb in blocks.items():
            for op in op_strings.get(bra_chg - ket_chg, ()):
                rho = density(op, bb, kb)
                densities.setdefault(op, {})[(bra_chg, ket_chg)] = rho
                if bra_chg != ket_chg:
                    conj.setdefault(conjugate_string(op), {})[(ket_chg, bra_chg)] = conjugate(rho)
    for op, per in conj.items():
        densities[op] = per
    densities["n_elec"] = {chg: n_elec_0 - chg for chg in states}
    densities["n_states"] = {chg: blocks[chg].n_states for chg in states}
    densities["n_states_bra"] = densities["n_states"]
    densities["KetCoeffs"] = {(chg, chg): blocks[chg].coeffs for chg in states}
    return densities


# ---------------------------------------------------------------------------
# Order-0 XR terms in the hermitian formulation
# (hermitian-XRCC/diagrams/S{T,U,V}_{1,2}mer_0.py with the catalogs of
# S{T,U,V}_diagrams.py; blocks as in XR_term.py:22-165 and
# get_xr_result.py:86-132).  Each diagram is a bilinear form
# s * rhoA W rhoB^T in two fragments' densities: the first product is the
# reference's precontracted intermediate (precontract.py:26-94, e.g.
# "ca0pr_Vp1r1" = sum_pr ca0[i0,j0,p,r] V[p,q,r,s]), the second contracts it
# with the other fragment, both on libxrhip's FP64 GEMM.
# ---------------------------------------------------------------------------
from .gemm import mm


def precontract(rho, W):
    """(nI*nJ, cols) intermediate: rho flattened over its state pair times W
    (rows of W run over rho's orbital indices in rho's order)."""
    return mm(rho.reshape(rho.shape[0] * rho.shape[1], -1), W)


def _flat(R):
    return R.reshape(R.shape[0] * R.shape[1], -1)


# name: (Dchg of role A, Dchg of role B), permutations [(phase, (roleA, roleB))]
_DIMER_DIAGRAMS = {
    "v0101": ((0, 0), [(+1, (0, 1))]),
    "v0001": ((-1, +1), [(+1, (0, 1)), (-1, (1, 0))]),
    "v0100": ((+1, -1), [(+1, (0, 1)), (-1, (1, 0))]),
    "v0011": ((-2, +2), [(+1, (0, 1)), (+1, (1, 0))]),
    "u100": ((0, 0), [(+1, (0, 1)), (+1, (1, 0))]),
    "u001": ((-1, +1), [(+1, (0, 1)), (-1, (1, 0))]),
    "u101": ((-1, +1), [(+1, (0, 1)), (-1, (1, 0))]),
    "t01": ((-1, +1), [(+1, (0, 1)), (-1, (1, 0))]),
}


def _diagram(name, A, B, rA, rB, kA, kB, ints, n_j0):
    """Block [iA, iB, jA, jB] of one diagram with roles A, B (fragment
    indices), densities rA, rB (dicts op -> {charges: tensor}) at charge
    keys kA, kB."""
    dA = ints.V(A, A, A, A).shape[0]
    dB = ints.V(B, B, B, B).shape[0]
    if name == "u100":
        R = rA["ca"][kA]
        nIB = rB["ca"][kB].shape[0]
        e = precontract(R, ints.U(B, A, A).reshape(-1, 1)).reshape(R.shape[0], R.shape[1])
        eye = torch.eye(nIB, dtype=torch.float64, device=R.device)
        return e.view(e.shape[0], 1, e.shape[1], 1) * eye.view(1, nIB, 1, nIB)   # [a, c, b, d]
    if name == "v0101":
        X, Y, s = rA["ca"][kA], rB["ca"][kB], 4.0
        W = ints.V(A, B, A, B).permute(0, 2, 1, 3).reshape(dA * dA, dB * dB)
    elif name == "v0001":
        X, Y, s = rA["cca"][kA], rB["a"][kB], 2.0 * (-1) ** (n_j0 + 1)
        W = ints.V(A, A, A, B).reshape(dA ** 3, dB)
    elif name == "v0100":
        X, Y, s = rA["caa"][kA], rB["c"][kB], 2.0 * (-1) ** n_j0
        W = ints.V(A, B, A, A).permute(0, 3, 2, 1).reshape(dA ** 3, dB)
    elif name == "v0011":
        X, Y, s = rA["cc"][kA], rB["aa"][kB], 1.0
        W = ints.V(A, A, B, B).permute(0, 1, 3, 2).reshape(dA * dA, dB * dB)
    else:
        X, Y = rA["c"][kA], rB["a"][kB]
        s = float((-1) ** n_j0)
        W = {"u001": ints.U(A, A, B), "u101": ints.U(B, A, B), "t01": ints.T(A, B)}[name]
    P = precontract(X, W.contiguous())               # (iA jA) x (B indices)
    blk = mm(P, _flat(Y), transB=1, alpha=s)          # (iA jA) x (iB jB)
    nIA, nJA, nIB, nJB = X.shape[0], X.shape[1], Y.shape[0], Y.shape[1]
    return blk.view(nIA, nJA, nIB, nJB).permute(0, 2, 1, 3)


def dimer_block4(system, m0, m1, charges):
    """Sum of the order-0 pair diagrams for fragments (m0, m1) and charges
    ((ci0, cj0), (ci1, cj1)): tensor [i0, i1, j0, j1], or None."""
    F = [system.frags[m0], system.frags[m1]]
    rho = [F[0].rho, F[1].rho]
    frags = (m0, m1)
    dchg = [charges[0][0] - charges[0][1], charges[1][0] - charges[1][1]]
    n_j0 = (F[0].n_elec_ref - charges[0][1]) % 2
    out = None
    for name, (dchgs, perms) in _DIMER_DIAGRAMS.items():
        for phase, perm in perms:
            if (dchg[perm[0]], dchg[perm[1]]) != dchgs:
                continue
            a, b = perm
            blk = _diagram(name, frags[a], frags[b], rho[a], rho[b], tuple(charges[a]), tuple(charges[b]),
                           system.ints, n_j0)
            if perm == (1, 0):
                blk = blk.permute(1, 0, 3, 2)
            blk = phase * blk
            out = blk.contiguous() if out is None else out + blk
    return out


def monomer_block(system, m, ci, cj):
    """t00 + u000 + v0000 for fragment m between charges ci, cj (nonzero
    only for ci == cj): (ni x nj)."""
    if ci != cj:
        return None
    F = system.frags[m]
    R = F.rho["ca"][(ci, cj)]
    h = (system.ints.T(m, m) + system.ints.U(m, m, m)).reshape(-1, 1)
    e1 = precontract(R, h).reshape(R.shape[0], R.shape[1])
    return e1 + F.E2[ci]                  # v0000 = sum ccaa V, the sigma-form monomer term


def monomer_matrix(system, m, charges=None):
    """XR_term.monomer_matrix for order 0: states grouped by charge in the
    order of `charges`."""
    system._need()
    F = system.frags[m]
    charges = list(charges) if charges is not None else list(F.charges)
    sizes = [F.n_states[c] for c in charges]
    n = sum(sizes)
    out = torch.zeros(n, n, dtype=torch.float64, device=system.device)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    for a, ci in enumerate(charges):
        blk = monomer_block(system, m, ci, ci)
        out[off[a]:off[a + 1], off[a]:off[a + 1]] = blk
    return out


def dimer_matrix(system, m0, m1, charge_blocks=None):
    """XR_term.dimer_matrix for the order-0 pair diagrams (the S2H2 of
    get_xr_result.py:121-131): rows and columns run over the charge blocks
    (chg0, chg1) in the given order, each over (i0, i1) row-major."""
    system._need()
    F0, F1 = system.frags[m0], system.frags[m1]
    if charge_blocks is None:
        charge_blocks = [(a, b) for a in F0.charges for b in F1.charges]
    sizes = [F0.n_states[a] * F1.n_states[b] for a, b in charge_blocks]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    n = int(off[-1])
    out = torch.zeros(n, n, dtype=torch.float64, device=system.device)
    for x, (ci0, ci1) in enumerate(charge_blocks):
        for y, (cj0, cj1) in enumerate(charge_blocks):
            blk = dimer_block4(system, m0, m1, ((ci0, cj0), (ci1, cj1)))
            if blk is None:
                continue
            out[off[x]:off[x + 1], off[y]:off[y + 1]] = blk.reshape(sizes[x], sizes[y])
    return out


def product_order(system, m0, m1, charge_blocks=None):
    """Index map from the charge-block order of dimer_matrix to the product
    order of the general path (position_0 * d_1 + position_1)."""
    F0, F1 = system.frags[m0], system.frags[m1]
    if charge_blocks is None:
        charge_blocks = [(a, b) for a in F0.charges for b in F1.charges]
    idx = []
    for a, b in charge_blocks:
        p0 = F0.pos[a].cpu().numpy()
        p1 = F1.pos[b].cpu().numpy()
        idx.append((p0[:, None] * F1.d + p1[None, :]).ravel())
    return torch.as_tensor(np.concatenate(idx), device=system.device)
