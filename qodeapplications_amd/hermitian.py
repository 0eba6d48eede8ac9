"""Fragment densities and the order-0 XR Hamiltonian in the layout of the
hermitian-XRCC path.

build_tensors (drop-in for Be-states/densities.py:236) keeps one tensor per
operator string and charge pair, rho[op][(bra_chg, ket_chg)] with axes
(i, j, p, q, ...), builds the strings of charge differences 2, 1, 0 (aa;
a, caa; ca, ccaa at xr_order 0) and derives the rest by conjugation:
for bra_chg != ket_chg, <j|op^+|i> = <i|op|j> with op^+ the reversed string
with c <-> a swapped and the index axes reversed (:192-206); for equal
charges the reference computes i >= j and fills j > i the same way (:179,
196-198).  With the 'bra_det' / 'ket_det' options the bra (ket) side is
every determinant (or the rows of dets[chg]) and the strings of charge
differences -1 and -2 are built directly (:99-101, 103-118).

Here every computed tensor comes from xr_density on the GPU (both
triangles directly: the kernels are exact for every (i, j)), and the
conjugate strings are transposed copies of them.  The reference's optional
compression ('compress', 'nat-orbs', 'abs-anti' options through Qode's
compress module) is not built (it raises); thresh, Qode field_op's
coefficient screening, is accepted and not applied (the densities here
are exact), and n_threads is accepted and ignored (all parallelism is on
the GPU).

get_xr_H (drop-in for hermitian-XRCC/get_xr_result.py:45 at xr_order 0)
returns the monomer matrices H1 and the pair matrix H2 of a dimer from two
such density dicts, with the reference's state reorder (:300-353).
"""
import numpy as np
import torch

from .blocks import IntegralCache
from .density import ChargeBlock, conjugate, conjugate_string, density
from .gemm import mm

OPS_XR0 = {2: ("aa",), 1: ("a", "caa"), 0: ("ca", "ccaa")}


def _parse_options(options):
    """The tokens of Be-states/densities.py:_token_parser (:31-67):
    'key' -> True, 'key=v1,v2' -> the values (True/False/int/float/str)."""
    out = {}
    for opt in options or []:
        if "=" not in opt:
            out[opt] = True
            continue
        key, value = opt.split("=")
        vals = []
        for v in value.split(","):
            if v in ("True", "False"):
                vals.append(v == "True")
                continue
            for cast in (int, float):
                try:
                    v = cast(v)
                    break
                except ValueError:
                    pass
            vals.append(v)
        out[key] = vals[0] if len(vals) == 1 and isinstance(vals[0], bool) else tuple(vals)
    return out


def build_tensors(states, n_orbs, n_elec_0, thresh=1e-10, options=None, xr_order=0, dets=None, n_threads=1,
                  *, n_core=0, device=0, op_strings=None):
    """states: {chg: obj with .configs (occupation lists in lexical order,
    core included) and .coeffs (n_states x n_configs)}; n_elec_0: electrons
    of the neutral fragment.  Returns the densities dict of the hermitian
    path (device tensors), with 'n_elec', 'n_states', 'n_states_bra' and
    'KetCoeffs' as Be-states/densities.py:211-231 fills them.

    options: 'bra_det', 'ket_det' are supported; 'compress', 'nat-orbs' and
    'abs-anti' raise NotImplementedError.  xr_order 1 adds 'caaa' (built)
    and 'ccaaa' (5-operator, not built: raises).  dets: {chg: rows of
    coefficients over the charge block's determinants} replacing the
    identity of the _det options.  n_core, device and op_strings (a
    {charge difference: strings} table replacing the order's) are
    keyword-only extensions."""
    opts = _parse_options(options)
    for key in ("compress", "nat-orbs", "abs-anti"):
        if opts.get(key):
            raise NotImplementedError(f"build_tensors option {key!r} (compression through Qode) is not built")
    if op_strings is None:
        op_strings = {2: ["aa"], 1: ["a", "caa"], 0: ["ca", "ccaa"]}
        if xr_order >= 1:
            op_strings[2].append("caaa")
            op_strings[1].append("ccaaa")
        if xr_order >= 2:
            raise NotImplementedError(f"densities for xr_order {xr_order} are not built")
    else:
        op_strings = {k: list(v) for k, v in op_strings.items()}
    bra_det, ket_det = bool(opts.get("bra_det")), bool(opts.get("ket_det"))
    if bra_det or ket_det:
        op_strings[-1] = ["c", "cca"]
        op_strings[-2] = ["cc"]
    dets = dets or {}
    blocks = {chg: ChargeBlock.from_states(st, n_orbs, n_core, device) for chg, st in states.items()}
    dev = next(iter(blocks.values())).coeffs.device

    def side(chg, as_det):
        b = blocks[chg]
        if not as_det:
            return b
        if chg in dets:
            z = torch.as_tensor(np.ascontiguousarray(dets[chg], dtype=np.float64), device=dev)
        else:
            z = torch.eye(b.n_configs, dtype=torch.float64, device=dev)
        return b.with_coeffs(z)

    densities, conj = {}, {}
    for bra_chg in blocks:
        for ket_chg in blocks:
            for op in op_strings.get(bra_chg - ket_chg, ()):
                if op == "ccaaa" or op == "cccaa":
                    raise NotImplementedError("5-operator densities (xr_order 1) are not built")
                rho = density(op, side(bra_chg, bra_det), side(ket_chg, ket_det))
                densities.setdefault(op, {})[(bra_chg, ket_chg)] = rho
                if bra_chg != ket_chg and not (bra_det or ket_det):
                    conj.setdefault(conjugate_string(op), {})[(ket_chg, bra_chg)] = conjugate(rho)
    for op, per in conj.items():
        densities[op] = per
    densities["n_elec"] = {chg: n_elec_0 - chg for chg in states}
    densities["n_states"] = {chg: blocks[chg].n_states for chg in states}
    if bra_det and not ket_det:
        densities["n_states_bra"] = {chg: (len(dets[chg]) if chg in dets else blocks[chg].n_configs)
                                     for chg in states}
    elif ket_det and not bra_det:
        densities["n_states"] = {chg: (len(dets[chg]) if chg in dets else blocks[chg].n_configs)
                                 for chg in states}
        densities["n_states_bra"] = {chg: blocks[chg].n_states for chg in states}
    else:
        densities["n_states_bra"] = densities["n_states"]
    kc = {}
    for chg in states:
        z = blocks[chg].coeffs
        if chg in dets:
            D = torch.as_tensor(np.ascontiguousarray(dets[chg], dtype=np.float64), device=dev)
            z = mm(z, D, transB=1)          # einsum("kp,ip->ik", dets, coeffs)
        kc[(chg, chg)] = z
    densities["KetCoeffs"] = kc
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


def pair_diagrams(rho0, rho1, n_elec0, frags, ints, charges):
    """Sum of the order-0 pair diagrams for the fragments frags = (m0, m1)
    (integral block indices), with density dicts rho0, rho1, charges
    ((ci0, cj0), (ci1, cj1)) and n_elec0 = {charge: electrons} of m0:
    tensor [i0, i1, j0, j1], or None when no diagram connects them."""
    rho = [rho0, rho1]
    dchg = [charges[0][0] - charges[0][1], charges[1][0] - charges[1][1]]
    n_j0 = n_elec0[charges[0][1]] % 2
    out = None
    for name, (dchgs, perms) in _DIMER_DIAGRAMS.items():
        for phase, perm in perms:
            if (dchg[perm[0]], dchg[perm[1]]) != dchgs:
                continue
            a, b = perm
            blk = _diagram(name, frags[a], frags[b], rho[a], rho[b], tuple(charges[a]), tuple(charges[b]),
                           ints, n_j0)
            if perm == (1, 0):
                blk = blk.permute(1, 0, 3, 2)
            blk = phase * blk
            out = blk.contiguous() if out is None else out + blk
    return out


def dimer_block4(system, m0, m1, charges):
    """pair_diagrams for fragments (m0, m1) of an XRSystem."""
    F0, F1 = system.frags[m0], system.frags[m1]
    n_elec0 = {c: F0.n_elec_ref - c for c in F0.charges}
    return pair_diagrams(F0.rho, F1.rho, n_elec0, (m0, m1), system.ints, charges)


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


# ---------------------------------------------------------------------------
# get_xr_H drop-in (hermitian-XRCC/get_xr_result.py:45-355, xr_order 0)
# ---------------------------------------------------------------------------

def _dev(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)


def get_xr_H(ints, dens, xr_order, monomer_charges, bra_det=False, ket_det=False):
    """(H1, H2) of the dimer XR Hamiltonian at order 0 from two density dicts
    of build_tensors: H1[m] (monomer, rows / columns grouped by charge in
    monomer_charges[m] order: XR_term.monomer_matrix), H2 = the pair
    diagrams S2H2 (get_xr_result.py:110-131) reordered from the charge-block
    order to the product order i0 * n1 + i1 (fragment states numbered in
    charge-block order), as the reference's bra / ket maps do (:300-353).

    ints: the reference's (symm_ints, bior_ints, nuc_rep) triple or one
    integrals object; at order 0 only the biorthogonal integrals' T, U, V
    blocks enter (equal to the plain ones for orthonormal orbitals), indexed
    [m, n], [a, m, n], [m, n, o, p].  Densities may be device tensors or
    arrays.  xr_order >= 1 and the bra_det / ket_det variants (which the
    reference uses only inside state_gradients; grad.herm_state_gradient
    gives that gradient) raise NotImplementedError."""
    if xr_order != 0:
        raise NotImplementedError(f"get_xr_H at xr_order {xr_order} is not built")
    if bra_det or ket_det:
        raise NotImplementedError("get_xr_H with bra_det / ket_det (state-gradient densities) is not built")
    bior = ints[1] if isinstance(ints, (tuple, list)) else ints
    sample = dens[0]["ca"][next(iter(dens[0]["ca"]))]
    device = sample.device if isinstance(sample, torch.Tensor) else torch.device("cuda", 0)
    ic = IntegralCache(bior, device)
    rho = [{op: {k: _dev(v, device) for k, v in per.items()} for op, per in d.items()
            if op not in ("n_elec", "n_states", "n_states_bra", "KetCoeffs")} for d in dens]
    H1 = []
    for m in (0, 1):
        nb = [dens[m]["n_states_bra"][c] for c in monomer_charges[m]]
        nk = [dens[m]["n_states"][c] for c in monomer_charges[m]]
        ob = np.concatenate([[0], np.cumsum(nb)]).astype(int)
        ok = np.concatenate([[0], np.cumsum(nk)]).astype(int)
        out = torch.zeros(int(ob[-1]), int(ok[-1]), dtype=torch.float64, device=device)
        h = (ic.T(m, m) + ic.U(m, m, m)).reshape(-1, 1)
        V = ic.V(m, m, m, m)
        Vp = V.permute(0, 1, 3, 2).contiguous().reshape(-1, 1)      # [p, q, s, r] = V[p, q, r, s]
        for a, c in enumerate(monomer_charges[m]):
            R = rho[m]["ca"][(c, c)]
            blk = precontract(R, h).reshape(R.shape[0], R.shape[1])            # t00 + u000
            if "ccaa" in rho[m] and (c, c) in rho[m]["ccaa"]:
                R4 = rho[m]["ccaa"][(c, c)]
                blk = blk + precontract(R4, Vp).reshape(R4.shape[0], R4.shape[1])   # v0000
            out[ob[a]:ob[a + 1], ok[a]:ok[a + 1]] = blk
        H1.append(out)
    blocks = [(c0, c1) for c0 in monomer_charges[0] for c1 in monomer_charges[1]]
    nb = [dens[0]["n_states_bra"][a] * dens[1]["n_states_bra"][b] for a, b in blocks]
    nk = [dens[0]["n_states"][a] * dens[1]["n_states"][b] for a, b in blocks]
    ob = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    ok = np.concatenate([[0], np.cumsum(nk)]).astype(int)
    S2H2 = torch.zeros(int(ob[-1]), int(ok[-1]), dtype=torch.float64, device=device)
    for x, (ci0, ci1) in enumerate(blocks):
        for y, (cj0, cj1) in enumerate(blocks):
            blk = pair_diagrams(rho[0], rho[1], dens[0]["n_elec"], (0, 1), ic, ((ci0, cj0), (ci1, cj1)))
            if blk is not None:
                S2H2[ob[x]:ob[x + 1], ok[y]:ok[y + 1]] = blk.reshape(nb[x], nk[y])
    # charge-block order -> product order (get_xr_result.py:300-353)
    def mapping(key):
        d0 = [dens[0][key][c] for c in monomer_charges[0]]
        d1 = [dens[1][key][c] for c in monomer_charges[1]]
        n1 = sum(d1)
        m2 = np.empty(sum(d0) * n1, dtype=np.int64)
        idx, b0 = 0, 0
        for a in d0:
            b1 = 0
            for b in d1:
                for i in range(a):
                    for j in range(b):
                        m2[(b0 + i) * n1 + b1 + j] = idx
                        idx += 1
                b1 += b
            b0 += a
        return torch.as_tensor(m2, device=device)
    mb, mk = mapping("n_states_bra"), mapping("n_states")
    H2 = S2H2.index_select(0, mb).index_select(1, mk)
    return H1, H2
