"""Fragment densities and the order-0 and order-1 XR Hamiltonian in the
layout of the hermitian-XRCC path.

build_tensors (drop-in for Be-states/densities.py:236) keeps one tensor per
operator string and charge pair, rho[op][(bra_chg, ket_chg)] with axes
(i, j, p, q, ...), builds the strings of charge differences 2, 1, 0 (aa;
a, caa; ca, ccaa at xr_order 0; with caaa and ccaaa at xr_order 1) and
derives the rest by conjugation:
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

get_xr_H (drop-in for hermitian-XRCC/get_xr_result.py:45 at xr_order 0 and
1) returns the monomer matrices H1 and the pair matrix H2 of a dimer from
two such density dicts, with the reference's state reorder (:300-353).
Every diagram (the order-0 ones and the order-1 s01 terms of the overlap
expansion) is a chain of libxrhip GEMMs over its densities and integral
blocks.
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
    'abs-anti' raise NotImplementedError.  xr_order 1 adds 'caaa' and the
    5-operator 'ccaaa' (with their conjugates ccca, cccaa); xr_order 2
    (6-operator strings) raises.  dets: {chg: rows of
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
# XR terms in the hermitian formulation, orders 0 and 1 of the overlap
# expansion (hermitian-XRCC/diagrams/S{T,U,V}_{1,2}mer_{0,1}.py and
# S_2mer_1.py with the catalogs of S{,T,U,V}_diagrams.py; blocks as in
# XR_term.py:22-165 and get_xr_result.py:86-198).  Each diagram is a
# product of two fragments' densities and one or two integral blocks: each
# integral is contracted into the density it shares the most indices with
# (the reference's precontracted intermediates, precontract.py:26-94, e.g.
# "ca0pr_Vp1r1" = sum_pr ca0[i0,j0,p,r] V[p,q,r,s]), then the two
# intermediates are contracted, every step one libxrhip FP64 GEMM.
# ---------------------------------------------------------------------------

def precontract(rho, W):
    """(nI*nJ, cols) intermediate: rho flattened over its state pair times W
    (rows of W run over rho's orbital indices in rho's order)."""
    return mm(rho.reshape(rho.shape[0] * rho.shape[1], -1), W)


def _flat(R):
    return R.reshape(R.shape[0] * R.shape[1], -1)


# Diagram table (hermitian-XRCC/diagrams/S_2mer_1.py and
# S{T,U,V}_2mer_{0,1}.py, the commented einsum form of each; charge
# conditions and fragment permutations from the S{,T,U,V}_diagrams.py
# catalogs).  name: (prefactor, exponent e of the sign (-1)^(n_j0 + e) or
# None, (string, indices) of role 0, (string, indices) of role 1 or None for
# a delta on role 1's states, integral factors (kind, roles, indices), the
# roles' charge changes, [(phase, permutation)]).  Integral kinds: S and T
# with roles (x, y); U with (nucleus, x, y); V with (w, x, y, z).
_P2 = [(+1, (0, 1)), (+1, (1, 0))]
_PM = [(+1, (0, 1)), (-1, (1, 0))]
DIAGRAMS = {
    "t01":   (1, 0, ("c", "p"), ("a", "q"), [("T", (0, 1), "pq")], (-1, 1), _PM),
    "u100":  (1, None, ("ca", "pq"), None, [("U", (1, 0, 0), "pq")], (0, 0), _P2),
    "u001":  (1, 0, ("c", "p"), ("a", "q"), [("U", (0, 0, 1), "pq")], (-1, 1), _PM),
    "u101":  (1, 0, ("c", "p"), ("a", "q"), [("U", (1, 0, 1), "pq")], (-1, 1), _PM),
    "v0101": (4, None, ("ca", "pr"), ("ca", "qs"), [("V", (0, 1, 0, 1), "pqrs")], (0, 0), [(+1, (0, 1))]),
    "v0001": (2, 1, ("cca", "pqr"), ("a", "s"), [("V", (0, 0, 0, 1), "pqrs")], (-1, 1), _PM),
    "v0100": (2, 0, ("caa", "psr"), ("c", "q"), [("V", (0, 1, 0, 0), "pqrs")], (1, -1), _PM),
    "v0011": (1, None, ("cc", "pq"), ("aa", "sr"), [("V", (0, 0, 1, 1), "pqrs")], (-2, 2), _P2),
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
for _nuc in (0, 1):       # the s01-U diagrams: the s01-T ones with U of nucleus 0 or 1 (SU_2mer_1.py)
    for _t, _u in (("s01t10", f"s01u{_nuc}10"), ("s01t00", f"s01u{_nuc}00"), ("s01t11", f"s01u{_nuc}11"),
                   ("s01t01", f"s01u{_nuc}01")):
        _f, _js, _A, _B, _ints, _dch, _perms = DIAGRAMS[_t]
        DIAGRAMS[_u] = (_f, _js, _A, _B, [_ints[0], ("U", (_nuc,) + _ints[1][1], _ints[1][2])], _dch, _perms)

# diagram lists by order of the overlap expansion (diagram_lists.py:11-73)
ST2 = {0: ["t01"], 1: ["s01t10", "s01t00", "s01t11", "s01t01"]}
SU2 = {0: ["u100", "u001", "u101"],
       1: ["s01u010", "s01u000", "s01u011", "s01u001", "s01u110", "s01u100", "s01u111", "s01u101"]}
SV2 = {0: ["v0101", "v0001", "v0100", "v0011"],
       1: ["s01v0100", "s01v1101", "s01v0000", "s01v0101", "s01v1100", "s01v1111", "s01v0001", "s01v0111"]}
S2 = {1: ["s01"]}
_DIMER_DIAGRAMS = ST2[0] + SU2[0] + SV2[0]


def _tdot(X, xl, Y, yl):
    """X (axes named by the letters xl) times Y (yl) contracted over their
    shared letters, as one libxrhip GEMM: result axes X's rest, then Y's."""
    shared = [c for c in xl if c in yl]
    xr = [c for c in xl if c not in shared]
    yr = [c for c in yl if c not in shared]
    dim = dict(zip(xl, X.shape))
    dim.update(zip(yl, Y.shape))
    prod = lambda cs: int(np.prod([dim[c] for c in cs])) if cs else 1
    Xp = X.permute(*[xl.index(c) for c in xr + shared]).reshape(prod(xr), prod(shared))
    Yp = Y.permute(*[yl.index(c) for c in shared + yr]).reshape(prod(shared), prod(yr))
    R = mm(Xp.contiguous(), Yp.contiguous())
    return R.view(*[dim[c] for c in xr + yr]), "".join(xr + yr)


def _blk(arr, *idx):
    """Fragment block of an integral array indexed [m, n] or [m][n]."""
    try:
        return arr[idx]
    except (TypeError, IndexError, KeyError):
        for i in idx:
            arr = arr[i]
        return arr


class _Ints:
    """Device copies of the fragment-blocked integral blocks of one
    integrals object (.S, .T, .U, .V; V may come from another attribute)."""

    def __init__(self, ints, device, V=None, S=None):
        self.src, self.device = ints, device
        self.Vsrc = V if V is not None else getattr(ints, "V", None)
        self.Ssrc = S if S is not None else getattr(ints, "S", None)
        self._c = {}

    def get(self, kind, key):
        k = (kind,) + tuple(key)
        t = self._c.get(k)
        if t is None:
            arr = {"S": self.Ssrc, "T": getattr(self.src, "T", None), "U": getattr(self.src, "U", None),
                   "V": self.Vsrc}[kind]
            t = self._c[k] = _dev(_blk(arr, *key), self.device).contiguous()
        return t

    # the IntegralCache interface the order-0 system path uses
    def T(self, *k):
        return self.get("T", k)

    def U(self, *k):
        return self.get("U", k)

    def V(self, *k):
        return self.get("V", k)


def _diagram(name, frags, rho, charges, n_j0, ints):
    """Block [i0, i1, j0, j1] of one diagram summed over the permutations
    whose charge condition holds (None if none does): every precontraction
    (density x integral) and the final product are libxrhip GEMMs."""
    f, js, (opA, ixA), B, terms, dchgs, perms = DIAGRAMS[name]
    dchg = (charges[0][0] - charges[0][1], charges[1][0] - charges[1][1])
    out = None
    for phase, perm in perms:
        if (dchg[perm[0]], dchg[perm[1]]) != tuple(dchgs):
            continue
        a, b = perm
        A = rho[a][opA][tuple(charges[a])]
        PA, la = A, "IJ" + ixA
        PB, lb = None, None
        if B is not None:
            PB, lb = rho[b][B[0]][tuple(charges[b])], "KL" + B[1]
        for kind, roles, idx in terms:
            W = ints.get(kind, [frags[(a, b)[r]] for r in roles])
            na = len(set(idx) & set(la))
            nb = len(set(idx) & set(lb)) if lb else -1
            if na >= nb:
                PA, la = _tdot(PA, la, W, idx)
            else:
                PB, lb = _tdot(PB, lb, W, idx)
        sign = float(f * ((-1) ** (n_j0 + js) if js is not None else 1)) * phase
        if PB is None:                                 # delta on role b's states
            R = PA.permute(*[la.index(c) for c in "IJ"])
            nb_ = rho[b]["ca"][(charges[b][0], charges[b][0])].shape[0]
            eye = torch.eye(nb_, dtype=torch.float64, device=R.device)
            blk = sign * R[:, None, :, None] * eye[None, :, None, :]
        else:
            R, lr = _tdot(PA, la, PB, lb)
            blk = sign * R.permute(*[lr.index(c) for c in "IKJL"])
        if perm == (1, 0):
            blk = blk.permute(1, 0, 3, 2)
        out = blk.contiguous() if out is None else out + blk
    return out


def pair_diagrams(rho0, rho1, n_elec0, frags, ints, charges, names=None):
    """Sum of the named dimer diagrams (default: the order-0 ones) for the
    fragments frags = (m0, m1) (integral block indices), density dicts
    rho0, rho1, charges ((ci0, cj0), (ci1, cj1)) and n_elec0 = {charge:
    electrons} of m0: tensor [i0, i1, j0, j1], or None when no diagram
    connects them.  ints: an _Ints (or IntegralCache-like with .get)."""
    if not hasattr(ints, "get"):
        ints = _CacheAdapter(ints)
    rho = [rho0, rho1]
    n_j0 = n_elec0[charges[0][1]] % 2
    out = None
    for name in (names if names is not None else _DIMER_DIAGRAMS):
        blk = _diagram(name, frags, rho, charges, n_j0, ints)
        if blk is not None:
            out = blk if out is None else out + blk
    return out


class _CacheAdapter:
    """blocks.IntegralCache seen through the _Ints.get interface."""

    def __init__(self, cache):
        self.c = cache

    def get(self, kind, key):
        return getattr(self.c, kind)(*key).contiguous()


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
# get_xr_H drop-in (hermitian-XRCC/get_xr_result.py:45-355, xr_order 0 and 1)
# ---------------------------------------------------------------------------

def _dev(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)


def _monomer_terms(rho_m, c, ints, m, terms):
    """t00 + u000 + v0000 (those named) of fragment m at charge c: (n x n)."""
    R = rho_m["ca"][(c, c)]
    out = torch.zeros(R.shape[0], R.shape[1], dtype=torch.float64, device=R.device)
    h = None
    if "t00" in terms:
        h = ints.get("T", (m, m))
    if "u000" in terms:
        u = ints.get("U", (m, m, m))
        h = u if h is None else h + u
    if h is not None:
        out += precontract(R, h.reshape(-1, 1)).reshape(R.shape[0], R.shape[1])
    if "v0000" in terms and "ccaa" in rho_m and (c, c) in rho_m["ccaa"]:
        R4 = rho_m["ccaa"][(c, c)]
        Vp = ints.get("V", (m, m, m, m)).permute(0, 1, 3, 2).contiguous().reshape(-1, 1)  # [p,q,s,r] = V[p,q,r,s]
        out += precontract(R4, Vp).reshape(R4.shape[0], R4.shape[1])
    return out


def _dimer_matrix(rho, n_elec0, ints, names, blocks, nb, nk, monomer=()):
    """XR_term.dimer_matrix (XR_term.py:77-165) over the charge blocks: the
    named dimer diagrams and the fragment-order-1 monomer terms (each
    fragment's, where the other keeps its state)."""
    ob = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    ok = np.concatenate([[0], np.cumsum(nk)]).astype(int)
    dev = next(iter(rho[0]["ca"].values())).device
    M = torch.zeros(int(ob[-1]), int(ok[-1]), dtype=torch.float64, device=dev)
    for x, (ci0, ci1) in enumerate(blocks):
        for y, (cj0, cj1) in enumerate(blocks):
            ch = ((ci0, cj0), (ci1, cj1))
            blk = pair_diagrams(rho[0], rho[1], n_elec0, (0, 1), ints, ch, names) if names else None
            if monomer and ci0 == cj0 and ci1 == cj1:
                m0 = _monomer_terms(rho[0], ci0, ints, 0, monomer)
                m1 = _monomer_terms(rho[1], ci1, ints, 1, monomer)
                e0 = torch.eye(m0.shape[0], dtype=torch.float64, device=dev)
                e1 = torch.eye(m1.shape[0], dtype=torch.float64, device=dev)
                emb = m0[:, None, :, None] * e1[None, :, None, :] + e0[:, None, :, None] * m1[None, :, None, :]
                blk = emb if blk is None else blk + emb
            if blk is not None:
                M[ob[x]:ob[x + 1], ok[y]:ok[y + 1]] = blk.reshape(nb[x], nk[y])
    return M


def get_xr_H(ints, dens, xr_order, monomer_charges, bra_det=False, ket_det=False):
    """(H1, H2) of the dimer XR Hamiltonian (hermitian-XRCC/get_xr_result.py
    :45-355) from two density dicts of build_tensors: H1[m] (monomer, rows /
    columns grouped by charge in monomer_charges[m] order:
    XR_term.monomer_matrix) and H2 reordered from the charge-block order to
    the product order i0 * n1 + i1 (fragment states numbered in charge-block
    order), as the reference's bra / ket maps do (:300-353).

    ints: the reference's (symm_ints, bior_ints, nuc_rep) triple, or one
    integrals object used as both; blocks indexed [m, n] (S, T), [a, m, n]
    (U), [m, n, o, p] (V), fragment-blocked spin orbitals; bior_ints carries
    V_diff (or Vd) for order 1.

      order 0 (:86-132): H1 and the pair diagrams S2H2 from the
        biorthogonal T, U, V;
      order 1 (:133-198): H1 from the symmetric integrals;
        S2 = 1 + s01 (symmetric S), S2H2 = the order-0 T / U diagrams and
        monomer terms (symmetric) + the s01-T / s01-U diagrams
        (biorthogonal T, U) + the order-0 V diagrams and monomer term
        (V_diff) + the s01-V diagrams (biorthogonal V);
        H2 = S2^-1 S2H2 - the symmetric monomer terms.

    Densities may be device tensors or arrays (order 1 needs caaa, ccca,
    ccaaa, cccaa: build_tensors(..., xr_order=1)).  The bra_det / ket_det
    variants (which the reference uses only inside state_gradients;
    grad.herm_state_gradient gives that gradient) raise
    NotImplementedError, as does xr_order 2."""
    if xr_order not in (0, 1):
        raise NotImplementedError(f"get_xr_H at xr_order {xr_order} is not built")
    if bra_det or ket_det:
        raise NotImplementedError("get_xr_H with bra_det / ket_det (state-gradient densities) is not built")
    symm, bior = (ints[0], ints[1]) if isinstance(ints, (tuple, list)) else (ints, ints)
    sample = dens[0]["ca"][next(iter(dens[0]["ca"]))]
    device = sample.device if isinstance(sample, torch.Tensor) else torch.device("cuda", 0)
    rho = [{op: {k: _dev(v, device).contiguous() for k, v in per.items()} for op, per in d.items()
            if op not in ("n_elec", "n_states", "n_states_bra", "KetCoeffs")} for d in dens]
    n_elec0 = dens[0]["n_elec"]
    blocks = [(c0, c1) for c0 in monomer_charges[0] for c1 in monomer_charges[1]]
    nb = [dens[0]["n_states_bra"][a] * dens[1]["n_states_bra"][b] for a, b in blocks]
    nk = [dens[0]["n_states"][a] * dens[1]["n_states"][b] for a, b in blocks]
    if xr_order == 0:
        Ib = _Ints(bior, device)
        H1 = [_monomer_matrix(rho[m], dens[m], monomer_charges[m], Ib, m) for m in (0, 1)]
        H2b = _dimer_matrix(rho, n_elec0, Ib, _DIMER_DIAGRAMS, blocks, nb, nk)
    else:
        Sb = getattr(symm, "S")
        Vd = getattr(bior, "V_diff", None)
        if Vd is None:
            Vd = getattr(bior, "Vd")
        Is = _Ints(symm, device)
        Ib = _Ints(bior, device, S=Sb)
        Id = _Ints(bior, device, V=Vd, S=Sb)
        H1 = [_monomer_matrix(rho[m], dens[m], monomer_charges[m], Is, m) for m in (0, 1)]
        Smat = _dimer_matrix(rho, n_elec0, Is, S2[1], blocks, nb, nk)
        Smat += torch.eye(Smat.shape[0], dtype=torch.float64, device=device)
        SH = _dimer_matrix(rho, n_elec0, Is, ST2[0] + SU2[0], blocks, nb, nk, monomer=("t00", "u000"))
        SH += _dimer_matrix(rho, n_elec0, Ib, ST2[1] + SU2[1], blocks, nb, nk)
        SH += _dimer_matrix(rho, n_elec0, Id, SV2[0], blocks, nb, nk, monomer=("v0000",))
        SH += _dimer_matrix(rho, n_elec0, Ib, SV2[1], blocks, nb, nk)
        H2b = torch.linalg.solve(Smat, SH)
        H2b -= _dimer_matrix(rho, n_elec0, Is, [], blocks, nb, nk, monomer=("t00", "u000", "v0000"))
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
    H2 = H2b.index_select(0, mb).index_select(1, mk)
    return H1, H2


def _monomer_matrix(rho_m, dens_m, charges, ints, m):
    """XR_term.monomer_matrix with t00 + u000 + v0000."""
    nb = [dens_m["n_states_bra"][c] for c in charges]
    nk = [dens_m["n_states"][c] for c in charges]
    ob = np.concatenate([[0], np.cumsum(nb)]).astype(int)
    ok = np.concatenate([[0], np.cumsum(nk)]).astype(int)
    dev = next(iter(rho_m["ca"].values())).device
    out = torch.zeros(int(ob[-1]), int(ok[-1]), dtype=torch.float64, device=dev)
    for a, c in enumerate(charges):
        out[ob[a]:ob[a + 1], ok[a]:ok[a + 1]] = _monomer_terms(rho_m, c, ints, m, ("t00", "u000", "v0000"))
    return out
