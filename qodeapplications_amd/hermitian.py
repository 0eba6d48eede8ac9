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
import torch

from .density import ChargeBlock, density

OPS_XR0 = {2: ("aa",), 1: ("a", "caa"), 0: ("ca", "ccaa")}


def conjugate_string(op):
    """The string of op^+: reversed, with c and a swapped."""
    return op[::-1].translate(str.maketrans("ca", "ac"))


def conjugate(rho):
    """<j| op^+ |i>[rev idx] from <i| op |j>[idx]: tensor (nI, nJ, d, ..., d)
    -> (nJ, nI, d, ..., d) with the index axes reversed."""
    return rho.permute(1, 0, *range(rho.dim() - 1, 1, -1)).contiguous()


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
