"""Drop-in for general-XRCC/build_density_tensors.py.

Same names, arguments, return values and error behaviour for valid input,
computed by the HIP library on the GPU.  Invalid determinant lists, which
the reference silently mis-indexes (general-XRCC/density_tensors.c:49-62),
raise XRError here.
"""
import math

import numpy as np
import torch

from .density import OPS_BY_CHG_DIFF, charge_blocks, density, monomer_2e


def FCIcomboMat(n_elec, n_spin_orbs):
    """Matrix of C(j, i) for n_elec rows and n_spin_orbs columns, upper
    triangle only (general-XRCC/build_density_tensors.py:31-37)."""
    out = np.zeros((n_elec, n_spin_orbs), dtype=np.int64)
    for i in range(n_elec):
        for j in range(i, n_spin_orbs):
            out[i, j] = math.comb(j, i)
    return out


def numpy_storage_to_lists(nparray_1d, n_bra, n_ket, n_spin_orbs, n_ops):
    """Splits flat (I, J, idx) storage into per-(I, J) arrays
    (general-XRCC/build_density_tensors.py:42-53)."""
    tensor_shape = [n_spin_orbs] * n_ops
    tensor_size = n_spin_orbs ** n_ops
    if tensor_size * n_bra * n_ket != nparray_1d.shape[0]:
        raise ValueError
    blocks = np.asarray(nparray_1d, dtype=np.float64).reshape([n_bra, n_ket] + tensor_shape)
    return [[blocks[i, j] for j in range(n_ket)] for i in range(n_bra)]


def build_density_tensors(z_lists, n_orbs, Vints, n_core, n_threads=1):
    """Transition densities of one fragment (general-XRCC/build_density_tensors.py:57-160).

    z_lists: {charge: states} with states.configs (n_configs x n_elec,
    lexical order) and states.coeffs (n_states x n_configs).  Returns
    (densities, total_size): densities[op][(bra_chg, ket_chg)][i][j] is a
    numpy array of shape (2*n_orbs,)*len(op) for op in aa a caa ca c cca cc,
    and a float  sum V[p,q,r,s] <i|p+q+sr|j>  for 'ccaa' (:128-133, here
    computed without forming the 4-index tensor).  total_size counts the
    stored floats like the reference.  n_threads is accepted and unused, as
    in the reference (density_tensors.c:171).
    """
    blocks = charge_blocks(z_lists, n_orbs, n_core)
    dev = next(iter(blocks.values())).device if blocks else torch.device("cuda", 0)
    V = torch.as_tensor(np.ascontiguousarray(Vints, dtype=np.float64), device=dev)
    densities = {op: {} for op in ("aa", "a", "caa", "ca", "ccaa", "c", "cca", "cc")}
    total_size = 0
    for bra_chg in z_lists:
        for ket_chg in z_lists:
            bb, kb = blocks[bra_chg], blocks[ket_chg]
            nI, nJ = bb.n_states, kb.n_states
            for op in OPS_BY_CHG_DIFF.get(bra_chg - ket_chg, ()):
                if op == "ccaa":
                    E = monomer_2e(bb, kb, V).cpu().numpy()
                    densities[op][(bra_chg, ket_chg)] = [[float(E[i, j]) for j in range(nJ)]
                                                         for i in range(nI)]
                    total_size += nI * nJ
                else:
                    R = density(op, bb, kb).cpu().numpy()
                    densities[op][(bra_chg, ket_chg)] = [[R[i, j] for j in range(nJ)]
                                                         for i in range(nI)]
                    total_size += R.size
    return densities, total_size
