"""Writes be631g_states.npz (run on the CPU, in a checkout that has the
reference tree): the Be 6-31G fragment states the reference keeps as test
data (tests/ref_data/check_{configs,states}_{m}_{chg}.pkl, 9 spatial
orbitals, core orbital 0 of each spin frozen, charges +1 / 0 / -1 with
2 / 10 / 10 states), read with the non-executing decoder
qodeapplications_amd.formats.read_pickled_data, both as stored (bit-packed
configs in ascending order) and converted to the lexical layout; plus the
oracle's XR dimer of two such fragments with seeded synthetic integrals
(there are no Be integrals without psi4; regenerated from the same seed by
the tests): both monomer blocks, 4000 sampled pair-matrix elements, and the
lowest 5 eigenvalues of the assembled matrix.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from qodeapplications_amd import formats as F  # noqa: E402
from qodeapplications_amd import synthetic as syn  # noqa: E402

REF = "/root/reference/tests/ref_data"
N_ORBS, N_CORE, N_ELEC_REF = 9, 1, 4
CHARGES = (0, 1, -1)


def fragments(out):
    frags = []
    for m in (0, 1):
        states = {}
        for chg in CHARGES:
            cfg = F.read_pickled_data(os.path.join(REF, f"check_configs_{m}_{chg}.pkl"))
            vec = F.read_pickled_data(os.path.join(REF, f"check_states_{m}_{chg}.pkl"))
            z = np.stack([np.asarray(v, dtype=np.float64) for v in vec])
            lex, zl = F.colex_to_lex(cfg, z, N_ORBS, N_CORE)
            key = f"{m}_{'m' if chg < 0 else 'p'}{abs(chg)}"
            out["packed_" + key] = np.asarray(cfg, dtype=np.int64)
            out["coeffs_colex_" + key] = z
            out["configs_" + key] = lex
            out["coeffs_" + key] = zl
            states[chg] = syn.ChargeStates(configs=lex, coeffs=zl)
        frags.append(syn.Fragment(n_orbs=N_ORBS, n_core=N_CORE, n_elec_ref=N_ELEC_REF, states=states))
    return frags


def main():
    t0 = time.time()
    out = {}
    frags = fragments(out)
    rng = np.random.default_rng(631)
    n_sp = 2 * N_ORBS
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng, np.concatenate([np.arange(N_ORBS)] * 2))
    ints = syn.FragmentBlocked([N_ORBS, N_ORBS], h, V)
    for m, f in enumerate(frags):
        f.rho, _ = orc.build_density_tensors(f.states, f.n_orbs, ints.V[m, m, m, m], f.n_core)
    H1, H2, _ = orc.compute_H_blocks(frags, ints, np.zeros((2, 2)))
    E0, Hmat = orc.fci((H1, H2, None))
    out["H1_0"], out["H1_1"] = H1[0], H1[1]
    D = H2[0][1].shape[0]
    idx = rng.integers(0, D, size=(4000, 2))
    out["H2_idx"] = idx
    out["H2_val"] = H2[0][1][idx[:, 0], idx[:, 1]]
    out["evals"] = np.sort(np.linalg.eigvalsh(0.5 * (Hmat + Hmat.T)))[:5]
    np.savez_compressed(os.path.join(HERE, "be631g_states.npz"), **out)
    print(f"be631g: E0 = {E0:.12f}, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
