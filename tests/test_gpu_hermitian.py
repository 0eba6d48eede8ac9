"""Densities in the hermitian-XRCC layout (Be-states/densities.py:80-231):
the conjugate strings derived by transposition equal the strings computed
directly, equal-charge tensors are Hermitian in (i, j), the reference's
positional signature and _det options hold; and get_xr_H
(hermitian-XRCC/get_xr_result.py:45, order 0) equals the H_contractions
path element by element after the product-basis sign map
s(a, b) = (-1)^(N_a N_b): the hermitian diagrams' phase (-1)^(n_j0)
(electrons of fragment 0's ket) against H_contractions' sign2 =
(-1)^(electrons of fragment 1's bra) differ on every one-electron
transfer by (-1)^(N_total - 1), which that per-product-state sign
absorbs (N_a (N_a + N_b - 1) = N_a N_b mod 2)."""
import numpy as np
import pytest
import torch

from qodeapplications_amd import density as D
from qodeapplications_amd import synthetic as syn
from qodeapplications_amd.hermitian import build_tensors, conjugate_string

pytestmark = pytest.mark.gpu


def test_hermitian_layout_matches_direct_strings():
    rng = np.random.default_rng(21)
    frag = syn.make_fragment(5, [2, 3, 4], 4, 3, rng)
    # the reference's positional order: thresh, options, xr_order
    dens = build_tensors(frag.states, 5, 3, 1e-10, None, 0)
    assert set(dens) >= {"aa", "a", "caa", "ca", "ccaa", "cc", "c", "cca"}
    blocks = D.charge_blocks(frag.states, 5, 0)
    for op in ("cc", "c", "cca"):
        for (bc, kc), R in dens[op].items():
            direct = D.density(op, blocks[bc], blocks[kc])
            assert float((R - direct).abs().max()) < 1e-12, (op, bc, kc)
    for op in ("ca", "ccaa"):
        for (bc, kc), R in dens[op].items():
            k = R.dim() - 2
            Rt = R.permute(1, 0, *range(R.dim() - 1, 1, -1))
            assert float((R - Rt).abs().max()) < 1e-12
    assert dens["n_elec"] == {chg: 3 - chg for chg in frag.states}
    for chg in frag.states:
        assert torch.equal(dens["KetCoeffs"][(chg, chg)].cpu(), torch.as_tensor(frag.states[chg].coeffs))


def test_det_options_and_signature():
    rng = np.random.default_rng(22)
    frag = syn.make_fragment(3, [1, 2, 3], 2, 2, rng)
    blocks = D.charge_blocks(frag.states, 3, 0)
    dens = build_tensors(frag.states, 3, 2, 1e-10, ["bra_det"], 0)
    assert set(dens) >= {"aa", "a", "caa", "ca", "ccaa", "c", "cca", "cc"}
    for (bc, kc), R in dens["cca"].items():
        eye = blocks[bc].with_coeffs(torch.eye(blocks[bc].n_configs, dtype=torch.float64, device="cuda"))
        assert torch.equal(R, D.density("cca", eye, blocks[kc]))
    assert dens["n_states_bra"] == {c: frag.states[c].configs.shape[0] for c in frag.states}
    assert dens["n_states"] == {c: frag.states[c].coeffs.shape[0] for c in frag.states}
    dk = build_tensors(frag.states, 3, 2, options=["ket_det"])
    assert dk["n_states"] == {c: frag.states[c].configs.shape[0] for c in frag.states}
    dets = {c: np.eye(frag.states[c].configs.shape[0])[:2] for c in frag.states}
    dd = build_tensors(frag.states, 3, 2, options=["bra_det"], dets=dets)
    assert dd["n_states_bra"] == {c: 2 for c in frag.states}
    for c in frag.states:
        ref = frag.states[c].coeffs @ dets[c].T
        assert np.allclose(dd["KetCoeffs"][(c, c)].cpu().numpy(), ref, atol=1e-15)
    with pytest.raises(NotImplementedError):
        build_tensors(frag.states, 3, 2, options=["compress=SVD,1e-4"])
    with pytest.raises(NotImplementedError):
        build_tensors(frag.states, 3, 2, xr_order=2)


def _dimer_system(seed):
    from qodeapplications_amd.xr import XRSystem
    rng = np.random.default_rng(seed)
    orbs = [3, 3]
    frags = [syn.make_fragment(n, [1, 2, 3], 3, 2, rng) for n in orbs]
    h = syn.spatial_h(6, rng)
    V = syn.spatial_V(6, rng)
    U = np.stack([0.2 * syn.spatial_h(6, rng) for _ in orbs])
    ints = syn.FragmentBlocked(orbs, h, V, U)
    return frags, ints, XRSystem(frags, ints, np.zeros((2, 2)))


@pytest.mark.parametrize("which", ["c0", "with_U"])
def test_get_xr_H_equals_general_path_elementwise(which):
    """get_xr_H on build_tensors densities against the H_contractions path
    (checked against FCI in test_oracle / test_gpu_hamiltonian): monomer
    matrices element by element, the pair matrix element by element after
    the product-basis sign map, and the same spectrum."""
    from qodeapplications_amd import hermitian as Hm
    from qodeapplications_amd.xr import XRSystem
    if which == "c0":
        w = syn.config0(seed=0)
        frags, ints = w.fragments, w.integrals
        X = XRSystem(frags, ints, np.zeros((2, 2)))
    else:
        frags, ints, X = _dimer_system(7)
    charges = [list(f.states) for f in frags]
    dens = [Hm.build_tensors(f.states, f.n_orbs, f.n_elec_ref) for f in frags]
    H1h, H2h = Hm.get_xr_H((None, ints, 0.0), dens, 0, charges)
    H1, H2, _ = X.blocks()
    order = []
    for m, F in enumerate(X.frags):
        o = torch.cat([F.pos[c] for c in charges[m]])
        order.append(o)
        assert float((H1h[m] - H1[m][o][:, o]).abs().max()) < 1e-12
    d1 = X.frags[1].d
    idx = (order[0][:, None] * d1 + order[1][None, :]).reshape(-1)
    G = H2[0][1][idx][:, idx]
    ne = [torch.tensor([f.n_elec_ref - c for c in charges[m] for _ in range(f.states[c].coeffs.shape[0])],
                       device="cuda") for m, f in enumerate(frags)]
    par = (ne[0][:, None] * ne[1][None, :]).reshape(-1) % 2
    sgn = 1.0 - 2.0 * par.to(torch.float64)
    mapped = sgn[:, None] * G * sgn[None, :]
    assert float((H2h - mapped).abs().max()) < 1e-12
    # the spectra of the two assembled XR matrices agree
    d0 = X.frags[0].d
    I0 = torch.eye(d0, dtype=torch.float64, device="cuda")
    I1 = torch.eye(d1, dtype=torch.float64, device="cuda")
    M_h = torch.kron(H1h[0], I1) + torch.kron(I0, H1h[1]) + H2h
    M_gen = X.xr_matrix(H=(H1, H2, None))
    ev_h = torch.linalg.eigvalsh(0.5 * (M_h + M_h.T))
    ev_g = torch.linalg.eigvalsh(0.5 * (M_gen + M_gen.T))
    assert float((ev_h - ev_g).abs().max()) < 1e-10


def test_order0_system_diagrams_reproduce_xr_dimer():
    """The XRSystem-based order-0 diagrams (hermitian.dimer_matrix) are the
    same pair matrix as get_xr_H's before the reorder."""
    from qodeapplications_amd import hermitian as Hm
    frags, ints, X = _dimer_system(9)
    S2H2 = Hm.dimer_matrix(X, 0, 1)
    charges = [list(f.states) for f in frags]
    dens = [Hm.build_tensors(f.states, f.n_orbs, f.n_elec_ref) for f in frags]
    _, H2h = Hm.get_xr_H(ints, dens, 0, charges)
    # both in the hermitian sign convention; H2h in product order of the
    # charge-block numbering, S2H2 in charge-block order
    blocks = [(a, b) for a in charges[0] for b in charges[1]]
    n = [f.states for f in frags]
    pos = []
    off0 = {c: sum(n[0][x].coeffs.shape[0] for x in charges[0][:charges[0].index(c)]) for c in charges[0]}
    off1 = {c: sum(n[1][x].coeffs.shape[0] for x in charges[1][:charges[1].index(c)]) for c in charges[1]}
    d1 = sum(n[1][c].coeffs.shape[0] for c in charges[1])
    for a, b in blocks:
        for i in range(n[0][a].coeffs.shape[0]):
            for j in range(n[1][b].coeffs.shape[0]):
                pos.append((off0[a] + i) * d1 + off1[b] + j)
    pos = torch.tensor(pos, device="cuda")
    assert float((S2H2 - H2h[pos][:, pos]).abs().max()) < 1e-12
