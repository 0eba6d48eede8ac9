"""Chains of more than two fragments: the matrix-free XR operator and the
N-fragment adjoint.

* the operator applied to the identity equals the matrix the reference
  loops assemble (general-XRCC/hamiltonian.py:21-83, here the braket kernel
  checked against the oracle's loops in test_gpu_hamiltonian.py);
* its Davidson root equals the dense eigensolver's lowest eigenvalue;
* dE/dz (plain and generalised eigenvalue) and dE/dkappa against central
  finite differences (step 1e-5), on the dense and the Davidson path.
"""
import numpy as np
import pytest
import torch

from qodeapplications_amd import synthetic as syn
from qodeapplications_amd.eigen import DENSE_MAX
from qodeapplications_amd.grad import XRAdjoint, energy_and_gradient
from qodeapplications_amd.orbitals import OrbitalRotation, energy_and_gradients, integral_keys
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu


def _chain(seed, orbs, elec=(1, 2, 3), ns=2):
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, list(elec), ns, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    U = np.stack([0.2 * syn.spatial_h(n_sp, rng) for _ in orbs])
    nf = len(orbs)
    nuc = np.array([[0.0 if m == n else 0.3 / abs(m - n) for n in range(nf)] for m in range(nf)])
    return frags, (list(orbs), h, V, U), nuc


def _system(seed, orbs, elec=(1, 2, 3), ns=2, precision="f64"):
    frags, (o, h, V, U), nuc = _chain(seed, orbs, elec, ns)
    return XRSystem(frags, syn.FragmentBlocked(o, h, V, U), nuc, precision=precision), (o, h, V, U)


@pytest.mark.parametrize("orbs", [(2, 3), (2, 2, 2), (2, 2, 2, 2), (3, 2, 3)])
def test_operator_equals_assembled_matrix(orbs):
    X, _ = _system(40, orbs)
    H = X.blocks()
    M = X.xr_matrix(H=H)
    A = X.operator(H=H).dense()
    assert float((A - M).abs().max()) < 1e-13


@pytest.mark.parametrize("orbs,precision", [((2, 3), "f64"), ((2, 2, 2), "f64"), ((3, 2, 3), "split_f32")])
def test_sector_blocks_match_dense_pair_matrices(orbs, precision):
    """Pair terms built straight into the operator's fixed-charge blocks equal
    the same blocks cut out of the dense pair matrices (same sums, same
    order), and the operator built that way equals the assembled matrix."""
    X, _ = _system(42, orbs, ns=3, precision=precision)
    H = X.blocks()
    ref = X.operator(H=H)
    op = X.operator()
    assert set(op.pairs) == set(ref.pairs)
    for key, P in op.pairs.items():
        R = ref.pairs[key]
        assert P["segs"] == R["segs"] and P["segq"] == R["segq"]
        for a, b in zip(P["blocks"], R["blocks"]):
            assert torch.equal(a, b.to(a.dtype))
            assert torch.equal(a, a.T)
    if precision == "f64":
        assert float((op.dense() - X.xr_matrix(H=H)).abs().max()) < 1e-13


def test_operator_equals_assembled_matrix_c0():
    w = syn.config0(seed=0)
    X = XRSystem(w.fragments, w.integrals, w.nuc)
    H = X.blocks()
    assert float((X.operator(H=H).dense() - X.xr_matrix(H=H)).abs().max()) < 1e-13


def test_operator_block_columns():
    """A block of columns through the operator equals the same columns of
    the assembled matrix (the column index travels with the packed axes)."""
    X, _ = _system(41, (2, 3, 2), ns=3)
    H = X.blocks()
    M = X.xr_matrix(H=H)
    op = X.operator(H=H)
    g = torch.Generator(device="cuda").manual_seed(3)
    V = torch.randn(op.n, 5, dtype=torch.float64, device="cuda", generator=g)
    Y = torch.empty_like(V)
    op.matvec(V, Y)
    assert float((Y - M @ V).abs().max()) < 1e-12


@pytest.mark.parametrize("orbs,b", [((2, 3, 2, 2), 1), ((2, 3, 2, 2), 3), ((3, 2, 3), 2)])
def test_sector_matvec_matches_assembled(orbs, b):
    """The compact per-charge operator (monomer axis kernel, multi-item
    packs incl. the LDS-tiled transposes, grouped pair GEMM, unpacks) against
    the assembled matrix, every total charge, b columns."""
    X, _ = _system(46, orbs, ns=3)
    H = X.blocks()
    M = X.xr_matrix(H=H)
    op = X.operator(H=H)
    g = torch.Generator(device="cuda").manual_seed(b)
    for Q in op.total_charges():
        sp = op.sector_space(Q)
        V = torch.randn(sp.size, b, dtype=torch.float64, device="cuda", generator=g)
        Y = torch.empty_like(V)
        sp.matvec(V, Y)
        Vf = torch.zeros(op.n, b, dtype=torch.float64, device="cuda")
        Yf = torch.zeros(op.n, b, dtype=torch.float64, device="cuda")
        for j in range(b):
            col = torch.zeros(op.n, dtype=torch.float64, device="cuda")
            sp.to_product(V[:, j].contiguous(), col)
            Vf[:, j] = col
            col = torch.zeros(op.n, dtype=torch.float64, device="cuda")
            sp.to_product(Y[:, j].contiguous(), col)
            Yf[:, j] = col
        ref = M @ Vf
        assert float((Yf - ref).abs().max()) < 1e-12 * max(1.0, float(ref.abs().max())), Q


@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_axis_apply_multi_matches_einsum(beta):
    """xr_axis_apply_multi (the monomer terms and eigenbasis changes of the
    sector operator) against an FP64 einsum, over 20 items (two launches) of
    every column layout: post 1 (contiguous tile), post < 64 (several p per
    tile), post >= 64 with a ragged last tile; n from 1 to 64; U and U^T."""
    import ctypes
    from qodeapplications_amd._lib import context
    g = torch.Generator(device="cuda").manual_seed(int(beta) + 7)
    shapes = [(1, 32, 70000), (1024, 32, 32), (32768, 32, 1), (3, 64, 100), (5, 1, 7), (7, 3, 1),
              (2, 17, 65), (9, 33, 20), (1, 64, 1), (4, 5, 129), (100, 8, 3), (1, 2, 64),
              (13, 31, 1), (2, 64, 64), (6, 16, 63), (1, 4, 1000), (8, 40, 2), (1, 1, 1),
              (3, 48, 5), (50, 12, 1)]
    items = []
    for q, (pre, n, post) in enumerate(shapes):
        U = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        src = torch.randn(pre * n * post, dtype=torch.float64, device="cuda", generator=g)
        dst = torch.randn(pre * n * post, dtype=torch.float64, device="cuda", generator=g)
        items.append((U, src, dst, pre, n, post, q % 2))
    ref = []
    for U, src, dst, pre, n, post, tr in items:
        A = U.T if tr else U
        ref.append(torch.einsum("ij,pjq->piq", A, src.view(pre, n, post)).reshape(-1) + beta * dst)
    k = len(items)
    I, L, P = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
    arr = lambda T, vals: (T * k)(*vals)
    context(0).call("xr_axis_apply_multi", k, arr(P, [it[0].data_ptr() for it in items]),
                    arr(I, [it[4] for it in items]), arr(I, [it[6] for it in items]),
                    arr(L, [it[3] for it in items]), arr(L, [it[5] for it in items]),
                    arr(P, [it[1].data_ptr() for it in items]), arr(P, [it[2].data_ptr() for it in items]),
                    float(beta))
    torch.cuda.synchronize()
    for q, (it, r) in enumerate(zip(items, ref)):
        assert float((it[2] - r).abs().max()) < 1e-12 * max(1.0, float(r.abs().max())), shapes[q]


def test_davidson_root_matches_dense():
    X, _ = _system(42, (4, 4, 4), ns=7)
    H = X.blocks()
    op = X.operator(H=H)
    assert op.n > DENSE_MAX
    w, v = op.lowest(k=1)
    ref = torch.linalg.eigvalsh(X.xr_matrix(H=H))[0]
    assert abs(float(w[0]) - float(ref)) < 1e-9
    r = torch.empty_like(v)
    op.matvec(v, r)
    assert float((r - w[0] * v).norm()) < 1e-7


def _set_coeffs(X, coeffs):
    for (m, c), z in coeffs.items():
        X.frags[m].blocks[c] = X.frags[m].blocks[c].with_coeffs(z)


def _generalised_energy(X):
    X.build_densities()
    H = X.blocks()
    M = X.xr_matrix(H=H)
    S = None
    for F in X.frags:
        Sm = torch.zeros(F.d, F.d, dtype=torch.float64, device=X.device)
        for c in F.charges:
            z = F.blocks[c].coeffs
            p = F.pos[c]
            Sm[p[:, None], p[None, :]] = z @ z.T
        S = Sm if S is None else torch.kron(S, Sm)
    Li = torch.linalg.inv(torch.linalg.cholesky(S))
    return float(torch.linalg.eigvalsh(Li @ (0.5 * (M + M.T)) @ Li.T)[0])


@pytest.mark.parametrize("orbs,ns,lagrangian", [((2, 2, 2), 2, False), ((2, 2, 2), 2, True),
                                                ((2, 2, 2, 2), 2, True), ((4, 4, 4), 7, False)])
def test_chain_state_gradient_matches_finite_differences(orbs, ns, lagrangian):
    X, _ = _system(43, orbs, ns=ns)
    E, v, zbar, _ = energy_and_gradient(X, lagrangian=lagrangian)
    base = {(m, c): X.frags[m].blocks[c].coeffs.clone() for m in range(len(orbs)) for c in X.frags[m].charges}
    g = torch.Generator(device="cuda").manual_seed(9)
    dirs = {k: torch.randn(t.shape, dtype=torch.float64, device="cuda", generator=g) for k, t in base.items()}
    # unit direction: keeps the O(eps^2) term of the central difference small
    norm = float(sum(float((d * d).sum()) for d in dirs.values())) ** 0.5
    dirs = {k: d / norm for k, d in dirs.items()}
    eps = 1e-5

    def e_at(t):
        _set_coeffs(X, {k: base[k] + t * dirs[k] for k in base})
        return _generalised_energy(X) if lagrangian else XRAdjoint(X).forward()

    fd = (e_at(eps) - e_at(-eps)) / (2 * eps)
    an = sum(float((zbar[m][c] * dirs[(m, c)]).sum()) for (m, c) in base)
    _set_coeffs(X, base)
    assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)


@pytest.mark.parametrize("orbs", [(2, 2, 2), (2, 3, 2, 2)])
def test_chain_orbital_gradient_matches_finite_differences(orbs):
    X, (o, h, V, U) = _system(44, orbs)
    rot = OrbitalRotation(o, h, V, U)
    E, gz, gk = energy_and_gradients(X, rot, lagrangian=True)
    g = torch.Generator(device="cuda").manual_seed(11)
    dirs = [torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g) for n in o]
    keys = integral_keys(len(o))

    def e_at(t):
        blocks = rot.blocks([t * d for d in dirs], keys)
        for k in keys:
            X.ints.set_block(k, blocks[k])
        return XRAdjoint(X).forward()

    eps = 1e-5
    fd = (e_at(eps) - e_at(-eps)) / (2 * eps)
    an = sum(float((a * d).sum()) for a, d in zip(gk, dirs))
    assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)
    for a in gk:
        assert float((a + a.T).abs().max()) < 1e-12
    assert abs(e_at(0.0) - E) < 1e-12


def test_chain_split_precision_energy_and_gradient():
    """Split-FP32 two-electron contractions on a 3-fragment chain: energy
    within 1e-9 Eh of FP64 and dE/dz within 1e-7 of it (relative to the
    gradient's largest entry)."""
    Xf, _ = _system(45, (3, 3, 3), ns=3)
    Xs, _ = _system(45, (3, 3, 3), ns=3, precision="split_f32")
    Ef, _, zf, _ = energy_and_gradient(Xf, lagrangian=True)
    Es, _, zs, _ = energy_and_gradient(Xs, lagrangian=True)
    assert abs(Ef - Es) < 1e-9
    scale = max(float(t.abs().max()) for per in zf for t in per.values())
    err = max(float((zf[m][c] - zs[m][c]).abs().max()) for m in range(3) for c in zf[m])
    assert err < 1e-7 * scale, (err, scale)
