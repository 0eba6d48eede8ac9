"""GPU checks of the hand-written adjoint.

* density / sigma adjoints: the densities are bilinear in (zbra, zket) and
  the sigma term linear in V, so f(x + d) - f(x) = grad . d holds exactly
  for f = <Rbar, rho>; a wrong index, sign or missing term breaks it;
* dE/dz and dE/dkappa of the lowest XR eigenvalue against central finite
  differences (step 1e-5), for the plain eigenvalue and for the
  generalised one with the tensor-product overlap.
"""
import numpy as np
import pytest
import torch

from qodeapplications_amd import density as D
from qodeapplications_amd import synthetic as syn
from qodeapplications_amd._lib import context, ptr
from qodeapplications_amd.grad import DimerAdjoint, XRAdjoint, energy_and_gradient
from qodeapplications_amd.orbitals import OrbitalRotation, dimer_integral_keys, energy_and_gradients
from qodeapplications_amd.xr import XRSystem

pytestmark = pytest.mark.gpu


def _blk(rng, n_orbs, n_elec, ns, n_core=0):
    return D.ChargeBlock.from_states(syn.random_states(n_orbs, n_elec, ns, rng, n_core), n_orbs, n_core)


def _dens_adjoint(op, bb, kb, Rbar):
    zb = torch.zeros_like(bb.coeffs)
    zk = torch.zeros_like(kb.coeffs)
    context(0).call("xr_density_adjoint", op.encode(), kb.n_orbs, kb.n_core, kb.n_elec, bb.n_states,
                    bb.n_configs, ptr(bb.coeffs), kb.n_states, kb.n_configs, ptr(kb.coeffs),
                    ptr(Rbar), ptr(zb), ptr(zk))
    return zb, zk


@pytest.mark.parametrize("op", D.SUPPORTED_OPS)
@pytest.mark.parametrize("n_orbs,n_core,n_ket", [(4, 0, 3), (4, 1, 4), (5, 0, 4)])
def test_density_adjoint_is_exact(op, n_orbs, n_core, n_ket):
    n_bra = n_ket + D.n_elec_change(op)
    if n_bra < 2 * n_core or n_bra > 2 * n_orbs:
        pytest.skip("empty bra space")
    rng = np.random.default_rng(31)
    kb = _blk(rng, n_orbs, n_ket, 3, n_core)
    bb = _blk(rng, n_orbs, n_bra, 4, n_core)
    R = D.density(op, bb, kb)
    Rbar = torch.randn(R.shape, dtype=torch.float64, device=R.device)
    f = lambda b, k: float((Rbar * D.density(op, b, k)).sum())
    zb, zk = _dens_adjoint(op, bb, kb, Rbar)
    db = torch.randn_like(bb.coeffs)
    dk = torch.randn_like(kb.coeffs)
    f0 = f(bb, kb)
    fb = f(bb.with_coeffs(bb.coeffs + db), kb)
    fk = f(bb, kb.with_coeffs(kb.coeffs + dk))
    scale = 1 + abs(f0)
    assert abs((fb - f0) - float((zb * db).sum())) < 1e-10 * scale
    assert abs((fk - f0) - float((zk * dk).sum())) < 1e-10 * scale


@pytest.mark.parametrize("n_orbs,n_core,n_elec", [(4, 0, 3), (4, 1, 4), (5, 0, 4)])
def test_monomer_2e_adjoint_is_exact(n_orbs, n_core, n_elec):
    rng = np.random.default_rng(32)
    kb = _blk(rng, n_orbs, n_elec, 3, n_core)
    bb = _blk(rng, n_orbs, n_elec, 4, n_core)
    ints = syn.FragmentBlocked([n_orbs], syn.spatial_h(n_orbs, rng), syn.spatial_V(n_orbs, rng))
    V = torch.as_tensor(ints.V[0, 0, 0, 0], device="cuda")
    Eb = torch.randn(4, 3, dtype=torch.float64, device="cuda")
    f = lambda b, k, VV: float((Eb * D.monomer_2e(b, k, VV)).sum())
    zb, zk, Vb = torch.zeros_like(bb.coeffs), torch.zeros_like(kb.coeffs), torch.zeros_like(V)
    context(0).call("xr_monomer_2e_adjoint", n_orbs, n_core, n_elec, 4, bb.n_configs, ptr(bb.coeffs),
                    3, kb.n_configs, ptr(kb.coeffs), ptr(V), ptr(Eb), ptr(zb), ptr(zk), ptr(Vb))
    f0 = f(bb, kb, V)
    db, dk, dV = torch.randn_like(bb.coeffs), torch.randn_like(kb.coeffs), torch.randn_like(V)
    scale = 1 + abs(f0)
    assert abs(f(bb.with_coeffs(bb.coeffs + db), kb, V) - f0 - float((zb * db).sum())) < 1e-10 * scale
    assert abs(f(bb, kb.with_coeffs(kb.coeffs + dk), V) - f0 - float((zk * dk).sum())) < 1e-10 * scale
    assert abs(f(bb, kb, V + dV) - f0 - float((Vb * dV).sum())) < 1e-10 * scale


@pytest.mark.parametrize("same", [True, False])
@pytest.mark.parametrize("n_orbs,n_core,n_elec", [(4, 0, 3), (5, 1, 5)])
def test_monomer_2e_keep_matches_recompute(n_orbs, n_core, n_elec, same):
    """The sigma intermediate kept by xr_monomer_2e_keep and handed to
    xr_monomer_2e_adjoint_keep gives the same energy matrix and the same
    adjoints, bit for bit, as recomputing it (a block with itself takes the
    one-gather path, two blocks the general one)."""
    rng = np.random.default_rng(33)
    kb = _blk(rng, n_orbs, n_elec, 3, n_core)
    bb = kb if same else _blk(rng, n_orbs, n_elec, 4, n_core)
    ints = syn.FragmentBlocked([n_orbs], syn.spatial_h(n_orbs, rng), syn.spatial_V(n_orbs, rng))
    V = torch.as_tensor(ints.V[0, 0, 0, 0], device="cuda")
    E0 = D.monomer_2e(bb, kb, V)
    E1, W = D.monomer_2e(bb, kb, V, keep=True)
    assert torch.equal(E0, E1)
    nI = bb.coeffs.shape[0]
    Eb = torch.randn(nI, 3, dtype=torch.float64, device="cuda")
    outs = []
    for keep in (False, True):
        zb, zk, Vb = torch.zeros_like(bb.coeffs), torch.zeros_like(kb.coeffs), torch.zeros_like(V)
        if same:
            zk = zb
        args = (n_orbs, n_core, n_elec, nI, bb.n_configs, ptr(bb.coeffs), 3, kb.n_configs, ptr(kb.coeffs),
                ptr(V), ptr(Eb), ptr(zb), ptr(zk), ptr(Vb))
        if keep:
            context(0).call("xr_monomer_2e_adjoint_keep", *args, ptr(W), W.numel())
        else:
            context(0).call("xr_monomer_2e_adjoint", *args)
        outs.append((zb, zk, Vb))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("hermitian", [False, True])
def test_monomer_2e_adjoint_shared_buffer(hermitian):
    """A block with itself into one shared coefficient adjoint equals the sum
    of the two separate adjoints: with a V lacking the symmetry
    V[pqrs] = V[rspq] the general path runs (no promise made); with a
    Hermitian V and xr_assume_hermitian_v the one-gather path does."""
    rng = np.random.default_rng(34)
    n_orbs, n_core, n_elec = 4, 0, 3
    kb = _blk(rng, n_orbs, n_elec, 3, n_core)
    dim = 2 * n_orbs
    if hermitian:
        ints = syn.FragmentBlocked([n_orbs], syn.spatial_h(n_orbs, rng), syn.spatial_V(n_orbs, rng))
        V = torch.as_tensor(ints.V[0, 0, 0, 0], device="cuda")
    else:
        V = torch.as_tensor(rng.standard_normal((dim,) * 4), device="cuda")
        assert float((V - V.permute(2, 3, 0, 1)).abs().max()) > 0.1
    Eb = torch.randn(3, 3, dtype=torch.float64, device="cuda")
    ctx = context(0)
    args = lambda zb, zk, Vb: (n_orbs, n_core, n_elec, 3, kb.n_configs, ptr(kb.coeffs), 3, kb.n_configs,
                               ptr(kb.coeffs), ptr(V), ptr(Eb), ptr(zb), ptr(zk), ptr(Vb))
    zb, zk, Vb2 = torch.zeros_like(kb.coeffs), torch.zeros_like(kb.coeffs), torch.zeros_like(V)
    ctx.call("xr_monomer_2e_adjoint", *args(zb, zk, Vb2))
    zs, Vb1 = torch.zeros_like(kb.coeffs), torch.zeros_like(V)
    ctx.call("xr_assume_hermitian_v", 1 if hermitian else 0)
    try:
        ctx.call("xr_monomer_2e_adjoint", *args(zs, zs, Vb1))
    finally:
        ctx.call("xr_assume_hermitian_v", 0)
    scale = float((zb + zk).abs().max())
    assert float((zs - (zb + zk)).abs().max()) <= 1e-12 * max(1.0, scale)
    assert float((Vb1 - Vb2).abs().max()) <= 1e-12 * max(1.0, float(Vb2.abs().max()))
    # and it is the derivative of sum(Eb * E2) with bra and ket moving together
    f = lambda z: float((Eb * D.monomer_2e(kb.with_coeffs(z), kb.with_coeffs(z), V)).sum())
    # f is quadratic in z: the central difference is exact up to rounding
    dz = torch.randn_like(kb.coeffs) * 1e-3
    lin = float((zs * dz).sum())
    cd = 0.5 * (f(kb.coeffs + dz) - f(kb.coeffs - dz))
    assert abs(cd - lin) <= 1e-10 * max(abs(lin), 1.0)


def _dimer(seed, orbs=(3, 3), elec=(1, 2, 3), ns=3):
    rng = np.random.default_rng(seed)
    frags = [syn.make_fragment(n, list(elec), ns, 2, rng) for n in orbs]
    n_sp = sum(orbs)
    h = syn.spatial_h(n_sp, rng)
    V = syn.spatial_V(n_sp, rng)
    U = np.stack([0.2 * syn.spatial_h(n_sp, rng) for _ in orbs])
    nuc = np.array([[0.0, 0.3], [0.3, 0.0]])
    return frags, (list(orbs), h, V, U), nuc


def _energy(X, coeffs, generalised):
    for (m, c), z in coeffs.items():
        X.frags[m].blocks[c] = X.frags[m].blocks[c].with_coeffs(z)
    adj = DimerAdjoint(X)
    E = adj.forward()
    if not generalised:
        return E
    # lowest eigenvalue of (H, S) with S the overlap of the product states
    H1, H2, _ = X.blocks()
    M = X.xr_matrix(H=(H1, H2, None))
    s = []
    for F in X.frags:
        Sm = torch.zeros(F.d, F.d, dtype=torch.float64, device=X.device)
        for c in F.charges:
            z = F.blocks[c].coeffs
            p = F.pos[c]
            Sm[p[:, None], p[None, :]] = z @ z.T
        s.append(Sm)
    S = torch.kron(s[0], s[1])
    L = torch.linalg.cholesky(S)
    Li = torch.linalg.inv(L)
    Ht = Li @ (0.5 * (M + M.T)) @ Li.T
    return float(torch.linalg.eigvalsh(Ht)[0])


@pytest.mark.parametrize("orbs", [(3, 3), (2, 4)])
def test_grouped_pair_adjoint_matches_per_combination(orbs):
    """The charge-conserving and one-electron-transfer combinations of a
    pair, taken together (partner sums first, each output in its cheaper
    association; uneven fragments pick different associations per term),
    give the adjoint of taking every combination on its own."""
    frags, (orbs_, h, V, U), nuc = _dimer(41, orbs=orbs)
    X = XRSystem(frags, syn.FragmentBlocked(orbs_, h, V, U), nuc)
    got = []
    for group in (True, False):
        adj = XRAdjoint(X, group_combos=group)
        adj.forward()
        got.append(adj.backward())
    (z1, i1), (z2, i2) = got
    for m in range(2):
        for c in z1[m]:
            ref = z2[m][c]
            assert float((z1[m][c] - ref).abs().max()) <= 1e-11 * max(1.0, float(ref.abs().max())), (m, c)
    assert set(i1) == set(i2)
    for k, ref in i2.items():
        assert float((i1[k] - ref).abs().max()) <= 1e-11 * max(1.0, float(ref.abs().max())), k


@pytest.mark.parametrize("lagrangian", [False, True])
def test_state_gradient_matches_finite_differences(lagrangian):
    frags, (orbs, h, V, U), nuc = _dimer(33)
    ints = syn.FragmentBlocked(orbs, h, V, U)
    X = XRSystem(frags, ints, nuc)
    E, v, zbar, _ = energy_and_gradient(X, lagrangian=lagrangian)
    base = {(m, c): X.frags[m].blocks[c].coeffs.clone() for m in range(2) for c in X.frags[m].charges}
    rng = torch.Generator(device="cuda").manual_seed(5)
    dirs = {k: torch.randn(t.shape, dtype=torch.float64, device="cuda", generator=rng) for k, t in base.items()}
    eps = 1e-5
    plus = {k: base[k] + eps * dirs[k] for k in base}
    minus = {k: base[k] - eps * dirs[k] for k in base}
    fd = (_energy(X, plus, lagrangian) - _energy(X, minus, lagrangian)) / (2 * eps)
    an = sum(float((zbar[m][c] * dirs[(m, c)]).sum()) for (m, c) in base)
    _energy(X, base, False)
    assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)


def test_orbital_gradient_matches_finite_differences():
    frags, (orbs, h, V, U), nuc = _dimer(34)
    ints = syn.FragmentBlocked(orbs, h, V, U)
    X = XRSystem(frags, ints, nuc)
    rot = OrbitalRotation(orbs, h, V, U)
    E, gz, gk = energy_and_gradients(X, rot, lagrangian=True)
    g = torch.Generator(device="cuda").manual_seed(7)
    dirs = [torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g) for n in orbs]
    keys = dimer_integral_keys()

    def e_at(t):
        ks = [(t * d) for d in dirs]
        blocks = rot.blocks(ks, keys)
        for k in keys:
            X.ints.set_block(k, blocks[k])
        return DimerAdjoint(X).forward()

    eps = 1e-5
    fd = (e_at(eps) - e_at(-eps)) / (2 * eps)
    an = sum(float((a * d).sum()) for a, d in zip(gk, dirs))
    assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)
    # the generator gradient is antisymmetric
    for a in gk:
        assert float((a + a.T).abs().max()) < 1e-12
    assert abs(e_at(0.0) - E) < 1e-12
