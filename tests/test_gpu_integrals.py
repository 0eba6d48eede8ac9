"""Device integral transforms of the orbital loop
(qodeapplications_amd/integrals.py, hermitian-XRCC/get_ints.py:57-155)
against the oracle's restatement (oracle/integrals_ref.py) on the same
non-orthogonal two-fragment orbitals, to 1e-11."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed=5, n=6):
    from qodeapplications_amd import synthetic as syn
    rng = np.random.default_rng(seed)
    h = syn.spatial_h(n, rng)
    V = syn.spatial_V(n, rng)
    C = np.eye(n) + 0.25 * rng.standard_normal((n, n))
    S = C.T @ C
    return S, C.T @ h @ C, np.einsum("ap,bq,cr,ds,abcd->pqrs", C, C, C, C, V, optimize=True)


def test_transforms_match_oracle():
    from oracle import integrals_ref as ir
    from qodeapplications_amd import integrals as I
    S, h, V = _problem()
    ints = I.SupersystemIntegrals([3, 3], S, h, V)
    bi = ints.biorthogonal()
    hb, Vb = ir.biorthogonal(S, h, V)
    np.testing.assert_allclose(bi.h.cpu().numpy(), hb, atol=1e-11)
    np.testing.assert_allclose(bi.V.cpu().numpy(), Vb, atol=1e-11)
    np.testing.assert_allclose(bi.S.cpu().numpy(), np.eye(6), atol=1e-11)
    cp = ints.core_projected([0, 3])
    L, R = ir.core_projection(S, [0, 3])
    h2, V2 = ir.transform(h, V, L, R)
    np.testing.assert_allclose(cp.h.cpu().numpy(), h2, atol=1e-11)
    np.testing.assert_allclose(cp.V.cpu().numpy(), V2, atol=1e-11)
    np.testing.assert_allclose(cp.S.cpu().numpy(), L @ S @ R, atol=1e-11)
    np.testing.assert_allclose(ints.v_half().cpu().numpy(), ir.v_half(S, V), atol=1e-11)
    np.testing.assert_allclose(ints.v_diff().cpu().numpy(), ir.v_half(S, V) - Vb, atol=1e-11)


def test_linear_orbital_update():
    import torch
    from qodeapplications_amd import integrals as I
    rng = np.random.default_rng(2)
    C = rng.standard_normal((7, 5))
    x = rng.standard_normal((5, 5))
    x = x - x.T
    out = I.linear_orbital_update(torch.as_tensor(C, device="cuda"), torch.as_tensor(x, device="cuda"), 0.1)
    np.testing.assert_allclose(out.cpu().numpy(), C @ (np.eye(5) + 0.1 * x), atol=1e-12)
