"""Out-of-memory paths and the store-or-recompute choice (forced from the
host, no GPU allocation actually fails):

  * with xr_set_memory_limit below what a density's scratch needs, the call
    fails with XR_ENOMEM (a clean error, the context stays usable) and runs
    again once the cap is lifted;
  * an XRSystem whose sigma budget is below every monomer sigma
    intermediate recomputes them in the adjoint, with the same energy and
    gradient as the one that stores them."""
import numpy as np
import pytest
import torch

from qodeapplications_amd import density as D
from qodeapplications_amd import synthetic as syn
from qodeapplications_amd._lib import XR_ENOMEM, XRError, context

pytestmark = pytest.mark.gpu


def test_memory_limit_gives_enomem_and_recovers():
    rng = np.random.default_rng(1)
    st = syn.random_states(7, 4, 6, rng)
    sb = syn.random_states(7, 5, 5, rng)
    kb = D.ChargeBlock(7, 0, st.configs, st.coeffs)
    bb = D.ChargeBlock(7, 0, sb.configs, sb.coeffs)
    ctx = context(0)
    ctx.call("xr_trim")
    ctx.call("xr_set_memory_limit", 4096)
    try:
        with pytest.raises(XRError) as e:
            D.density("cca", bb, kb)
        assert e.value.code == XR_ENOMEM
    finally:
        ctx.call("xr_set_memory_limit", 0)
    ref = D.density("cca", bb, kb)
    again = D.density("cca", bb, kb)
    torch.cuda.synchronize()
    assert torch.equal(ref, again)


def test_sigma_recomputed_under_budget_matches_stored():
    from qodeapplications_amd.grad import energy_and_gradient
    from qodeapplications_amd.xr import XRSystem
    rng = np.random.default_rng(3)
    orbs = [3, 4]
    frags = [syn.make_fragment(n, [2, 3, 4], 3, 3, rng) for n in orbs]
    ints = syn.FragmentBlocked(orbs, syn.spatial_h(7, rng), syn.spatial_V(7, rng))
    keep = XRSystem(frags, ints, np.zeros((2, 2)))
    recompute = XRSystem(frags, ints, np.zeros((2, 2)), sigma_budget=0)
    E1, _, z1, i1 = energy_and_gradient(keep)
    assert any(F.E2W for F in keep.frags)
    E2, _, z2, i2 = energy_and_gradient(recompute)
    assert not any(F.E2W for F in recompute.frags)
    assert abs(E1 - E2) < 1e-12
    for m in range(2):
        for c in z1[m]:
            assert float((z1[m][c] - z2[m][c]).abs().max()) < 1e-10
    for k in i1:
        assert float((i1[k] - i2[k]).abs().max()) < 1e-10
