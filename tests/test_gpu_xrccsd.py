"""XR-CCSD on the GPU (xrccsd.ccsd, drop-in for general-XRCC/excitonic.py:
ccsd) against the oracle's brute-force restatement (oracle.xr_ccsd, pinned
in tests/test_oracle_ccsd.py: exact for dimers) on weakly coupled dimers,
trimers and 4-fragment chains whose fragment states are rotated into their
monomer eigenbases."""
import numpy as np
import pytest

import oracle.oracle as orc
from qodeapplications_amd.xrccsd import ccsd
from test_oracle_ccsd import xr_blocks

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("orbs,seed,ns", [((2, 2), 3, 3), ((2, 2), 4, 3), ((2, 2, 2), 5, 2), ((2, 2, 2, 2), 6, 2)])
def test_xrccsd_matches_oracle(orbs, seed, ns):
    H1, H2, ch = xr_blocks(orbs, seed, n_states=ns)
    E_ref, t1, t2 = orc.xr_ccsd(H1, H2)

    class Out:
        lines = []

        def log(self, *a):
            self.lines.append(" ".join(str(x) for x in a))

    out = Out()
    E, amp = ccsd((H1, H2), out, charges=ch)
    assert abs(E - E_ref) < 1e-9, (E, E_ref)
    for m, t in enumerate(t1):
        assert float(np.abs(amp.singles[m].cpu().numpy() - t).max()) < 1e-7
    for (m, k), t in t2.items():
        assert float(np.abs(amp.doubles[(m, k)].cpu().numpy() - t).max()) < 1e-7
    assert any("CCSD Energy" in line for line in out.lines)
