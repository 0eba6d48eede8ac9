"""Drop-in for the XR-FCI part of general-XRCC/excitonic.py (fci, :56-82)."""
import torch

from .eigen import lowest_eigenpairs
from .hamiltonian import braket_loops


def fci(H, out=None, device=0):
    """Lowest eigenvalue of the assembled XR matrix.  H = (H1, H2, H3); all
    fragments need the same number of states (the reference raises
    AssertionError otherwise).  out: optional object with .log(...)."""
    monomer_Hamiltonians = H[0]
    N_frag = len(monomer_Hamiltonians)
    states_per_frag = monomer_Hamiltonians[0].shape[0]
    for h in monomer_Hamiltonians:
        if h.shape[0] != states_per_frag:
            raise AssertionError
    if out is not None:
        out.log("Building ...")
    n = states_per_frag ** N_frag
    Hmat = torch.zeros(n, n, dtype=torch.float64, device=torch.device("cuda", device))
    braket_loops(Hmat, N_frag, states_per_frag, H, device=device)
    if out is not None:
        out.log("Diagonalizing ...")
    H1d = [torch.as_tensor(h, dtype=torch.float64, device=Hmat.device) for h in monomer_Hamiltonians]
    w, _ = lowest_eigenpairs(0.5 * (Hmat + Hmat.T), k=1, H1=H1d)
    E = float(w[0])
    if out is not None:
        out.log("FCI Energy  =", E)
    return E
