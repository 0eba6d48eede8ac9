"""Drop-in for general-XRCC/hamiltonian.py (braket_loops), assembled on the GPU."""
import numpy as np
import torch

from .xr import assemble


def braket_loops(Hmat, N_frag, states_per_frag, H, parities=None, device=0):
    """Adds the tensor-product XR matrix built from H = (H1, H2, H3) into
    Hmat (general-XRCC/hamiltonian.py:82-83).  H blocks may be numpy arrays
    or device tensors; H3 may be None or hold None entries.  parities
    (optional, per fragment the electron-count parity of each state) adds
    the spectator sign the reference omits."""
    dev = torch.device("cuda", device)
    t = lambda x: None if x is None else torch.as_tensor(np.asarray(x, dtype=np.float64) if not isinstance(x, torch.Tensor) else x, dtype=torch.float64, device=dev).contiguous()
    H1, H2, H3 = H
    H1d = [t(h) for h in H1]
    H2d = [[t(H2[m][n]) if (H2[m] is not None and n > m) else None for n in range(N_frag)] for m in range(N_frag)]
    H3d = None
    if H3 is not None:
        H3d = [[[t(H3[a][b][c]) if (H3[a] is not None and H3[a][b] is not None and a < b < c) else None
                 for c in range(N_frag)] for b in range(N_frag)] for a in range(N_frag)]
    if any(h.shape[0] != states_per_frag for h in H1d):
        raise AssertionError("states_per_frag does not match H1")
    par = None if parities is None else [torch.as_tensor(np.asarray(p), dtype=torch.int8, device=dev) for p in parities]
    M = assemble(H1d, H2d, H3d, par)
    if isinstance(Hmat, torch.Tensor):
        Hmat += M.to(Hmat.device)
    else:
        Hmat += M.cpu().numpy()
    return Hmat
