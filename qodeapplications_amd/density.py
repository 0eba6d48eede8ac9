"""Device-resident fragment state blocks and transition densities.

Thin host layer over libxrhip: coefficients and determinant masks live in
HBM as torch tensors (the container), every density, excitation list and
monomer energy is computed by the HIP kernels behind the C ABI.
"""

import math

import numpy as np
import torch

from ._lib import context, ptr

# operator strings built for each charge difference bra_chg - ket_chg
# (general-XRCC/build_density_tensors.py:90-157)
OPS_BY_CHG_DIFF = {+2: ("aa",), +1: ("a", "caa"), 0: ("ca", "ccaa"), -1: ("c", "cca"), -2: ("cc",)}
# with the 4-operator strings of Be-states/_attic_/deprecated/density_tensors.c:560-688
OPS_BY_CHG_DIFF_FULL = {+2: ("aa", "caaa"), +1: ("a", "caa"), 0: ("ca", "ccaa"),
                        -1: ("c", "cca"), -2: ("cc", "ccca")}
SUPPORTED_OPS = ("a", "c", "ca", "aa", "cc", "caa", "cca", "ccaa", "caaa", "ccca", "ccaaa", "cccaa")


def n_elec_change(op):
    return sum(+1 if o == "c" else -1 for o in op)


def _device(device):
    return torch.device("cuda", device if isinstance(device, int) else torch.device(device).index or 0)


class ChargeBlock:
    """One charge block of one fragment on the GPU: the states' coefficients
    (n_states x n_configs, float64) and the determinants as 64-bit masks,
    checked to be the complete lexical enumeration."""

    def __init__(self, n_orbs, n_core, configs, coeffs, device=0, check=True):
        self.device = _device(device)
        self.n_orbs = int(n_orbs)
        self.n_core = int(n_core)
        self.dim = 2 * self.n_orbs
        cfg = np.ascontiguousarray(np.asarray(configs, dtype=np.int64))
        if cfg.ndim != 2:
            raise ValueError("configs must be (n_configs x n_elec)")
        self.n_elec = cfg.shape[1]
        self.coeffs = torch.as_tensor(np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64)),
                                      device=self.device)
        if self.coeffs.ndim != 2 or self.coeffs.shape[1] != cfg.shape[0]:
            raise ValueError("coeffs must be (n_states x n_configs) matching configs")
        self.n_states, self.n_configs = self.coeffs.shape
        self.masks = torch.empty(cfg.shape[0], dtype=torch.int64, device=self.device)
        if check:
            ctx = context(self.device.index)
            dcfg = torch.as_tensor(cfg, device=self.device)
            ctx.call("xr_configs_to_masks", ptr(dcfg), cfg.shape[0], self.n_elec, self.n_orbs,
                     self.n_core, ptr(self.masks))

    @classmethod
    def from_states(cls, states, n_orbs, n_core, device=0):
        return cls(n_orbs, n_core, states.configs, states.coeffs, device)

    def with_coeffs(self, coeffs):
        """Same determinants, new coefficients (already on the device)."""
        new = object.__new__(ChargeBlock)
        new.__dict__.update(self.__dict__)
        new.coeffs = coeffs.contiguous()
        new.n_states = coeffs.shape[0]
        return new


def density(op, bra, ket, out=None):
    """<I| op |J> for all state pairs: tensor (nI, nJ, dim, ..., dim) float64."""
    if op not in SUPPORTED_OPS:
        raise ValueError(f"unsupported operator string {op!r}")
    if bra.n_orbs != ket.n_orbs or bra.n_core != ket.n_core:
        raise ValueError("bra and ket blocks belong to different fragments")
    if bra.n_elec != ket.n_elec + n_elec_change(op):
        raise ValueError(f"{op}: bra has {bra.n_elec} electrons, ket {ket.n_elec}")
    dim = ket.dim
    shape = (bra.n_states, ket.n_states) + (dim,) * len(op)
    if out is None:
        out = torch.empty(shape, dtype=torch.float64, device=ket.device)
    elif tuple(out.shape) != shape or not out.is_contiguous() or out.dtype != torch.float64:
        raise ValueError("out has the wrong shape / layout")
    ctx = context(ket.device.index)
    ctx.call("xr_density", op.encode(), ket.n_orbs, ket.n_core, ket.n_elec,
             bra.n_states, bra.n_configs, ptr(bra.coeffs),
             ket.n_states, ket.n_configs, ptr(ket.coeffs), ptr(out))
    return out


def monomer_2e(bra, ket, V, keep=False):
    """E[I,J] = sum V[p,q,r,s] <I|p+ q+ s r|J> (nI x nJ), V as (dim,)*4.
    keep: also return the intermediate W = Veff Y (or None) that
    monomer_2e_adjoint can take instead of recomputing it."""
    if bra.n_elec != ket.n_elec:
        raise ValueError("monomer_2e needs one charge")
    V = torch.as_tensor(V, dtype=torch.float64, device=ket.device).contiguous()
    if tuple(V.shape) != (ket.dim,) * 4:
        raise ValueError("V must be (dim, dim, dim, dim)")
    out = torch.empty((bra.n_states, ket.n_states), dtype=torch.float64, device=ket.device)
    ctx = context(ket.device.index)
    if not keep or ket.n_elec < 2:
        ctx.call("xr_monomer_2e", ket.n_orbs, ket.n_core, ket.n_elec, bra.n_states, bra.n_configs,
                 ptr(bra.coeffs), ket.n_states, ket.n_configs, ptr(ket.coeffs), ptr(V), ptr(out))
        return (out, None) if keep else out
    w_len = math.comb(ket.dim, 2) * math.comb(ket.dim, ket.n_elec - 2) * ket.n_states
    W = torch.empty(w_len, dtype=torch.float64, device=ket.device)
    ctx.call("xr_monomer_2e_keep", ket.n_orbs, ket.n_core, ket.n_elec, bra.n_states, bra.n_configs,
             ptr(bra.coeffs), ket.n_states, ket.n_configs, ptr(ket.coeffs), ptr(V), ptr(out), ptr(W), w_len)
    return out, W


def excitation_list(op, ket, q0=0, q1=None):
    """(Q, idx, P, sign) for ket determinants [q0, q1) in the reference loop
    order (general-XRCC/density_tensors.c:162-556), as device tensors."""
    if q1 is None:
        q1 = ket.n_configs
    ctx = context(ket.device.index)
    offsets = torch.empty(q1 - q0 + 1, dtype=torch.int64, device=ket.device)
    import ctypes
    total = ctypes.c_int64(0)
    ctx.call("xr_excitation_count", op.encode(), ket.n_orbs, ket.n_core, ket.n_elec,
             ptr(ket.masks), q0, q1, ptr(offsets), ctypes.byref(total))
    n = total.value
    e_q = torch.empty(n, dtype=torch.int64, device=ket.device)
    e_idx = torch.empty(n, dtype=torch.int64, device=ket.device)
    e_p = torch.empty(n, dtype=torch.int64, device=ket.device)
    e_s = torch.empty(n, dtype=torch.int8, device=ket.device)
    if n:
        ctx.call("xr_excitation_fill", op.encode(), ket.n_orbs, ket.n_core, ket.n_elec,
                 ptr(ket.masks), q0, q1, ptr(offsets), ptr(e_q), ptr(e_idx), ptr(e_p), ptr(e_s))
    return e_q, e_idx, e_p, e_s


def config_rank(block):
    out = torch.empty(block.n_configs, dtype=torch.int64, device=block.device)
    context(block.device.index).call("xr_config_rank", ptr(block.masks), block.n_configs,
                                     block.n_orbs, block.n_core, ptr(out))
    return out


def charge_blocks(z_lists, n_orbs, n_core, device=0):
    return {chg: ChargeBlock.from_states(st, n_orbs, n_core, device) for chg, st in z_lists.items()}


def conjugate_string(op):
    """The string of op^+: reversed, with c and a swapped."""
    return op[::-1].translate(str.maketrans("ca", "ac"))


def conjugate(rho, out=None, accumulate=False):
    """<J| op^+ |I>[reversed indices] from <I| op |J>[indices] for real
    states: (nI, nJ, d, ..., d) -> (nJ, nI, d, ..., d) with the index axes
    reversed (xr_conjugate; into out, added to it with accumulate).  An
    involution, so it also maps the adjoint back."""
    rho = rho.contiguous()
    k = rho.dim() - 2
    nI, nJ = int(rho.shape[0]), int(rho.shape[1])
    d = int(rho.shape[2]) if k else 1
    if out is None:
        out = torch.empty((nJ, nI) + tuple(rho.shape[2:]), dtype=rho.dtype, device=rho.device)
    if tuple(out.shape) != (nJ, nI) + tuple(rho.shape[2:]) or not out.is_contiguous():
        raise ValueError("conjugate: out must be a contiguous (nJ, nI, d, ...) tensor")
    context(rho.device.index).call("xr_conjugate", ptr(rho), nI, nJ, d, k, ptr(out), 1 if accumulate else 0)
    return out


def build_density_arrays(blocks, ops_table=OPS_BY_CHG_DIFF_FULL, skip=("ccaa",), derived=None,
                         conjugates=True, out=None):
    """All densities of one fragment on the device:
    rho[op][(bra_chg, ket_chg)] = tensor (nI, nJ, dim, ...).  Operator strings
    in `skip` are not built (the monomer 4-index term goes through
    monomer_2e instead).  A density whose conjugate string between the
    swapped charges is in the table is not built again but taken as the
    conjugate of that one (<I|op|J> = <J|op^+|I> for real states: half the
    GEMM work of the charge-changing strings); derived, if given, receives
    {(op, (bra, ket)): (op^+, (ket, bra))} for them.  conjugates=False builds
    every string directly.  out: a previous result whose tensors are
    overwritten where they have the same shape (fixed device addresses)."""
    rho = {}
    pairs = [(b, k) for b in blocks for k in blocks]
    pairs.sort(key=lambda bk: bk[1] - bk[0])        # bra - ket >= 0 first: the conjugate sources
    for bra_chg, ket_chg in pairs:
        for op in ops_table.get(bra_chg - ket_chg, ()):
            if op in skip:
                continue
            cop = conjugate_string(op)
            src = rho.get(cop, {}).get((ket_chg, bra_chg)) if cop != op and conjugates else None
            old = (out or {}).get(op, {}).get((bra_chg, ket_chg))
            if src is not None:
                if old is not None and old.shape != (src.shape[1], src.shape[0]) + tuple(src.shape[2:]):
                    old = None
                rho.setdefault(op, {})[(bra_chg, ket_chg)] = conjugate(src, out=old)
                if derived is not None:
                    derived[(op, (bra_chg, ket_chg))] = (cop, (ket_chg, bra_chg))
            else:
                bb, kb = blocks[bra_chg], blocks[ket_chg]
                if old is not None and tuple(old.shape) != (bb.n_states, kb.n_states) + (kb.dim,) * len(op):
                    old = None
                rho.setdefault(op, {})[(bra_chg, ket_chg)] = density(op, bb, kb, out=old)
    return rho
