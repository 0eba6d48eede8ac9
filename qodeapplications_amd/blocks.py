"""Charge-block contractions of fragment densities with integral blocks.

One call evaluates a reference matrix-element function
(general-XRCC/build_H.py:42-188) for every state pair of a charge
combination at once, through the batched H_contractions families of
libxrhip.  Densities are device tensors rho[op][(bra_chg, ket_chg)] of shape
(nI, nJ, dim, ...).
"""
import numpy as np
import torch

from ._lib import context, ptr


def _flat(R):
    return R.reshape(R.shape[0] * R.shape[1], -1).contiguous()


def _dev(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)


class IntegralCache:
    """Device copies of the fragment-blocked integrals the contractions use:
    T[m,n], U[a,m,n] and V[m,n,o,p] (general-XRCC/build_H.py:41)."""

    def __init__(self, integrals, device):
        self.ints = integrals
        self.device = device
        self._c = {}

    def T(self, *k):
        return self._get(("T",) + k, lambda: self.ints.T[k])

    def U(self, *k):
        return self._get(("U",) + k, lambda: self.ints.U[k])

    def V(self, *k):
        return self._get(("V",) + k, lambda: self.ints.V[k])

    def h_sum(self, key, parts):
        """Sum of one-electron blocks, e.g. T[m1,m2] + U[m1,m1,m2] + U[m2,m1,m2]."""
        def make():
            acc = None
            for kind, k in parts:
                a = getattr(self, kind)(*k)
                acc = a.clone() if acc is None else acc + a
            return acc
        return self._get(("h",) + key, make)

    def set_block(self, key, t):
        """Replaces one cached block (e.g. with rotated integrals) and drops
        the sums built from the old ones."""
        self._c[key] = t.detach().to(device=self.device, dtype=torch.float64).contiguous()
        for k in [k for k in self._c if k[0] == "h"]:
            del self._c[k]

    def _get(self, key, make):
        t = self._c.get(key)
        if t is None:
            t = self._c[key] = _dev(make(), self.device)
        return t


def monomer_1e(Rca, h):
    """sum_pq h[p,q] Rca[a,p,q] for every state pair a -> (nI, nJ)."""
    nI, nJ, n = Rca.shape[0], Rca.shape[1], Rca.shape[2]
    out = torch.empty(nI * nJ, dtype=torch.float64, device=Rca.device)
    ctx = context(Rca.device.index)
    ctx.call("xr_monomer_1e", n, nI * nJ, ptr(Rca.contiguous()), ptr(h), ptr(out))
    return out.view(nI, nJ)


def dimer_block(m1, m2, n_elec_ref2, rho1, rho2, keys1, keys2, ints):
    """All elements of build_matrix_elements.dimer for one charge combination
    (general-XRCC/build_H.py:56-102) except the Kronecker-delta terms.

    Returns (terms, extras).  terms: [(Hb, swapped, scale)], the element is
    sum scale * Hb[a, b] with a = I1*nJ1+J1, b = I2*nJ2+J2 (Hb[b, a] when
    swapped).  extras lists the terms that only apply on spectator-diagonal
    elements: ("e2", v): v[b] where I1 == J1 (extPot of fragment 1 on 2),
    ("e1", v): v[a] where I2 == J2.  The nuclear term (both diagonal) is
    left to the caller.
    """
    (i1c, j1c), (i2c, j2c) = keys1, keys2
    d = i1c - j1c
    sign2 = (-1) ** ((n_elec_ref2 - i2c) % 2)
    dev = ints.device
    ctx = context(dev.index)
    L = lambda name, *a: ctx.call(name, *a)
    if d == -2:
        R1, R2 = rho1["cc"][keys1], rho2["aa"][keys2]
        n1, n2 = R1.shape[2], R2.shape[2]
        nA, nB = R1.shape[0] * R1.shape[1], R2.shape[0] * R2.shape[1]
        out = torch.empty(nA, nB, dtype=torch.float64, device=dev)
        L("xr_dimer_2min2pls", n1, n2, nA, ptr(_flat(R1)), nB, ptr(_flat(R2)), ptr(ints.V(m1, m1, m2, m2)), ptr(out))
        return [(out, False, 1.0)], []
    if d == +2:
        R2, R1 = rho2["cc"][keys2], rho1["aa"][keys1]
        n1, n2 = R1.shape[2], R2.shape[2]
        nA, nB = R1.shape[0] * R1.shape[1], R2.shape[0] * R2.shape[1]
        out = torch.empty(nB, nA, dtype=torch.float64, device=dev)
        L("xr_dimer_2min2pls", n2, n1, nB, ptr(_flat(R2)), nA, ptr(_flat(R1)), ptr(ints.V(m2, m2, m1, m1)), ptr(out))
        return [(out, True, 1.0)], []
    if d == -1:
        Rc1, Rcca1 = rho1["c"][keys1], rho1["cca"][keys1]
        Ra2, Rcaa2 = rho2["a"][keys2], rho2["caa"][keys2]
        n1, n2 = Rc1.shape[2], Ra2.shape[2]
        nA, nB = Rc1.shape[0] * Rc1.shape[1], Ra2.shape[0] * Ra2.shape[1]
        h = ints.h_sum((m1, m2, "1min"), [("T", (m1, m2)), ("U", (m1, m1, m2)), ("U", (m2, m1, m2))])
        out = torch.empty(nA, nB, dtype=torch.float64, device=dev)
        L("xr_dimer_1min1pls", n1, n2, nA, ptr(_flat(Rc1)), ptr(_flat(Rcca1)), nB, ptr(_flat(Ra2)),
          ptr(_flat(Rcaa2)), ptr(h), ptr(ints.V(m1, m1, m1, m2)), ptr(ints.V(m1, m2, m2, m2)), ptr(out))
        return [(out, False, float(sign2))], []
    if d == +1:
        Rc2, Rcca2 = rho2["c"][keys2], rho2["cca"][keys2]
        Ra1, Rcaa1 = rho1["a"][keys1], rho1["caa"][keys1]
        n1, n2 = Ra1.shape[2], Rc2.shape[2]
        nA, nB = Ra1.shape[0] * Ra1.shape[1], Rc2.shape[0] * Rc2.shape[1]
        h = ints.h_sum((m2, m1, "1pls"), [("T", (m2, m1)), ("U", (m1, m2, m1)), ("U", (m2, m2, m1))])
        out = torch.empty(nB, nA, dtype=torch.float64, device=dev)
        L("xr_dimer_1min1pls", n2, n1, nB, ptr(_flat(Rc2)), ptr(_flat(Rcca2)), nA, ptr(_flat(Ra1)),
          ptr(_flat(Rcaa1)), ptr(h), ptr(ints.V(m2, m2, m2, m1)), ptr(ints.V(m2, m1, m1, m1)), ptr(out))
        return [(out, True, float(-sign2))], []
    # d == 0: exciton-exciton coupling plus the spectator terms
    Rca1, Rca2 = rho1["ca"][keys1], rho2["ca"][keys2]
    n1, n2 = Rca1.shape[2], Rca2.shape[2]
    nA, nB = Rca1.shape[0] * Rca1.shape[1], Rca2.shape[0] * Rca2.shape[1]
    out = torch.empty(nA, nB, dtype=torch.float64, device=dev)
    L("xr_dimer_ExEx", n1, n2, nA, ptr(_flat(Rca1)), nB, ptr(_flat(Rca2)), ptr(ints.V(m1, m2, m1, m2)), ptr(out))
    e2 = monomer_1e(Rca2, ints.U(m1, m2, m2)).reshape(-1)
    e1 = monomer_1e(Rca1, ints.U(m2, m1, m1)).reshape(-1)
    return [(out, False, 1.0)], [("e2", e2), ("e1", e1)]


def trimer_block(ms, n_elec_ref, rho, keys, ints):
    """build_matrix_elements.trimer (general-XRCC/build_H.py:103-188) for one
    charge combination of fragments ms = (m1, m2, m3), or None if the
    charges do not couple.  Returns (Hb, order, scale, extra): Hb[x, y, z]
    runs over the state pairs of fragments order = (x, y, z); extra is None
    or (spectator, u, (y, z), scale) for the U term that only applies where
    the spectator fragment is diagonal."""
    ic = [k[0] for k in keys]
    jc = [k[1] for k in keys]
    diffs = sorted(ic[k] - jc[k] for k in range(3))
    if diffs not in ([-2, 1, 1], [-1, 0, 1], [-1, -1, 2]):
        return None
    sign = [1, (-1) ** ((n_elec_ref[1] - ic[1]) % 2), (-1) ** ((n_elec_ref[2] - ic[2]) % 2)]
    dev = ints.device
    ctx = context(dev.index)
    R = lambda k, op: rho[k][op][keys[k]]
    npair = lambda T: T.shape[0] * T.shape[1]

    def run(name, a, b, c, opa, opb, opc, V, f):
        Ra, Rb, Rc = R(a, opa), R(b, opb), R(c, opc)
        out = torch.empty(npair(Ra), npair(Rb), npair(Rc), dtype=torch.float64, device=dev)
        ctx.call(name, Ra.shape[2], Rb.shape[2], Rc.shape[2], npair(Ra), ptr(_flat(Ra)), npair(Rb),
                 ptr(_flat(Rb)), npair(Rc), ptr(_flat(Rc)), ptr(V), ptr(out))
        return out

    m = ms
    for a, b, c, s in ((0, 1, 2, sign[2]), (1, 0, 2, sign[1] * sign[2]), (2, 0, 1, sign[1])):
        if ic[a] == jc[a] - 2:
            return run("xr_trimer_2min1pls1pls", a, b, c, "cc", "a", "a", ints.V(m[a], m[a], m[b], m[c]), s), (a, b, c), float(s), None
    for a, b, c, s in ((0, 1, 2, sign[2]), (1, 0, 2, sign[1] * sign[2]), (2, 0, 1, sign[1])):
        if ic[a] == jc[a] + 2:
            return run("xr_trimer_2pls1min1min", a, b, c, "aa", "c", "c", ints.V(m[b], m[c], m[a], m[a]), s), (a, b, c), float(s), None
    cases = ((0, 1, 2, +sign[2]), (0, 2, 1, -sign[2]), (1, 0, 2, +sign[1] * sign[2]),
             (1, 2, 0, -sign[1] * sign[2]), (2, 0, 1, +sign[1]), (2, 1, 0, -sign[1]))
    for a, b, c, s in cases:
        if ic[a] == jc[a] and ic[b] == jc[b] - 1:
            Hb = run("xr_trimer_Ex1min1pls", a, b, c, "ca", "c", "a", ints.V(m[a], m[b], m[a], m[c]), s)
            Rb, Rcc = R(b, "c"), R(c, "a")
            u = torch.empty(npair(Rb), npair(Rcc), dtype=torch.float64, device=dev)
            ctx.call("xr_dimer_1min1pls_1e", Rb.shape[2], Rcc.shape[2], npair(Rb), ptr(_flat(Rb)),
                     npair(Rcc), ptr(_flat(Rcc)), ptr(ints.U(m[a], m[b], m[c])), ptr(u))
            return Hb, (a, b, c), float(s), (a, u, (b, c), float(s))
    return None


def scatter(out, ldo, frags, rows, cols, d, block, strides, diag=None, scale=1.0):
    """Adds block (or the constant `scale` if block is None) into out through
    xr_block_scatter; frags lists (nI, nJ) per fragment."""
    import ctypes
    nf = len(frags)
    diag = diag or [0] * nf
    nI = (ctypes.c_int64 * nf)(*[f[0] for f in frags])
    nJ = (ctypes.c_int64 * nf)(*[f[1] for f in frags])
    st = (ctypes.c_int64 * nf)(*strides)
    dg = (ctypes.c_int * nf)(*diag)
    rp = (ctypes.c_void_p * nf)(*[r.data_ptr() for r in rows])
    cp = (ctypes.c_void_p * nf)(*[c.data_ptr() for c in cols])
    dd = (ctypes.c_int64 * nf)(*d)
    ctx = context(out.device.index)
    ctx.call("xr_block_scatter", nf, nI, nJ, st, dg, rp, cp, dd,
             None if block is None else ptr(block.contiguous()), float(scale), ptr(out), ldo)
