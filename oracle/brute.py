"""Brute-force Fock-space operators (TEST INFRASTRUCTURE ONLY).

Independent of the list-walking algorithm in xr_oracle.c: every operator is
an explicit matrix over all 2^dim occupation bit strings, with the sign
convention of general-XRCC/density_tensors.c:80-113 written as a bit
formula, a_p|n> = (-1)^popcount(n >> (p+1)) |n - p>.  Used to pin the
oracle's densities and, through the dimer FCI comparison, the whole chain
build_density_tensors -> build_matrix_elements -> braket_loops -> fci.
Only for dim <= 12.
"""
import itertools

import numpy as np


def popcount(x):
    return bin(x).count("1")


def annihilator(p, dim):
    n = 1 << dim
    M = np.zeros((n, n))
    for s in range(n):
        if (s >> p) & 1:
            M[s & ~(1 << p), s] = -1.0 if popcount(s >> (p + 1)) & 1 else 1.0
    return M


def operators(dim):
    a = [annihilator(p, dim) for p in range(dim)]
    c = [x.T.copy() for x in a]
    return c, a


def config_bits(cfg):
    b = 0
    for p in cfg:
        b |= 1 << int(p)
    return b


def embed(z, configs, dim):
    """State coefficients (n_states x n_configs) -> Fock-space row vectors."""
    out = np.zeros((z.shape[0], 1 << dim))
    for k, cfg in enumerate(configs):
        out[:, config_bits(cfg)] = z[:, k]
    return out


def density(ops, zbra_f, zket_f, dim, cache=None):
    """<I| ops |J> for every index tuple, as (nI, nJ, dim, ..., dim)."""
    c, a = cache if cache is not None else operators(dim)
    k = len(ops)
    out = np.zeros((zbra_f.shape[0], zket_f.shape[0]) + (dim,) * k)
    for idx in itertools.product(range(dim), repeat=k):
        M = np.eye(1 << dim)
        for o, p in zip(ops, idx):
            M = M @ (c[p] if o == "c" else a[p])
        out[(slice(None), slice(None)) + idx] = zbra_f @ M @ zket_f.T
    return out


def density_fast(ops, zbra_f, zket_f, dim):
    """The same <I| ops |J> as density(), with the operators applied right
    to left to the ket vectors as signed index maps (every index tuple of a
    level at once) instead of explicit matrix products: usable for 5-operator
    strings on dim <= 8."""
    n = 1 << dim
    states = np.arange(n)
    maps = []
    for p in range(dim):
        has = (states >> p) & 1 == 1
        sg = np.array([-1.0 if popcount(int(s) >> (p + 1)) & 1 else 1.0 for s in states])
        maps.append((has, sg))
    V = zket_f.T[None]                          # (tuples, 2^dim, nJ), leading index = later operators
    for o in reversed(ops):
        out = np.zeros((dim,) + V.shape)
        for p in range(dim):
            has, sg = maps[p]
            bit = 1 << p
            if o == "a":                         # a_p: |s> -> sg |s - p> for s with p occupied
                src = states[has]
                out[p][:, src ^ bit] = sg[src][None, :, None] * V[:, src]
            else:                                # c_p = a_p^T: |t> -> sg(t + p) |t + p> for t with p empty
                src = states[~has]
                out[p][:, src | bit] = sg[src | bit][None, :, None] * V[:, src]
        V = out.reshape((-1,) + V.shape[1:])
    k = len(ops)
    res = np.einsum("is,tsj->ijt", zbra_f, V)
    return res.reshape((zbra_f.shape[0], zket_f.shape[0]) + (dim,) * k)


def hamiltonian(h, V, cache=None):
    """Full Fock-space H = sum h_pq p+q + sum V_pqrs p+q+sr."""
    dim = h.shape[0]
    c, a = cache if cache is not None else operators(dim)
    H = np.zeros((1 << dim, 1 << dim))
    for p in range(dim):
        for q in range(dim):
            if h[p, q] != 0.0:
                H += h[p, q] * (c[p] @ a[q])
    cc = [[c[p] @ c[q] for q in range(dim)] for p in range(dim)]
    aa = [[a[s] @ a[r] for r in range(dim)] for s in range(dim)]
    for p, q, r, s in itertools.product(range(dim), repeat=4):
        v = V[p, q, r, s]
        if v != 0.0:
            H += v * (cc[p][q] @ aa[s][r])
    return H


def number_operator(dim):
    n = 1 << dim
    return np.array([popcount(s) for s in range(n)], dtype=np.int64)


def density_elements(ops, tuples, zbra, zket, bra_configs, ket_configs):
    """<I| o_1(p_1) ... o_k(p_k) |J> for the index tuples (p_1..p_k), the
    operators applied right to left to every ket determinant as a bit mask
    (a_p|n> = (-1)^popcount(n >> (p+1)) |n - p>, c_p likewise on the mask
    before the bit is set), the results looked up among the bra
    determinants.  zbra, zket: one state's coefficients over bra / ket
    configs (tuples of occupied orbitals).  O(n_ket) per tuple, no tensor
    built: usable at full determinant counts."""
    def masks(configs):
        out = np.zeros(len(configs), dtype=np.int64)
        for k, cfg in enumerate(configs):
            for p in cfg:
                out[k] |= np.int64(1) << np.int64(int(p))
        return out
    km = masks(ket_configs)
    bm = masks(bra_configs)
    order = np.argsort(bm)
    bs = bm[order]
    zk = np.asarray(zket, dtype=np.float64)
    zb = np.asarray(zbra, dtype=np.float64)
    vals = np.empty(len(tuples))
    for t, idx in enumerate(tuples):
        m = km.copy()
        sign = np.ones(len(km))
        ok = np.ones(len(km), dtype=bool)
        for o, p in zip(reversed(ops), reversed(idx)):
            bit = np.int64(1) << np.int64(int(p))
            occ = (m & bit) != 0
            ok &= occ if o == "a" else ~occ
            above = np.bitwise_count((m >> np.int64(int(p) + 1)).astype(np.uint64))
            sign *= np.where(above & 1, -1.0, 1.0)
            m = m ^ bit
        pos = np.searchsorted(bs, m)
        pos_c = np.minimum(pos, len(bs) - 1)
        found = ok & (pos < len(bs)) & (bs[pos_c] == m)
        vals[t] = float(np.sum(zk[found] * sign[found] * zb[order[pos_c[found]]]))
    return vals
