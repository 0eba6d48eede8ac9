"""Seeded synthetic fragments, states and integrals for the XR workloads.

There are no molecular integrals or CI vectors to load, so every workload
is generated from a seed:

* fragment states: per charge block, N(0,1) columns over all lexically
  ordered determinants, orthonormalised by QR and stored as rows
  (n_states x n_configs), the layout of `states.coeffs` that
  general-XRCC/build_density_tensors.py:77 consumes;
* one-electron h: symmetric, over the supersystem spatial orbitals;
* two-electron V: 8-fold symmetric chemist-notation (pq|rs), damped by
  exp(-|p-q|/3) exp(-|r-s|/3) (local orbital indices) and, for chains, by
  exp(-span/2) over the fragments the four indices touch;
* spin-orbital blocks in the convention the contractions assume
  (general-XRCC/H_contractions.c prefactors 1, 2, 4): H = sum h_pq p+q +
  sum V_pqrs p+q+sr with V_pqrs = (1/4)(<pq|rs> - <pq|sr>),
  <pq|rs> = (pr|qs) d(sp,sr) d(sq,ss); each fragment lists its alpha spin
  orbitals first, then its beta spin orbitals.
"""
import itertools
import math
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------------------
# determinants and states
# ---------------------------------------------------------------------------

def lex_configs(n_orbs, n_elec, n_core=0):
    """Ascending occupation lists of every determinant with the core spin
    orbitals filled, in lexical order over the valence spin orbitals (the
    order in which general-XRCC/density_tensors.c:29-62 ranks them)."""
    core = list(range(n_core)) + list(range(n_orbs, n_orbs + n_core))
    valence = [p for p in range(2 * n_orbs) if p not in core]
    n_val = n_elec - 2 * n_core
    if n_val < 0 or n_val > len(valence):
        return np.zeros((0, n_elec), dtype=np.int64)
    n_cfg = math.comb(len(valence), n_val)
    out = np.empty((n_cfg, n_elec), dtype=np.int64)
    core_arr = np.array(core, dtype=np.int64)
    for k, combo in enumerate(itertools.combinations(valence, n_val)):
        out[k] = np.sort(np.concatenate([core_arr, np.array(combo, dtype=np.int64)]))
    return out


@dataclass
class ChargeStates:
    """Fragment states of one charge: .configs (n_configs x n_elec) and
    .coeffs (n_states x n_configs), as general-XRCC/Be631g.py:36-52 loads them."""
    configs: np.ndarray
    coeffs: np.ndarray


def random_states(n_orbs, n_elec, n_states, rng, n_core=0):
    cfg = lex_configs(n_orbs, n_elec, n_core)
    n_cfg = cfg.shape[0]
    n_states = min(n_states, n_cfg)
    A = rng.standard_normal((n_cfg, n_states))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))[None, :]
    return ChargeStates(configs=cfg, coeffs=np.ascontiguousarray(Q.T))


# ---------------------------------------------------------------------------
# integrals
# ---------------------------------------------------------------------------

def spatial_h(n, rng, frag_of=None, inter_decay=None):
    A = rng.standard_normal((n, n))
    h = 0.5 * (A + A.T)
    if frag_of is not None and inter_decay is not None:
        f = np.asarray(frag_of)
        h = h * np.exp(-np.abs(f[:, None] - f[None, :]) / inter_decay)
    return h


def spatial_V(n, rng, local_index=None, frag_of=None, inter_decay=None):
    """8-fold symmetric (pq|rs), ~N(0,1) damped by exp(-|p-q|/3)exp(-|r-s|/3)."""
    B = rng.standard_normal((n, n, n, n))
    perms = [(0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
             (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)]
    V = sum(B.transpose(p) for p in perms) / np.sqrt(8.0)
    li = np.arange(n) if local_index is None else np.asarray(local_index)
    d = np.exp(-np.abs(li[:, None] - li[None, :]) / 3.0)
    V = V * d[:, :, None, None] * d[None, None, :, :]
    if frag_of is not None and inter_decay is not None:
        f = np.asarray(frag_of)
        F = np.stack(np.meshgrid(f, f, f, f, indexing="ij"))
        span = F.max(axis=0) - F.min(axis=0)
        V = V * np.exp(-span / inter_decay)
    return V


class FragmentBlocked:
    """Fragment-blocked spin-orbital integrals, indexable like the reference's
    integrals.T[m,n], integrals.U[a,m,n], integrals.V[m,n,o,p]
    (general-XRCC/build_H.py:41, general-XRCC/get_ints.py:162-172).  Blocks are
    built on first use from the spatial arrays and cached."""

    def __init__(self, frag_orbs, h_sp, V_sp, U_sp=None):
        self.frag_orbs = list(frag_orbs)
        self.offsets = np.concatenate([[0], np.cumsum(self.frag_orbs)]).astype(int)
        self.h_sp = h_sp
        self.V_sp = V_sp
        self.U_sp = U_sp            # (n_frag, n_sp, n_sp) or None
        self.T = _BlockView(self, "T")
        self.U = _BlockView(self, "U")
        self.V = _BlockView(self, "V")
        self._cache = {}

    def sl(self, m):
        return slice(self.offsets[m], self.offsets[m + 1])

    def n_so(self, m):
        return 2 * self.frag_orbs[m]

    def one_body(self, A, m1, m2):
        n1, n2 = self.frag_orbs[m1], self.frag_orbs[m2]
        out = np.zeros((2 * n1, 2 * n2))
        blk = A[self.sl(m1), self.sl(m2)]
        out[:n1, :n2] = blk
        out[n1:, n2:] = blk
        return out

    def two_body(self, m1, m2, m3, m4):
        n = [self.frag_orbs[m] for m in (m1, m2, m3, m4)]
        s = [self.sl(m) for m in (m1, m2, m3, m4)]
        C = self.V_sp[s[0], s[2], s[1], s[3]].transpose(0, 2, 1, 3)   # (pr|qs)
        X = self.V_sp[s[0], s[3], s[1], s[2]].transpose(0, 2, 3, 1)   # (ps|qr)
        out = np.zeros((2 * n[0], 2 * n[1], 2 * n[2], 2 * n[3]))
        for s1, s2, s3, s4 in itertools.product((0, 1), repeat=4):
            blk = np.zeros((n[0], n[1], n[2], n[3]))
            if s1 == s3 and s2 == s4:
                blk = blk + C
            if s1 == s4 and s2 == s3:
                blk = blk - X
            if s1 == s3 and s2 == s4 or s1 == s4 and s2 == s3:
                out[s1 * n[0]:(s1 + 1) * n[0], s2 * n[1]:(s2 + 1) * n[1],
                    s3 * n[2]:(s3 + 1) * n[2], s4 * n[3]:(s4 + 1) * n[3]] = 0.25 * blk
        return out

    def block(self, kind, key):
        ck = (kind,) + tuple(key)
        if ck not in self._cache:
            if kind == "T":
                self._cache[ck] = self.one_body(self.h_sp, *key)
            elif kind == "U":
                a, m1, m2 = key
                if self.U_sp is None:
                    self._cache[ck] = np.zeros((self.n_so(m1), self.n_so(m2)))
                else:
                    self._cache[ck] = self.one_body(self.U_sp[a], m1, m2)
            else:
                self._cache[ck] = self.two_body(*key)
        return self._cache[ck]

    def full_spin_orbital(self):
        """Whole-supersystem h (T + sum_a U[a]) and V in the fragment-blocked
        spin-orbital order (for small brute-force checks)."""
        nf = len(self.frag_orbs)
        so = [2 * n for n in self.frag_orbs]
        off = np.concatenate([[0], np.cumsum(so)]).astype(int)
        D = off[-1]
        h = np.zeros((D, D))
        V = np.zeros((D, D, D, D))
        for m1 in range(nf):
            for m2 in range(nf):
                blk = self.block("T", (m1, m2)).copy()
                for a in range(nf):
                    blk += self.block("U", (a, m1, m2))
                h[off[m1]:off[m1 + 1], off[m2]:off[m2 + 1]] = blk
        for ms in itertools.product(range(nf), repeat=4):
            V[tuple(slice(off[m], off[m + 1]) for m in ms)] = self.block("V", ms)
        return h, V


class _BlockView:
    def __init__(self, owner, kind):
        self.owner, self.kind = owner, kind

    def __getitem__(self, key):
        return self.owner.block(self.kind, key)


# ---------------------------------------------------------------------------
# fragments and workloads
# ---------------------------------------------------------------------------

@dataclass
class Fragment:
    """One fragment: spatial orbitals, frozen-core count, reference electron
    count, and states per relative charge (charge = n_elec_ref - n_elec)."""
    n_orbs: int
    n_core: int
    n_elec_ref: int
    states: dict
    rho: dict = None
    state_indices: list = field(default_factory=list)

    def __post_init__(self):
        if not self.state_indices:
            ref = 0 if 0 in self.states else next(iter(self.states))
            idx = [(ref, 0)] + [(ref, i) for i in range(1, self.states[ref].coeffs.shape[0])]
            for chg in self.states:
                if chg != ref:
                    idx += [(chg, i) for i in range(self.states[chg].coeffs.shape[0])]
            self.state_indices = idx


@dataclass
class Workload:
    name: str
    fragments: list
    integrals: FragmentBlocked
    nuc: np.ndarray
    seed: int


def make_fragment(n_orbs, elec_counts, n_states, n_elec_ref, rng, n_core=0):
    """elec_counts: electron numbers to include; n_states: int or dict by count."""
    states = {}
    for ne in elec_counts:
        ns = n_states[ne] if isinstance(n_states, dict) else n_states
        states[n_elec_ref - ne] = random_states(n_orbs, ne, ns, rng, n_core)
    # order charge blocks by decreasing electron count of the reference first
    ordered = {}
    if 0 in states:
        ordered[0] = states[0]
    for chg in sorted(states):
        if chg != 0:
            ordered[chg] = states[chg]
    return Fragment(n_orbs=n_orbs, n_core=n_core, n_elec_ref=n_elec_ref, states=ordered)


def make_workload(name, frag_orbs, elec_counts, n_states, n_elec_ref, seed,
                  inter_decay=None, U=False):
    rng = np.random.default_rng(seed)
    frags = [make_fragment(n, elec_counts[i] if isinstance(elec_counts, list) and
                           isinstance(elec_counts[0], (list, tuple)) else elec_counts,
                           n_states[i] if isinstance(n_states, list) else n_states,
                           n_elec_ref, rng) for i, n in enumerate(frag_orbs)]
    n_sp = int(sum(frag_orbs))
    frag_of = np.concatenate([[m] * n for m, n in enumerate(frag_orbs)])
    local = np.concatenate([np.arange(n) for n in frag_orbs])
    h = spatial_h(n_sp, rng, frag_of, inter_decay)
    V = spatial_V(n_sp, rng, local, frag_of, inter_decay)
    U_sp = None
    if U:
        U_sp = np.stack([spatial_h(n_sp, rng) * 0.1 for _ in frag_orbs])
    ints = FragmentBlocked(frag_orbs, h, V, U_sp)
    nuc = np.zeros((len(frag_orbs), len(frag_orbs)))
    return Workload(name=name, fragments=frags, integrals=ints, nuc=nuc, seed=seed)


def config0(seed=0):
    """Plain FP64 dimer: 2 x 9 spatial orbitals, 3/4/5 electrons, 8 states per block."""
    return make_workload("dimer_9x9_e345_s8", [9, 9], [3, 4, 5], 8, 4, seed)


def config1_fragment(seed=1, n_states=32, elec_counts=(2, 3, 4, 5, 6)):
    """Single fragment density build: 13 spatial orbitals, 2-6 electrons, 32 states."""
    rng = np.random.default_rng(seed)
    return make_fragment(13, list(elec_counts), n_states, 4, rng)


def config2(seed=2):
    """6 uneven fragments (9,9,13,13,18,18 spatial orbitals), 3/4/5 electrons,
    8-64 states per charge block drawn from the seed."""
    rng = np.random.default_rng(seed)
    orbs = [9, 9, 13, 13, 18, 18]
    ns = [int(rng.choice([8, 16, 32, 64])) for _ in orbs]
    return make_workload("six_fragments_uneven", orbs, [3, 4, 5], ns, 4, seed)


def config3(seed=3, n_states=64):
    """Split-precision dimer: 2 x 18 spatial orbitals, 3/4/5 electrons, 64 states."""
    return make_workload("dimer_18x18_e345_s64", [18, 18], [3, 4, 5], n_states, 4, seed)


def config4(seed=4, n_states=32):
    """4-fragment chain, 13 spatial orbitals each, 3/4/5 electrons, 32 states,
    inter-fragment decay exp(-d/2)."""
    return make_workload("chain4_13_e345_s32", [13, 13, 13, 13], [3, 4, 5], n_states, 4, seed,
                         inter_decay=2.0)
