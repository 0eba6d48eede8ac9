"""Oracle samples at full workload size: densities of single (I, J) state
pairs and pair-matrix elements, written by tests/golden/make_sample_fixtures.py
and read by the GPU tests.

At the full sizes the oracle cannot build whole density tensors (one
3-index string of an 18-orbital fragment with 5 electrons takes ~20 s per
state pair on one core), so the fixture generator picks one (I, J) pair per
(fragment, bra charge, ket charge), lets the oracle compute every density of
exactly those pairs, and evaluates one element per charge combination of
every pair block with the oracle's build_matrix_elements.dimer on them.

What is stored per density: 1- and 2-index strings whole; 3-index strings
as 1024 entries at seeded positions plus the checksum sum(w * rho) with
seeded N(0, 1) weights w over every entry (an error anywhere in the tensor
moves the checksum).
"""
import numpy as np

import oracle.oracle as orc

N_SAMPLED_ENTRIES = 1024


class _Row:
    def __init__(self, pairs, i):
        self.pairs, self.i = pairs, i

    def __getitem__(self, j):
        return self.pairs.get(self.i, j)


class _Pairs:
    def __init__(self, op, frag, bra_chg, ket_chg, cache):
        self.op, self.frag, self.bc, self.kc, self.cache = op, frag, bra_chg, ket_chg, cache

    def get(self, i, j):
        key = (self.op, self.bc, self.kc, i, j)
        if key not in self.cache:
            sb, sk = self.frag.states[self.bc], self.frag.states[self.kc]
            zb = np.ascontiguousarray(sb.coeffs[i:i + 1])
            zk = np.ascontiguousarray(sk.coeffs[j:j + 1])
            self.cache[key] = orc.density(self.op, self.frag.n_orbs, self.frag.n_core, zb, zk, sk.configs)[0, 0]
        return self.cache[key]

    def __getitem__(self, i):
        return _Row(self, i)


class _Op:
    def __init__(self, op, frag, cache):
        self.op, self.frag, self.cache = op, frag, cache

    def __getitem__(self, charges):
        bc, kc = charges
        return _Pairs(self.op, self.frag, bc, kc, self.cache)


class OracleFragment:
    """A fragment as oracle.build_matrix_elements reads it (.n_elec_ref and
    .rho[op][(bra_chg, ket_chg)][I][J]), the oracle computing each density
    of a state pair on first use."""

    def __init__(self, frag):
        self.n_elec_ref = frag.n_elec_ref
        self.cache = {}
        self.rho = {op: _Op(op, frag, self.cache)
                    for op in ("a", "c", "ca", "aa", "cc", "caa", "cca", "ccaa", "caaa", "ccca")}


def sample_pairs(frag, rng):
    """One (I, J) state pair per (bra charge, ket charge) of a fragment."""
    out = {}
    for bc, sb in frag.states.items():
        for kc, sk in frag.states.items():
            out[(bc, kc)] = (int(rng.integers(sb.coeffs.shape[0])), int(rng.integers(sk.coeffs.shape[0])))
    return out


def _c(x):
    return f"m{-x}" if x < 0 else f"p{x}"


def dens_key(m, op, bc, kc):
    return f"d{m}_{op}_{_c(bc)}_{_c(kc)}"


def key_seed(key):
    """Seed of a density's sampled positions and checksum weights, fixed by
    its key (Python's own string hash is salted per process)."""
    import zlib
    return zlib.crc32(key.encode())


def entry_positions(size, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(size, size=min(size, N_SAMPLED_ENTRIES), replace=False))


def checksum_weights(size, seed):
    return np.random.default_rng(seed + 1).standard_normal(size)


def summarize(rho, seed):
    """(positions, values, checksum) of one state pair's density: the whole
    tensor for up to 2 indices, sampled entries + checksum beyond."""
    flat = np.ascontiguousarray(rho).reshape(-1)
    if rho.ndim <= 2:
        return np.arange(flat.size), flat.copy(), float(flat.sum())
    pos = entry_positions(flat.size, seed)
    return pos, flat[pos].copy(), float(checksum_weights(flat.size, seed) @ flat)


def check_density(got, pos, vals, chk, seed):
    """Compares one GPU density of a state pair with its stored summary;
    returns the largest deviation (entries absolute; the checksum relative
    to its weighted 1-norm, floored at 1)."""
    flat = np.ascontiguousarray(got).reshape(-1)
    err = float(np.abs(flat[pos] - vals).max()) if pos.size else 0.0
    if got.ndim > 2:
        w = checksum_weights(flat.size, seed)
        scale = max(float(np.abs(w) @ np.abs(flat)), 1.0)
        err = max(err, abs(float(w @ flat) - chk) / scale)
    return err
