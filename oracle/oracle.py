"""CPU oracle for the XR density build and Hamiltonian assembly.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP library is
compared against; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it.  The product package never imports it.

What it restates (reference file:line):
  * lexical configuration enumeration and rank:
      general-XRCC/atomic_states/mol_fci.py:99-127, fci_index.py:23-47,
      general-XRCC/density_tensors.c:29-62, build_density_tensors.py:31-37
  * the density kernels (C, oracle/xr_oracle.c):
      general-XRCC/density_tensors.c:162-556,
      Be-states/_attic_/deprecated/density_tensors.c:560-688
  * build_density_tensors host driver: general-XRCC/build_density_tensors.py:57-160
  * scalar contractions (C): general-XRCC/H_contractions.c:22-252
  * build_matrix_elements: general-XRCC/build_H.py:37-188
  * braket_loops: general-XRCC/hamiltonian.py:21-83
  * fci: general-XRCC/excitonic.py:56-82

How each part is pinned (tests/test_oracle_*.py): brute-force Fock-space
operator matrices built from bit strings (oracle/brute.py), density
invariants, and the dimer FCI-equivalence (XR eigenvalues == exact FCI
eigenvalues for complete fragment state spaces).
"""
import ctypes
import itertools
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I64P = ctypes.POINTER(ctypes.c_int64)
_DP = ctypes.POINTER(ctypes.c_double)
_I8P = ctypes.POINTER(ctypes.c_int8)
_i64 = ctypes.c_int64


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_config_rank.restype = _i64
        L.orc_excitations.restype = _i64
        for name in ("orc_monomer_1e", "orc_monomer_2e", "orc_dimer_2min2pls",
                     "orc_dimer_1min1pls_1e", "orc_dimer_1min1pls_2e", "orc_dimer_ExEx",
                     "orc_trimer_2min1pls1pls", "orc_trimer_2pls1min1min",
                     "orc_trimer_Ex1min1pls"):
            getattr(L, name).restype = ctypes.c_double
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------

def combo_table(n_e, n_so):
    """C(j, i) at row i, column j (upper triangle); FCIcomboMat layout
    (general-XRCC/build_density_tensors.py:31-37)."""
    out = np.zeros((max(n_e, 0), n_so), dtype=np.int64)
    for i in range(n_e):
        for j in range(i, n_so):
            out[i, j] = math.comb(j, i)
    return out


def lex_configs(n_orbs, n_elec, n_core=0):
    """All configurations of n_elec electrons in 2*n_orbs spin orbitals with
    the 2*n_core core spin orbitals occupied, as ascending occupation lists,
    ordered lexically over the valence spin orbitals (the enumeration the
    rank assumes; general-XRCC/atomic_states/mol_fci.py:99-127)."""
    core = list(range(n_core)) + list(range(n_orbs, n_orbs + n_core))
    valence = [p for p in range(2 * n_orbs) if p not in core]
    n_val = n_elec - 2 * n_core
    rows = []
    for combo in itertools.combinations(valence, n_val):
        rows.append(sorted(core + list(combo)))
    if not rows:
        return np.zeros((0, n_elec), dtype=np.int64)
    return np.array(rows, dtype=np.int64).reshape(len(rows), n_elec)


def config_rank(cfg, n_elec, n_orbs, n_core):
    comb = combo_table(n_elec - 2 * n_core, 2 * (n_orbs - n_core))
    c = np.ascontiguousarray(cfg, dtype=np.int64)
    return lib().orc_config_rank(_p(c, _I64P), _i64(n_elec), _i64(n_orbs), _i64(n_core),
                                 _p(comb, _I64P) if comb.size else None)


# ---------------------------------------------------------------------------
# densities
# ---------------------------------------------------------------------------

def n_elec_change(ops):
    return sum(+1 if o == "c" else -1 for o in ops)


def density(ops, n_orbs, n_core, zbra, zket, ket_configs):
    """Dense <I|ops|J>, shape (n_bra_states, n_ket_states, dim, ..., dim)."""
    ops_b = ops.encode()
    dim = 2 * n_orbs
    k = len(ops)
    zbra = np.ascontiguousarray(zbra, dtype=np.float64)
    zket = np.ascontiguousarray(zket, dtype=np.float64)
    cfg = np.ascontiguousarray(ket_configs, dtype=np.int64)
    n_elec_ket = cfg.shape[1]
    n_elec_bra = n_elec_ket + n_elec_change(ops)
    comb = combo_table(n_elec_bra - 2 * n_core, dim - 2 * n_core)
    if comb.size == 0:
        comb = np.zeros((1, 1), dtype=np.int64)
    nI, nP = zbra.shape
    nJ, nQ = zket.shape
    if cfg.shape[0] != nQ:
        raise ValueError("ket configs and coefficients disagree")
    out = np.zeros((nI, nJ) + (dim,) * k, dtype=np.float64)
    lib().orc_density(ops_b, _p(out, _DP), _i64(n_orbs), _i64(n_core),
                      _i64(nI), _i64(nP), _p(zbra, _DP), _p(comb, _I64P),
                      _i64(nJ), _i64(nQ), _p(zket, _DP), _i64(n_elec_ket), _p(cfg, _I64P))
    return out


def excitations(ops, n_orbs, n_core, ket_configs, q0=0, q1=None):
    """Excitation list (Q, idx, P, sign) in the reference loop order for ket
    configurations [q0, q1)."""
    cfg = np.ascontiguousarray(ket_configs, dtype=np.int64)
    if q1 is None:
        q1 = cfg.shape[0]
    n_elec_ket = cfg.shape[1]
    n_elec_bra = n_elec_ket + n_elec_change(ops)
    dim = 2 * n_orbs
    comb = combo_table(n_elec_bra - 2 * n_core, dim - 2 * n_core)
    if comb.size == 0:
        comb = np.zeros((1, 1), dtype=np.int64)
    L = lib()
    args = (ops.encode(), _i64(n_orbs), _i64(n_core), _p(comb, _I64P), _i64(n_elec_ket),
            _p(cfg, _I64P), _i64(q0), _i64(q1))
    n = L.orc_excitations(*args, _i64(0), None, None, None, None)
    e_q = np.zeros(n, np.int64)
    e_idx = np.zeros(n, np.int64)
    e_p = np.zeros(n, np.int64)
    e_s = np.zeros(n, np.int8)
    if n:
        L.orc_excitations(*args, _i64(n), _p(e_q, _I64P), _p(e_idx, _I64P), _p(e_p, _I64P),
                          _p(e_s, _I8P))
    return e_q, e_idx, e_p, e_s


# op string for each charge difference bra_chg - ket_chg
# (general-XRCC/build_density_tensors.py:90-157; 4-operator strings from
#  Be-states/_attic_/deprecated/density_tensors.c:560-688)
OPS_BY_CHG_DIFF = {+2: ("aa",), +1: ("a", "caa"), 0: ("ca", "ccaa"), -1: ("c", "cca"), -2: ("cc",)}
OPS_BY_CHG_DIFF_FULL = {+2: ("aa", "caaa"), +1: ("a", "caa"), 0: ("ca", "ccaa"),
                        -1: ("c", "cca"), -2: ("cc", "ccca")}


def monomer_2e(n_so, Rccaa, V):
    R = np.ascontiguousarray(Rccaa, dtype=np.float64)
    Vc = np.ascontiguousarray(V, dtype=np.float64)
    return lib().orc_monomer_2e(_i64(n_so), _p(R, _DP), _p(Vc, _DP))


def build_density_tensors(z_lists, n_orbs, Vints, n_core, n_threads=1):
    """Restates general-XRCC/build_density_tensors.py:57-160: returns
    (densities, total_size) with densities[op][(bra_chg, ket_chg)][i][j] an
    ndarray of shape (2*n_orbs,)*k, and the 'ccaa' entries replaced by the
    scalar sum V[p,q,r,s] R[p,q,s,r] (:128-133)."""
    n_so = 2 * n_orbs
    densities = {op: {} for op in ("aa", "a", "caa", "ca", "ccaa", "c", "cca", "cc")}
    total = 0
    for bra_chg in z_lists:
        for ket_chg in z_lists:
            zb = np.asarray(z_lists[bra_chg].coeffs, dtype=np.float64)
            zk = np.asarray(z_lists[ket_chg].coeffs, dtype=np.float64)
            nI, nJ = zb.shape[0], zk.shape[0]
            d = bra_chg - ket_chg
            for op in OPS_BY_CHG_DIFF.get(d, ()):
                R = density(op, n_orbs, n_core, zb, zk, z_lists[ket_chg].configs)
                if op == "ccaa":
                    densities[op][(bra_chg, ket_chg)] = [[monomer_2e(n_so, R[i, j], Vints)
                                                          for j in range(nJ)] for i in range(nI)]
                    total += nI * nJ
                else:
                    densities[op][(bra_chg, ket_chg)] = [[R[i, j] for j in range(nJ)]
                                                         for i in range(nI)]
                    total += R.size
    return densities, total


def build_density_arrays(z_lists, n_orbs, n_core, ops_table=OPS_BY_CHG_DIFF_FULL):
    """All dense densities as arrays rho[op][(bra_chg, ket_chg)] of shape
    (nI, nJ, dim, ..., dim)."""
    out = {}
    for bra_chg in z_lists:
        for ket_chg in z_lists:
            d = bra_chg - ket_chg
            for op in ops_table.get(d, ()):
                out.setdefault(op, {})[(bra_chg, ket_chg)] = density(
                    op, n_orbs, n_core, z_lists[bra_chg].coeffs, z_lists[ket_chg].coeffs,
                    z_lists[ket_chg].configs)
    return out


# ---------------------------------------------------------------------------
# matrix elements, Hamiltonian assembly, FCI
# ---------------------------------------------------------------------------

def _c(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class build_matrix_elements(object):
    """Restates general-XRCC/build_H.py:37-188 element by element, calling the
    C restatements of H_contractions.c."""

    def __init__(self, supersystem, integrals, nuc_repulsion):
        self.n_elec = [f.n_elec_ref for f in supersystem]
        self.rho = [f.rho for f in supersystem]
        self.T, self.U, self.V = integrals.T, integrals.U, integrals.V
        self.nuc = nuc_repulsion

    def monomer(self, m, I, J):
        (ic, ii), (jc, ji) = I, J
        if ic != jc:
            return 0.0
        L = lib()
        n = self.T[m, m].shape[0]
        Rca = _c(self.rho[m]["ca"][(ic, jc)][ii][ji])
        Rccaa = self.rho[m]["ccaa"][(ic, jc)][ii][ji]
        h = _c(self.T[m, m] + self.U[m, m, m])
        Enuc = self.nuc[m, m] if ii == ji else 0.0
        return Enuc + Rccaa + L.orc_monomer_1e(_i64(n), _p(Rca, _DP), _p(h, _DP))

    def dimer(self, frags, I, J):
        L = lib()
        m1, m2 = frags
        (i1c, i1), (i2c, i2) = I
        (j1c, j1), (j2c, j2) = J
        if not (i1c + i2c == j1c + j2c and j1c - 2 <= i1c <= j1c + 2):
            return 0.0
        rho, T, U, V = self.rho, self.T, self.U, self.V
        n1, n2 = T[m1, m1].shape[0], T[m2, m2].shape[0]
        sign2 = (-1) ** ((self.n_elec[m2] - i2c) % 2)
        r1 = lambda op: _c(rho[m1][op][(i1c, j1c)][i1][j1])
        r2 = lambda op: _c(rho[m2][op][(i2c, j2c)][i2][j2])
        P = lambda a: _p(a, _DP)
        if i1c == j1c - 2:
            return L.orc_dimer_2min2pls(_i64(n1), _i64(n2), P(r1("cc")), P(r2("aa")), P(_c(V[m1, m1, m2, m2])))
        if i1c == j1c - 1:
            h = _c(T[m1, m2] + U[m1, m1, m2] + U[m2, m1, m2])
            t = L.orc_dimer_1min1pls_1e(_i64(n1), _i64(n2), P(r1("c")), P(r2("a")), P(h))
            t += L.orc_dimer_1min1pls_2e(_i64(n1), _i64(n2), P(r1("c")), P(r1("cca")), P(r2("a")),
                                         P(r2("caa")), P(_c(V[m1, m1, m1, m2])), P(_c(V[m1, m2, m2, m2])))
            return +sign2 * t
        if i1c == j1c:
            Rca1, Rca2 = r1("ca"), r2("ca")
            t = L.orc_dimer_ExEx(_i64(n1), _i64(n2), P(Rca1), P(Rca2), P(_c(V[m1, m2, m1, m2])))
            if i1 == j1:
                t += L.orc_monomer_1e(_i64(n2), P(Rca2), P(_c(U[m1, m2, m2])))
            if i2 == j2:
                t += L.orc_monomer_1e(_i64(n1), P(Rca1), P(_c(U[m2, m1, m1])))
            if i1 == j1 and i2 == j2:
                t += self.nuc[m1, m2]
            return t
        if i1c == j1c + 1:
            h = _c(T[m2, m1] + U[m1, m2, m1] + U[m2, m2, m1])
            t = L.orc_dimer_1min1pls_1e(_i64(n2), _i64(n1), P(r2("c")), P(r1("a")), P(h))
            t += L.orc_dimer_1min1pls_2e(_i64(n2), _i64(n1), P(r2("c")), P(r2("cca")), P(r1("a")),
                                         P(r1("caa")), P(_c(V[m2, m2, m2, m1])), P(_c(V[m2, m1, m1, m1])))
            return -sign2 * t
        # i1c == j1c + 2
        return L.orc_dimer_2min2pls(_i64(n2), _i64(n1), P(r2("cc")), P(r1("aa")), P(_c(V[m2, m2, m1, m1])))

    def trimer(self, frags, I, J):
        L = lib()
        m = list(frags)
        ic = [x[0] for x in I]
        ii = [x[1] for x in I]
        jc = [x[0] for x in J]
        ji = [x[1] for x in J]
        diffs = sorted(ic[k] - jc[k] for k in range(3))
        if diffs not in ([-2, 1, 1], [-1, 0, 1], [-1, -1, 2]):
            return 0.0
        rho, T, U, V = self.rho, self.T, self.U, self.V
        n = [T[x, x].shape[0] for x in m]
        sign = [1, (-1) ** ((self.n_elec[m[1]] - ic[1]) % 2), (-1) ** ((self.n_elec[m[2]] - ic[2]) % 2)]
        R = lambda k, op: _c(rho[m[k]][op][(ic[k], jc[k])][ii[k]][ji[k]])
        P = lambda a: _p(a, _DP)
        N = lambda *ks: [_i64(n[k]) for k in ks]
        # general-XRCC/build_H.py:115-186, one branch per charge pattern
        for a, b, c, s in ((0, 1, 2, sign[2]), (1, 0, 2, sign[1] * sign[2]), (2, 0, 1, sign[1])):
            if ic[a] == jc[a] - 2:
                return s * L.orc_trimer_2min1pls1pls(*N(a, b, c), P(R(a, "cc")), P(R(b, "a")), P(R(c, "a")),
                                                     P(_c(V[m[a], m[a], m[b], m[c]])))
        for a, b, c, s in ((0, 1, 2, sign[2]), (1, 0, 2, sign[1] * sign[2]), (2, 0, 1, sign[1])):
            if ic[a] == jc[a] + 2:
                return s * L.orc_trimer_2pls1min1min(*N(a, b, c), P(R(a, "aa")), P(R(b, "c")), P(R(c, "c")),
                                                     P(_c(V[m[b], m[c], m[a], m[a]])))
        cases = ((0, 1, 2, +sign[2]), (0, 2, 1, -sign[2]),
                 (1, 0, 2, +sign[1] * sign[2]), (1, 2, 0, -sign[1] * sign[2]),
                 (2, 0, 1, +sign[1]), (2, 1, 0, -sign[1]))
        for a, b, c, s in cases:
            if ic[a] == jc[a] and ic[b] == jc[b] - 1:
                t = L.orc_trimer_Ex1min1pls(*N(a, b, c), P(R(a, "ca")), P(R(b, "c")), P(R(c, "a")),
                                            P(_c(V[m[a], m[b], m[a], m[c]])))
                if ii[a] == ji[a]:
                    t += L.orc_dimer_1min1pls_1e(*N(b, c), P(R(b, "c")), P(R(c, "a")),
                                                 P(_c(U[m[a], m[b], m[c]])))
                return s * t
        return 0.0


def _ket_loop(Hmat, bra, I, ket_now, J, N_frag, spf, H, par=None):
    """Restates general-XRCC/hamiltonian.py:21-63.  With par (par[m][i] =
    electron-count parity of state i of fragment m) the one-electron
    transfer couplings between non-adjacent fragments get the sign of the
    spectator fragments in between, which the reference omits."""
    n = len(ket_now) + 1
    ket_next = list(ket_now) + [None]
    if n == N_frag:
        H1, H2, H3 = H
        ket = ket_next
        for i in range(spf):
            ket[-1] = i
            tr = [(m, b, k) for m, (b, k) in enumerate(zip(bra, ket)) if b != k]
            if len(tr) == 0:
                for M in range(N_frag):
                    Hmat[I, J] += H1[M][bra[M], ket[M]]
                    for N in range(M + 1, N_frag):
                        Hmat[I, J] += H2[M][N][bra[M] * spf + bra[N], ket[M] * spf + ket[N]]
            if len(tr) == 1:
                (M, Mb, Mk), = tr
                Hmat[I, J] += H1[M][Mb, Mk]
                for N in range(0, M):
                    Hmat[I, J] += H2[N][M][bra[N] * spf + Mb, ket[N] * spf + Mk]
                for N in range(M + 1, N_frag):
                    Hmat[I, J] += H2[M][N][Mb * spf + bra[N], Mk * spf + ket[N]]
            if len(tr) == 2:
                (M, Mb, Mk), (N, Nb, Nk) = tr
                sgn = 1.0
                if par is not None and par[M][Mb] != par[M][Mk]:
                    for O in range(M + 1, N):
                        if par[O][bra[O]]:
                            sgn = -sgn
                Hmat[I, J] += sgn * H2[M][N][Mb * spf + Nb, Mk * spf + Nk]
                for O in range(0, M):
                    Hmat[I, J] += H3[O][M][N][(bra[O] * spf + Mb) * spf + Nb, (ket[O] * spf + Mk) * spf + Nk]
                for O in range(M + 1, N):
                    Hmat[I, J] += H3[M][O][N][(Mb * spf + bra[O]) * spf + Nb, (Mk * spf + ket[O]) * spf + Nk]
                for O in range(N + 1, N_frag):
                    Hmat[I, J] += H3[M][N][O][(Mb * spf + Nb) * spf + bra[O], (Mk * spf + Nk) * spf + ket[O]]
            if len(tr) == 3:
                (M, Mb, Mk), (N, Nb, Nk), (O, Ob, Ok) = tr
                Hmat[I, J] += H3[M][N][O][(Mb * spf + Nb) * spf + Ob, (Mk * spf + Nk) * spf + Ok]
            J += 1
    else:
        stride = spf ** (N_frag - n)
        for i in range(spf):
            ket_next[-1] = i
            _ket_loop(Hmat, bra, I, ket_next, J, N_frag, spf, H, par)
            J += stride


def _bra_loop(Hmat, bra_now, I, N_frag, spf, H, par=None):
    """Restates general-XRCC/hamiltonian.py:65-79."""
    n = len(bra_now) + 1
    bra_next = list(bra_now) + [None]
    if n == N_frag:
        for i in range(spf):
            bra_next[-1] = i
            _ket_loop(Hmat, bra_next, I, [], 0, N_frag, spf, H, par)
            I += 1
    else:
        stride = spf ** (N_frag - n)
        for i in range(spf):
            bra_next[-1] = i
            _bra_loop(Hmat, bra_next, I, N_frag, spf, H, par)
            I += stride


def braket_loops(Hmat, N_frag, states_per_frag, H, par=None):
    _bra_loop(Hmat, [], 0, N_frag, states_per_frag, H, par)


def state_parities(supersystem):
    """par[m][i]: parity of the electron count of state i (state_indices order)."""
    return [[(f.n_elec_ref - chg) % 2 for chg, _ in f.state_indices] for f in supersystem]


def fci(H, par=None):
    """Lowest eigenvalue of the assembled XR matrix (general-XRCC/excitonic.py:56-82).
    The reference calls numpy.linalg.eig and sorts; the matrix is symmetric
    for Hermitian integrals, so eigvalsh gives the same spectrum."""
    H1 = H[0]
    N_frag = len(H1)
    spf = H1[0].shape[0]
    Hmat = np.zeros((spf ** N_frag, spf ** N_frag))
    braket_loops(Hmat, N_frag, spf, H, par)
    vals = np.linalg.eigvals(Hmat)
    vals = np.sort(vals.real)
    return vals[0], Hmat


def compute_H_blocks(supersystem, integrals, nuc, with_trimers=False):
    """H1[M] (spf x spf) and H2[M][N] (spf^2 x spf^2) from per-element calls,
    in the layout of general-XRCC/test_H.py:63-108."""
    bme = build_matrix_elements(supersystem, integrals, nuc)
    nf = len(supersystem)
    si = [f.state_indices for f in supersystem]
    H1 = []
    for M in range(nf):
        d = len(si[M])
        h = np.zeros((d, d))
        for i, I in enumerate(si[M]):
            for j, J in enumerate(si[M]):
                h[i, j] = bme.monomer(M, I, J)
        H1.append(h)
    H2 = [[None] * nf for _ in range(nf)]
    H3 = [[[None] * nf for _ in range(nf)] for _ in range(nf)]
    for M in range(nf):
        for N in range(M + 1, nf):
            basis = [(a, b) for a in si[M] for b in si[N]]
            h = np.zeros((len(basis), len(basis)))
            for i, I in enumerate(basis):
                for j, J in enumerate(basis):
                    h[i, j] = bme.dimer((M, N), I, J)
            H2[M][N] = h
            if with_trimers:
                for O in range(N + 1, nf):
                    b3 = [(a, b, c) for a in si[M] for b in si[N] for c in si[O]]
                    h3 = np.zeros((len(b3), len(b3)))
                    for i, I in enumerate(b3):
                        for j, J in enumerate(b3):
                            h3[i, j] = bme.trimer((M, N, O), I, J)
                    H3[M][N][O] = h3
    return H1, H2, H3


# ---------------------------------------------------------------------------
# XR-CCSD by brute force (general-XRCC/excitonic.py:28-53)
# ---------------------------------------------------------------------------

class _NoTrimers:
    """H3 of a Hamiltonian without three-fragment terms: every element 0."""

    def __getitem__(self, k):
        return 0.0 if isinstance(k, tuple) else self


def xr_ccsd(H1, H2, thresh=1e-11):
    """Restates the XR-CCSD of general-XRCC/excitonic.py:28-53 (Qode's CCSD
    on H_012_operator) by brute force: the dense XR matrix from braket_loops,
    T (singles t_m[i] |i><0|_m, doubles t_mn[i, j] |i><0|_m |j><0|_n, i, j != 0)
    applied with numpy tensor contractions, exp(+-T) by the finite series,
    the residuals <mu| e^-T H e^T |0> for every single and double mu, solved
    by scipy.optimize.root from T = 0 (not the Jacobi / DIIS of the GPU
    path).  H1: monomer matrices; H2[m][n]: pair matrices (m < n).  Returns
    (E, singles, doubles)."""
    from scipy.optimize import root
    nf = len(H1)
    dims = [h.shape[0] for h in H1]
    n = int(np.prod(dims))
    spf = dims[0]
    if any(d != spf for d in dims):
        raise ValueError("braket_loops needs equal state counts")
    H = np.zeros((n, n))
    braket_loops(H, nf, spf, (H1, H2, _NoTrimers()))
    pairs = [(m, k) for m in range(nf) for k in range(m + 1, nf)]

    def unpack(x):
        t1, o = [], 0
        for d in dims:
            t = np.zeros(d)
            t[1:] = x[o:o + d - 1]
            o += d - 1
            t1.append(t)
        t2 = {}
        for m, k in pairs:
            t = np.zeros((dims[m], dims[k]))
            c = (dims[m] - 1) * (dims[k] - 1)
            t[1:, 1:] = x[o:o + c].reshape(dims[m] - 1, dims[k] - 1)
            o += c
            t2[(m, k)] = t
        return t1, t2

    def apply_T(t1, t2, x, sign):
        X = x.reshape(dims)
        Y = np.zeros_like(X)
        for m in range(nf):
            A = np.zeros((dims[m], dims[m]))
            A[:, 0] = t1[m]
            Y += np.moveaxis(np.tensordot(A, X, axes=(1, m)), 0, m)
        for (m, k), t in t2.items():
            S = np.take(np.take(X, 0, axis=k), 0, axis=m)      # fragments m, k in state 0
            full = np.multiply.outer(t, S)                      # axes (m, k, others)
            Y += np.moveaxis(full, [0, 1], [m, k])
        return sign * Y.reshape(-1)

    def exp_T(t1, t2, x, sign):
        out, term = x.copy(), x
        for j in range(1, nf + 1):
            term = apply_T(t1, t2, term, sign) / j
            out = out + term
        return out

    ref = np.zeros(n)
    ref[0] = 1.0
    strides = [int(np.prod(dims[m + 1:])) for m in range(nf)]

    def residual(x):
        t1, t2 = unpack(x)
        W = exp_T(t1, t2, H @ exp_T(t1, t2, ref, 1.0), -1.0)
        r = [W[np.arange(1, d) * strides[m]] for m, d in enumerate(dims)]
        for m, k in pairs:
            i = np.arange(1, dims[m]) * strides[m]
            j = np.arange(1, dims[k]) * strides[k]
            r.append(W[i[:, None] + j[None, :]].reshape(-1))
        return np.concatenate(r), W[0]

    n_amp = sum(d - 1 for d in dims) + sum((dims[m] - 1) * (dims[k] - 1) for m, k in pairs)
    sol = root(lambda x: residual(x)[0], np.zeros(n_amp), method="hybr", tol=thresh)
    r, E = residual(sol.x)
    if float(np.abs(r).max()) > 1e3 * thresh:
        raise RuntimeError(f"oracle XR-CCSD did not converge (max residual {np.abs(r).max():.2e})")
    t1, t2 = unpack(sol.x)
    return float(E), t1, t2
